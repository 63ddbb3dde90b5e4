"""GPU parity of the persistent, graph-captured batch (Gaussian smoothing +
eig3 of many uneven 6-component volumes): smoothed fields against the FP64
same-mode oracle, eigenvalues against FP64 Jacobi, and graph replays --
interleaved with eager work and synchronisations -- against the first run."""
import numpy as np
import pytest

from tiralib_amd import _lib, api, synth
from tiralib_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu

SHAPES = [(32, 48, 40), (17, 20, 33), (64, 16, 16), (5, 70, 9), (48, 48, 48)]


def _volumes(seed=0):
    host = []
    for i, s in enumerate(SHAPES):
        t6, _ = synth.eig_tensors(int(np.prod(s)), seed=2000 + seed + i)
        host.append(t6.reshape(s + (6,)))
    return host


@pytest.mark.parametrize("K,sigma", [(11, 1.5), (13, 2.0), (7, 1.0)])
def test_batch_gauss_eig3(hip, oracle, K, sigma):
    host = _volumes()
    dev = [DeviceArray.from_host(h) for h in host]
    taps = api.gaussian_taps(K, sigma)
    b = api.BatchGaussEig3(dev, taps, autotune=(K == 11))
    b.run()
    _lib.call("tira_hip_synchronize")
    plan = b.plan()
    print("plan", plan)
    assert plan["items"] > 0 and plan["zchunk"] in (8, 16, 32, 64) and 1 <= plan["blocks_per_cu"] <= 4
    for h, o, e in zip(host, b.out, b.evals):
        ref = oracle.gaussian3_same_packed6_f64(h, taps, nthreads=8)
        got = o.to_host()
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
        ev = e.to_host().reshape(-1, 3).astype(np.float64)
        ref_ev = oracle.eig3_packed6_f64(h.reshape(-1, 6))
        nw = np.abs(ev - ref_ev).max(1) / np.abs(ref_ev).max(1)
        assert nw.max() <= 1e-6, nw.max()
    b.close()


def test_batch_graph_replays(hip):
    host = _volumes(7)
    dev = [DeviceArray.from_host(h) for h in host]
    taps = api.gaussian_taps(11, 1.5)
    b = api.BatchGaussEig3(dev, taps, autotune=False)
    b.run()
    _lib.call("tira_hip_synchronize")
    first = [(o.to_host(), e.to_host()) for o, e in zip(b.out, b.evals)]
    scratch = DeviceArray.from_host(host[0])
    for rep in range(3):
        for o in b.out:
            o.zero()
        # eager work on the same stream between replays
        api.gaussian_filter3_tensor(scratch, taps, same=True)
        b.run()
        _lib.call("tira_hip_synchronize")
        for (fo, fe), o, e in zip(first, b.out, b.evals):
            np.testing.assert_array_equal(o.to_host(), fo)
            np.testing.assert_array_equal(e.to_host(), fe)
    b.close()
