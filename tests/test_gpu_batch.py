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
    assert plan["items"] > 0 and plan["zchunk"] in (8, 16, 32, 64) and 1 <= plan["blocks_per_cu"] <= 8
    assert plan["tile"][0] == 16 and plan["tile"][1] in (8, 16, 32)
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


@pytest.mark.parametrize("layout", [_lib.LAYOUT_PACKED6, _lib.LAYOUT_MAT9])
def test_eig3_batch_one_launch(hip, oracle, layout):
    """tira_eig3_batch_async over uneven arrays (including an empty one and one
    of 1 tensor) equals per-array tira_eig3_async bit for bit, and FP64 Jacobi
    to 1e-6 normwise."""
    counts = [int(np.prod(s)) for s in SHAPES] + [0, 1, 513]
    host = []
    for i, c in enumerate(counts):
        t6, _ = synth.eig_tensors(max(c, 1), seed=3000 + i)
        t6 = t6[:c]
        host.append(synth.mat9_from_packed6(t6) if layout == _lib.LAYOUT_MAT9 else t6)
    dev = [DeviceArray.from_host(h if h.size else np.zeros((1, 6), np.float32)) for h in host]
    got = api.eig3_batch(dev, counts=counts, layout=layout)
    singles = [api.eig3(d, layout=layout)[0].to_host() if c else None for d, c in zip(dev, counts)]
    _lib.call("tira_hip_synchronize")
    for h, c, g, s in zip(host, counts, got, singles):
        if not c:
            continue
        ev = g.to_host()
        assert np.array_equal(ev, s)
        t6 = h if layout == _lib.LAYOUT_PACKED6 else synth.packed6_from_mat9(h)
        ref = oracle.eig3_packed6_f64(np.ascontiguousarray(t6, np.float32))
        nw = np.abs(ev.astype(np.float64) - ref).max(1) / np.abs(ref).max(1)
        assert nw.max() <= 1e-6, nw.max()


def test_eig3_batch_many_segments(hip):
    """More arrays than one launch's segment table (64) holds."""
    counts = [37 + 11 * i for i in range(150)]
    host = [synth.eig_tensors(c, seed=4000 + i)[0] for i, c in enumerate(counts)]
    dev = [DeviceArray.from_host(h) for h in host]
    got = api.eig3_batch(dev)
    singles = [api.eig3(d)[0].to_host() for d in dev]
    _lib.call("tira_hip_synchronize")
    for g, s in zip(got, singles):
        assert np.array_equal(g.to_host(), s)


def test_batch_tile_shapes_agree(hip):
    """Every tile shape the autotuner chooses from (16 x 8, 16 x 16, 16 x 32)
    and z-chunk gives the same smoothed fields and eigenvalues, bit for bit
    (each output's sums run in the same order whatever the tile)."""
    host = _volumes(11)
    dev = [DeviceArray.from_host(h) for h in host]
    b = api.BatchGaussEig3(dev, api.gaussian_taps(13, 2.0), autotune=False)
    b.run()
    _lib.call("tira_hip_synchronize")
    first = [(o.to_host(), e.to_host()) for o, e in zip(b.out, b.evals)]
    for ty, zc, occ in ((8, 8, 4), (8, 64, 1), (32, 16, 1), (32, 64, 2), (16, 5, 3)):
        for o in b.out:
            o.zero()
        b.set_plan(ty, zc, occ)
        assert b.plan()["tile"] == [16, ty] and b.plan()["zchunk"] == zc
        b.run()
        _lib.call("tira_hip_synchronize")
        for (fo, fe), o, e in zip(first, b.out, b.evals):
            np.testing.assert_array_equal(o.to_host(), fo, err_msg=f"tile 16x{ty}, z-chunk {zc}")
            np.testing.assert_array_equal(e.to_host(), fe)
    b.close()
