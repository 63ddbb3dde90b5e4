"""GPU parity of the front end and the whole 3D pipeline: the fused
gradient + structure-tensor kernel bit for bit against the volume.h
restatement, and structure tensor -> Gaussian -> vote -> eig3 saliency
against the same chain evaluated by the FP64 oracle."""
import numpy as np
import pytest

from tiralib_amd import _lib, api, synth
from tiralib_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spacing", [(1.0, 1.0, 1.0), (0.7, 1.3, 2.1)])
def test_structure_tensor_bit_exact(hip, oracle, spacing):
    rng = np.random.default_rng(11)
    vol = rng.normal(size=(9, 13, 70)).astype(np.float32)
    got = api.structure_tensor3(DeviceArray.from_host(vol), spacing).to_host()
    _lib.call("tira_hip_synchronize")
    ref = oracle.structure_tensor3(vol, spacing)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("shape", [(5, 7, 66), (65537, 2, 3), (131073, 2, 3)])
def test_structure_tensor_unaligned_output(hip, oracle, shape):
    """An output that starts 4 bytes into a buffer, with the tiled kernel (two
    planes per thread; 65537 planes end in a one-plane column) and, for 131073
    planes (beyond the grid's z limit), the grid-stride one: the same bits as
    the restatement."""
    rng = np.random.default_rng(12)
    vol = rng.normal(size=shape).astype(np.float32)
    n = int(np.prod(shape))
    d_in = DeviceArray.from_host(vol)
    d_out = DeviceArray((n * 6 + 1,))
    _lib.call("tira_structure_tensor3_async", d_in.ptr, *shape, 1.0, 1.5, 0.75, d_out.ptr + 4)
    _lib.call("tira_hip_synchronize")
    got = d_out.to_host()[1:].reshape(shape + (6,))
    assert np.array_equal(got, oracle.structure_tensor3(vol, (1.0, 1.5, 0.75)))


def test_structure_tensor_rejects_thin_axes(hip):
    with pytest.raises(ValueError):
        api.structure_tensor3(DeviceArray((1, 8, 8)))


def test_vessel_volume(hip):
    shape = (48, 56, 64)
    v1, segs = synth.vessel_volume_device(shape, seed=3, noise_sigma=0.1)
    v2, _ = synth.vessel_volume_device(shape, seed=3, noise_sigma=0.1)
    a, b = v1.to_host(), v2.to_host()
    assert np.array_equal(a, b) and len(segs) > 0
    inside = a > 0.5
    assert 0 < inside.mean() < 0.5
    assert abs(a[~inside & (a < 0.45)].std() - 0.1) < 0.02


@pytest.mark.parametrize("mode", [_lib.PLATE_CLOSED, _lib.PLATE_SUM])
def test_pipeline_matches_fp64_chain(hip, oracle, mode):
    shape, smooth, sigma, w = (28, 30, 36), 1.0, 1.5, 7
    vol, _ = synth.vessel_volume_device(shape, seed=5, noise_sigma=0.1)
    sal = api.pipeline3(vol, smooth_sigma=smooth, vote_sigma=sigma, w=w, samples=20, plate_mode=mode)
    _lib.call("tira_hip_synchronize")
    got = sal.to_host().reshape(-1, 3)
    # the same chain in the oracle: FP32 structure tensor, FP64 Gaussian and vote
    t6 = oracle.structure_tensor3(vol.to_host())
    taps = api.gaussian_taps(7, smooth)
    sm = oracle.gaussian3_same_packed6_f64(t6, taps, nthreads=8).astype(np.float32)
    rng = np.random.default_rng(2)
    recv = np.sort(rng.choice(int(np.prod(shape)), 160, replace=False))
    v9 = oracle.tensorvote3_f64(synth.mat9_from_packed6(sm).reshape(shape + (9,)), sigma, 0.0, 1, w, samples=20,
                                recv=recv, nthreads=8)
    M = v9.reshape(-1, 3, 3)
    ev = np.linalg.eigvalsh(0.5 * (M + M.transpose(0, 2, 1)))
    ref = np.stack([ev[:, 2] - ev[:, 1], ev[:, 1] - ev[:, 0], ev[:, 0]], axis=1)
    err = np.abs(got[recv] - ref).max() / np.abs(ref).max()
    print("saliency max error / max saliency", err)
    assert err <= 1e-4


def test_pipeline_with_learned_plate(hip):
    shape, sigma = (24, 28, 32), 2.0
    vol, _ = synth.vessel_volume_device(shape, seed=6, noise_sigma=0.1)
    net = api.PlateMLP(sigma, 0.0, 1, 64, seed=8)
    net.train(nsamples=1 << 20, epochs=30, batch=8192, lr=1e-2, lr_final=1e-5, seed=9)
    a = api.pipeline3(vol, smooth_sigma=1.0, vote_sigma=sigma, plate_mode=_lib.PLATE_MLP, mlp=net, mlp_tol=1e30)
    b = api.pipeline3(vol, smooth_sigma=1.0, vote_sigma=sigma, plate_mode=_lib.PLATE_CLOSED)
    _lib.call("tira_hip_synchronize")
    x, y = a.to_host(), b.to_host()
    err = np.linalg.norm(x - y) / np.linalg.norm(y)
    print("MLP pipeline vs closed-form pipeline", err)
    assert err <= 5e-3
    net.close()
