"""GPU parity of the Gaussian filters: the reference's valid-mode
gaussian_filter3d / gaussian_convolve2 against the oracle restatement of its
CPU twins (filter.h:158-200), and the native fused 6-component same-mode
filter against an FP64 oracle."""
import numpy as np
import pytest

from tiralib_amd import api
from tiralib_amd._lib import DeviceArray

pytestmark = pytest.mark.gpu


def _rel_max(a, b):
    return np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("shape,sig", [((20, 22, 24), (1.0, 0.7, 1.5)), ((40, 33, 70), (2.0, 2.0, 2.0)),
                                       ((9, 64, 65), (0.5, 1.2, 0.8)), ((30, 30, 30), (2.7, 1.0, 0.3))])
def test_gaussian_filter3d_valid(hip, oracle, shape, sig):
    vol = np.random.default_rng(0).normal(size=shape).astype(np.float32)
    gpu = api.gaussian_filter3d(vol, *sig)
    ref = oracle.gaussian_convolve3(vol, *sig)
    assert gpu.shape == ref.shape
    e = _rel_max(gpu, ref)
    print(f"gaussian_filter3d {shape} {sig}: {e:.2e}")
    assert e <= 1e-5


@pytest.mark.parametrize("shape,sig", [((50, 61), (1.0, 1.5)), ((128, 100), (2.0, 0.5)), ((7, 300), (0.34, 1.0))])
def test_gaussian_convolve2_valid(hip, oracle, shape, sig):
    img = np.random.default_rng(1).normal(size=shape).astype(np.float32)
    gpu = api.gaussian_convolve2(img, *sig)
    kw, kh = int(sig[0] * 6), int(sig[1] * 6)
    gw, gh = oracle.gaussian1d(kw, 0, sig[0], 1), oracle.gaussian1d(kh, 0, sig[1], 1)
    ref = oracle.convolve3(img[None], gw, (kw, 1, 1))
    ref = oracle.convolve3(ref, gh, (1, kh, 1))[0]
    assert gpu.shape == ref.shape
    assert _rel_max(gpu, ref) <= 1e-5


@pytest.mark.parametrize("shape,sigma,K", [((37, 41, 45), 1.5, 11), ((64, 48, 32), 2.0, 13), ((5, 6, 7), 1.0, 7),
                                           ((33, 17, 129), 1.5, 9)])
def test_tensor6_same_mode(hip, oracle, shape, sigma, K):
    rng = np.random.default_rng(2)
    t6 = rng.normal(size=shape + (6,)).astype(np.float32)
    taps = api.gaussian_taps(K, sigma)
    np.testing.assert_array_equal(taps, oracle.gaussian1d(K, 0, sigma, 1))
    out = api.gaussian_filter3_tensor(DeviceArray.from_host(t6), taps, same=True).to_host()
    ref = oracle.gaussian3_same_packed6_f64(t6, taps, nthreads=8)
    e = _rel_max(out, ref)
    print(f"tensor6 same {shape} K={K}: {e:.2e}")
    assert e <= 2e-6
    # valid mode == interior of same mode
    if min(shape) < K:
        with pytest.raises(ValueError):
            api.gaussian_filter3_tensor(DeviceArray.from_host(t6), taps, same=False)
        return
    v = api.gaussian_filter3_tensor(DeviceArray.from_host(t6), taps, same=False).to_host()
    r = K // 2
    np.testing.assert_allclose(v, ref[r:-r, r:-r, r:-r], atol=2e-6 * np.abs(ref).max())


def test_tensor6_same_mode_deep_chunks(hip, oracle):
    """200 x 256 x 256 runs in 128-plane z chunks (a 128- and a 72-plane chunk per
    column): output blocks across the chunk seam (z = 128) and at two corners match the FP64
    oracle of an input crop with a K/2 margin (clipped at the volume, where the crop's zero
    padding is the volume's)."""
    shape = (200, 256, 256)
    t6 = np.random.default_rng(5).standard_normal(size=shape + (6,), dtype=np.float32)
    taps = api.gaussian_taps(13, 2.0)
    dev = DeviceArray.from_host(t6)
    out = api.gaussian_filter3_tensor(dev, taps, same=True).to_host()
    dev.free()
    r, b = 6, 16
    for corner in [(120, 100, 60), (0, 0, 0), (184, 240, 240)]:
        lo = [max(c - r, 0) for c in corner]
        hi = [min(c + b + r, s) for c, s in zip(corner, shape)]
        crop = np.ascontiguousarray(t6[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]])
        ref = oracle.gaussian3_same_packed6_f64(crop, taps, nthreads=8)
        ref = ref[tuple(slice(c - l, c - l + b) for c, l in zip(corner, lo))]
        got = out[corner[0]:corner[0] + b, corner[1]:corner[1] + b, corner[2]:corner[2] + b]
        e = _rel_max(got, ref)
        print(f"tensor6 same 200x256x256 block at {corner}: {e:.2e}")
        assert e <= 2e-6


def test_tensor_constant_field(hip):
    t = np.full((30, 31, 32, 6), 2.5, np.float32)
    taps = api.gaussian_taps(11, 1.5)
    out = api.gaussian_filter3_tensor(DeviceArray.from_host(t), taps, same=True).to_host()
    np.testing.assert_allclose(out[5:-5, 5:-5, 5:-5], 2.5, rtol=3e-6)
    assert np.all(out[0, 0, 0] < 2.5)  # zero padding at the corner
