"""GPU parity of the saliency threshold (reconstruct.py:14-63) and Lee 3D
thinning (thinning/lee_thinning.h:600-706) against the oracle.

Bar: bit-exact. The three neighbourhood tests agree with the oracle's
restatement on all 2^26 neighbourhoods; thinned volumes, Otsu thresholds and
masks are identical to the oracle's on the same inputs. Saliency is floating
point: within 1e-6 of numpy.linalg.eigh (FP64) relative to the field's
largest saliency."""
import numpy as np
import pytest
from scipy import ndimage

from tiralib_amd import _lib, api, reconstruct, synth
from tiralib_amd._lib import DeviceArray, call

pytestmark = pytest.mark.gpu


def test_classify_all_neighbourhoods(hip, oracle):
    n = 1 << 26
    d = DeviceArray((n,), np.uint8)
    try:
        call("tira_thin3_classify_async", 0, n, d.ptr)
        call("tira_hip_synchronize")
        gpu = d.to_host()
    finally:
        d.free()
    ref = oracle.thin_classify(0, n, nthreads=16)
    bad = np.nonzero(gpu != ref)[0]
    assert bad.size == 0, (bad[:10], gpu[bad[:10]], ref[bad[:10]])


def _blob(shape, seed, q=0.5, sigma=1.5):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.random(shape), sigma)
    return (f > np.quantile(f, q)).astype(np.uint8)


def _volumes():
    vols = [("blob", _blob((14, 13, 12), 1)), ("blob2", _blob((20, 33, 41), 2, 0.6)),
            ("slab", _blob((1, 40, 50), 3)), ("thin-y", _blob((30, 2, 28), 4, 0.3)),
            ("dense", _blob((24, 24, 24), 5, 0.2, 2.5)), ("sparse", _blob((31, 29, 37), 6, 0.85, 1.0))]
    bar = np.zeros((10, 10, 30), np.uint8)
    bar[3:6, 3:6, 3:27] = 1
    ball = np.zeros((17, 17, 17), np.uint8)
    z, y, x = np.mgrid[:17, :17, :17]
    ball[(x - 8) ** 2 + (y - 8) ** 2 + (z - 8) ** 2 <= 36] = 1
    vols += [("bar", bar), ("ball", ball), ("full", np.ones((9, 10, 11), np.uint8)),
             ("empty", np.zeros((5, 6, 7), np.uint8))]
    return vols


@pytest.mark.parametrize("name,vol", _volumes(), ids=[v[0] for v in _volumes()])
def test_lee_bit_exact(hip, oracle, name, vol):
    ref, it_ref = oracle.lee_thin3(vol)
    got, it = api.lee_thin3(vol.astype(np.int32) * 3)  # any nonzero value is object
    assert it == it_ref
    np.testing.assert_array_equal(got, ref.astype(np.int32))
    d = DeviceArray.from_host(vol)
    try:
        cnt = _lib.C.c_uint()
        call("tira_lee_thin3_u8", d.ptr, *vol.shape, _lib.C.byref(cnt))
        np.testing.assert_array_equal(d.to_host(), ref)
        assert cnt.value == it_ref
    finally:
        d.free()


def test_lee_vessel_tree(hip, oracle):
    shape = (64, 80, 96)
    vol, _ = synth.vessel_volume_device(shape, seed=7, noise_sigma=0.0)
    mask = (vol.to_host() > 0.5).astype(np.uint8)
    vol.free()
    assert mask.sum() > 1000
    ref, it_ref = oracle.lee_thin3(mask)
    got, it = api.lee_thin3(mask)
    print(f"{int(mask.sum())} voxels -> {int(ref.sum())} skeleton voxels in {it_ref} passes")
    assert it == it_ref
    np.testing.assert_array_equal(got, ref.astype(np.int32))


@pytest.mark.parametrize("n,kind", [(1, "exp"), (7, "exp"), (1000, "bimodal"), (1 << 20, "exp"),
                                    (5_000_003, "bimodal"), (4096, "zeros"), (4096, "const"), (300, "edges")])
def test_otsu_bit_exact(hip, oracle, n, kind):
    rng = np.random.default_rng(n)
    if kind == "exp":
        v = rng.exponential(2.0, n)
    elif kind == "bimodal":
        v = np.abs(np.where(rng.random(n) < 0.7, rng.normal(1, 0.3, n), rng.normal(6, 1.0, n)))
    elif kind == "zeros":
        v = np.zeros(n)
    elif kind == "const":
        v = np.full(n, 3.25)
    else:
        e = np.linspace(0.0, 3.0, 257).astype(np.float32)
        v = np.concatenate([e, e[::5], [np.float32(3.0)] * 3])
    v = v.astype(np.float32)
    t = api.otsu_threshold(v)
    assert t == oracle.otsu_threshold(v)
    # strided: the stick column of a saliency triple array
    trip = np.stack([v, v * 0.5, v + 1.0], -1).astype(np.float32)
    assert api.otsu_threshold(trip, stride=3, n=v.size) == t


def test_threshold_mask(hip, oracle):
    rng = np.random.default_rng(3)
    v = rng.exponential(1.0, 100_000).astype(np.float32)
    d, thr, m = DeviceArray.from_host(v), DeviceArray((1,)), DeviceArray((v.size,), np.uint8)
    try:
        call("tira_otsu_threshold_async", d.ptr, v.size, 1, thr.ptr)
        call("tira_threshold_async", d.ptr, v.size, 1, thr.ptr, m.ptr)
        call("tira_hip_synchronize")
        t = float(thr.to_host()[0])
        np.testing.assert_array_equal(m.to_host(), (v > t).astype(np.uint8))
        assert t == oracle.otsu_threshold(v)
    finally:
        for a in (d, thr, m):
            a.free()


def _field2(H, W, seed):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, np.pi, (H, W))
    l1 = rng.uniform(-1.0, 2.0, (H, W))
    l0 = rng.uniform(-0.5, 0.5, (H, W))
    c, s = np.cos(ang), np.sin(ang)
    T = np.zeros((H, W, 2, 2))
    T[..., 0, 0] = l1 * c * c + l0 * s * s
    T[..., 0, 1] = T[..., 1, 0] = (l1 - l0) * c * s
    T[..., 1, 1] = l1 * s * s + l0 * c * c
    return T


def test_saliency_threshold2(hip, oracle):
    T = _field2(48, 64, 11)
    s = reconstruct.stick_saliency(T.astype(np.float32))
    ev = np.sort(np.abs(np.linalg.eigh(T)[0]), -1)
    ref = ev[..., 1] - ev[..., 0]
    assert np.abs(s - ref).max() <= 1e-6 * np.abs(ref).max()
    mask = reconstruct.saliency_threshold(T.astype(np.float32))
    np.testing.assert_array_equal(mask, s > oracle.otsu_threshold(s))
    np.testing.assert_array_equal(reconstruct.saliency_threshold(T.astype(np.float32), threshold=0.7), s > 0.7)
    for r in (0, 1, 3):
        got = reconstruct.saliency_threshold(T.astype(np.float32), fill=True, close_radius=r)
        np.testing.assert_array_equal(got, oracle.fill_enclosed(s > oracle.otsu_threshold(s), r))


def _spiral(n):
    """A square spiral wall of side n: a one-pixel corridor opens on the left
    edge at row 1 and winds to the centre, so a flood from the border crosses
    tile edges many times."""
    m = np.zeros((n, n), bool)
    steps = ((0, 1), (1, 0), (0, -1), (-1, 0))
    y = x = d = 0
    m[0, 0] = True
    turns = 0
    while turns < 2:
        dy, dx = steps[d]
        y1, x1, y2, x2 = y + dy, x + dx, y + 2 * dy, x + 2 * dx
        free2 = not (0 <= y2 < n and 0 <= x2 < n) or not m[y2, x2]
        if 0 <= y1 < n and 0 <= x1 < n and not m[y1, x1] and free2:
            y, x, turns = y1, x1, 0
            m[y, x] = True
        else:
            d, turns = (d + 1) % 4, turns + 1
    return m


@pytest.mark.parametrize("shape,density,r", [
    ((1, 1), 0.5, 0), ((1, 70), 0.5, 1), ((70, 1), 0.3, 0), ((37, 29), 0.4, 0), ((37, 29), 0.2, 2),
    ((100, 257), 0.45, 0), ((257, 100), 0.12, 1), ((512, 384), 0.55, 0), ((512, 384), 0.08, 3),
    ((1023, 65), 0.3, 5), ((64, 64), 0.0, 0), ((64, 64), 1.0, 0), ((64, 64), 1.0, 2)])
def test_fill_enclosed_equals_scipy(hip, oracle, shape, density, r):
    """_fill_enclosed (reconstruct.py:98-116) against scipy's binary_closing +
    binary_fill_holes, bit-exact, on random masks of ragged shapes."""
    print(f"fill shape={shape} density={density} r={r}", flush=True)
    rng = np.random.default_rng(hash((shape, r)) % (1 << 32))
    b = rng.random(shape) < density
    got, sweeps = api.fill_holes2(b, r, return_sweeps=True)
    np.testing.assert_array_equal(got, oracle.fill_enclosed(b, r))
    assert sweeps >= 1


def test_fill_enclosed_spiral(hip, oracle):
    """A corridor winding from the border to the centre: the flood crosses tile
    edges hundreds of times, so hundreds of sweeps run. Shut at its mouth, the
    whole corridor is filled."""
    m = _spiral(301)
    got, sweeps = api.fill_holes2(m, 0, return_sweeps=True)
    print(f"open spiral: sweeps={sweeps}", flush=True)
    np.testing.assert_array_equal(got, oracle.fill_enclosed(m, 0))
    assert sweeps > 100
    m[1, 0] = True
    ref = oracle.fill_enclosed(m, 0)
    assert ref.all()
    np.testing.assert_array_equal(api.fill_holes2(m, 0), ref)


def test_fill_enclosed_device_pointers(hip, oracle):
    rng = np.random.default_rng(5)
    b = (rng.random((200, 333)) < 0.4).astype(np.uint8) * 3  # nonzero bytes count as 1
    d_in, d_out = DeviceArray.from_host(b), DeviceArray(b.shape, np.uint8)
    try:
        call("tira_fill_holes2", d_in.ptr, 200, 333, 1, d_out.ptr, None)
        got = d_out.to_host()
    finally:
        d_in.free(); d_out.free()
    assert set(np.unique(got)) <= {0, 1}
    np.testing.assert_array_equal(got.astype(bool), oracle.fill_enclosed(b != 0, 1))


def test_eig3_magnitude_saliency(hip):
    t6, _ = synth.eig_tensors(50_000, seed=12, lo=0.01, hi=10.0, degenerate_frac=0.05)
    t6 = t6.copy()
    t6[::3] *= -1.0  # negative definite tensors: ranking by magnitude differs from by value
    d, o = DeviceArray.from_host(t6), DeviceArray((t6.shape[0], 3))
    # reconstruct.py ranks numpy.linalg.eigh eigenvalues, which has no absolute
    # diagonal shortcut: tensors of entries ~1e-2 need the scale-invariant
    # thresholds (the reference's absolute 1e-6 would take their diagonal).
    try:
        call("tira_eig3_async", d.ptr, t6.shape[0], _lib.LAYOUT_PACKED6, o.ptr, None, None,
             _lib.EIG_SALIENCY | _lib.EIG_MAGNITUDE | _lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT, None)
        call("tira_hip_synchronize")
        got = o.to_host()
    finally:
        d.free(); o.free()
    m = synth.mat9_from_packed6(t6).reshape(-1, 3, 3).astype(np.float64)
    a = np.sort(np.abs(np.linalg.eigh(m)[0]), -1)
    ref = np.stack([a[:, 2] - a[:, 1], a[:, 1] - a[:, 0], a[:, 0]], -1)
    err = np.abs(got - ref).max(-1) / np.abs(a).max(-1)
    assert err.max() <= 2e-6, err.max()


def test_skeleton3_pipeline(hip, oracle):
    shape = (32, 32, 32)
    f = synth.tubes_sheets_volume(shape, n_straight=3, n_helix=2, n_sheets=1, seed=13)
    voted = api.tensorvote3(f, 1.5, 0.0, 9, 1, samples=20)
    d = DeviceArray.from_host(voted.reshape(-1, 9))
    try:
        r = api.skeleton3(d, shape, layout=_lib.LAYOUT_MAT9)
        sal = r["saliency"].to_host()
        sk = r["skeleton"].to_host()
        stick = np.ascontiguousarray(sal[:, 0])
        assert r["threshold"] == oracle.otsu_threshold(stick)
        ref, it = oracle.lee_thin3((stick > r["threshold"]).reshape(shape))
        np.testing.assert_array_equal(sk, ref)
        assert r["iterations"] == it and 0 < sk.sum() < (stick > r["threshold"]).sum()
        a = np.sort(np.abs(np.linalg.eigvalsh(voted.reshape(-1, 3, 3).astype(np.float64))), -1)
        assert np.abs(stick - (a[:, 2] - a[:, 1])).max() <= 1e-5 * a.max()
    finally:
        d.free()
        for k in ("saliency", "skeleton"):
            r[k].free()
