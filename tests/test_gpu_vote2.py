"""GPU parity of the 2D TK vote (tensorvote / tensorvote_prebuilt /
tensorvote_tk and the tensorvote_tk module) against the CPU oracle, on the
cases the reference's own CPU-vs-GPU script uses (64x64 stick / ball / mixed
fields, (sigma1, sigma2, p, N) in {(5,0,1,20), (5,2,1,20), (5,2,3,20)},
test_cpu_gpu.py:184-188) plus the 256x256 Bezier workload."""
import numpy as np
import pytest

from tiralib_amd import api, synth
from tiralib_amd import tensorvote_tk as tvk
from tests.conftest import rel_frob

pytestmark = pytest.mark.gpu
NT = 8


def _fields64(seed=0):
    rng = np.random.default_rng(seed)
    H = W = 64
    th = rng.uniform(0, np.pi, (H, W))
    q = np.stack([np.cos(th), np.sin(th)], -1)
    stick = np.einsum("hwi,hwj->hwij", q, q) * rng.uniform(0.2, 1.0, (H, W, 1, 1))
    ball = np.eye(2)[None, None] * rng.uniform(0.2, 1.0, (H, W, 1, 1))
    mixed = stick + 0.3 * ball
    return {k: v.reshape(H, W, 4).astype(np.float32) for k, v in (("stick", stick), ("ball", ball), ("mixed", mixed))}


@pytest.mark.parametrize("s1,s2,p,N", [(5, 0, 1, 20), (5, 2, 1, 20), (5, 2, 3, 20)])
@pytest.mark.parametrize("kind", ["stick", "ball", "mixed"])
def test_vote2_reference_cases(hip, oracle, kind, s1, s2, p, N):
    f = _fields64()[kind]
    w = 2 * int(3 * max(s1, s2)) + 1
    gpu = api.tensorvote(f, s1, s2, p, w, True, True, N)
    ref = oracle.tensorvote2(f, s1, s2, p, w, True, True, N, nthreads=NT)
    e = rel_frob(gpu, ref)
    print(f"{kind} {s1},{s2},{p},{N}: rel Frobenius {e:.2e}")
    assert e <= 1e-4
    assert np.array_equal(gpu[..., 1], gpu[..., 2])


def test_vote2_bezier_config(hip, oracle):
    """256x256, 8 Bezier curves + 2% noise sticks, sigma=3, w=19, stick only,
    then eigen2 saliency l1 - l0."""
    f = synth.bezier_field(seed=synth.SEED_CFG1)
    gpu = api.tensorvote(f, 3.0, 0.0, 1, 19, True, False, 20)
    ref = oracle.tensorvote2(f, 3.0, 0.0, 1, 19, True, False, 20, nthreads=NT)
    ref64 = oracle.tensorvote2_f64(f, 3.0, 0.0, 1, 19, True, False, 20, nthreads=NT)
    e32, e64 = rel_frob(gpu, ref), rel_frob(gpu, ref64)
    print(f"config-1 vote: vs FP32 oracle {e32:.2e}, vs FP64 {e64:.2e}")
    assert e32 <= 1e-4 and e64 <= 1e-4
    sal = api.evals2_symmetric(gpu.reshape(-1, 4)).astype(np.float64)
    sal_ref = oracle.evals2(ref.reshape(-1, 4)).astype(np.float64)
    s, sr = sal[:, 1] - sal[:, 0], sal_ref[:, 1] - sal_ref[:, 0]
    assert np.abs(s - sr).max() <= 1e-4 * np.abs(sr).max()


def test_vote2_prebuilt_and_module(hip, oracle):
    f = _fields64(1)["mixed"]
    a = api.tensorvote(f, 4.0, 1.0, 2, 0, True, True, 20)
    nb = oracle.build_neighbors2d(25, 4.0, 1.0, 2, 20)
    b = api.tensorvote_prebuilt(f, nb, 4.0, 1.0, 2, 25, True, True, 20)
    assert rel_frob(b, a) <= 1e-6
    c = tvk.tensorvote_tk(f, 4.0, 1.0, 2, 0, True, True, 20)
    np.testing.assert_array_equal(c, a)
    d = tvk.tensorvote_tk_prebuilt(f, nb.tobytes(), len(nb), 25, 4.0, 1.0, 2, True, True, 20)
    np.testing.assert_array_equal(d, b)


def test_vote2_tk_entry_and_flags(hip, oracle):
    f = _fields64(2)["mixed"]
    lam = oracle.evals2(f.reshape(-1, 4))
    th = oracle.evecs2polar(f.reshape(-1, 4), lam)
    for stick, plate in ((True, True), (True, False), (False, True)):
        g = api.tensorvote_tk(lam.reshape(64, 64, 2), th.reshape(64, 64, 2), (64, 64), 3.0, 0.5, 1, 19, stick,
                              plate, 20)
        r = np.zeros((64, 64, 4), np.float32)
        import ctypes as C
        nb = oracle.build_neighbors2d(19, 3.0, 0.5, 1, 20, stick, plate)
        oracle.lib().oracle_tensorvote2_tk(r, lam, th, 3.0, 0.5, 1, 19, 64, 64, int(stick), int(plate), 20,
                                           nb.ctypes.data, len(nb), NT)
        assert rel_frob(g, r) <= 1e-4
    z = api.tensorvote(f, 3.0, 0.0, 1, 9, False, False, 20)
    assert np.all(z == 0)


def test_vote2_odd_shapes(hip, oracle):
    rng = np.random.default_rng(9)
    for H, W in ((1, 1), (3, 70), (65, 5), (130, 129)):
        f = _fields64()["mixed"]
        g = np.resize(f, (H, W, 4)).astype(np.float32)
        a = api.tensorvote(g, 2.0, 0.0, 1, 9, True, True, 10)
        b = oracle.tensorvote2(g, 2.0, 0.0, 1, 9, True, True, 10)
        assert rel_frob(a, b) <= 1e-4


def test_vote2_table_kept_between_calls(hip):
    """The device copy of the neighbour table is reused while the table is
    unchanged: repeated calls, a parameter change, an ATV call in between and
    a workspace release all give the results of fresh calls."""
    from tiralib_amd._lib import call
    f = _fields64(3)["mixed"]
    a1 = api.tensorvote(f, 4.0, 1.0, 1, 25, True, True, 20)
    a2 = api.tensorvote(f, 4.0, 1.0, 1, 25, True, True, 20)
    np.testing.assert_array_equal(a1, a2)
    b = api.tensorvote(f, 3.0, 0.0, 1, 19, True, True, 20)  # another table
    assert not np.array_equal(a1, b)
    t1 = api.tensorvote(f, 2.0, 0.0, 1, 13, True, True, 20, is_atv=True)
    a3 = api.tensorvote(f, 4.0, 1.0, 1, 25, True, True, 20)
    np.testing.assert_array_equal(a1, a3)
    t2 = api.tensorvote(f, 3.0, 0.0, 1, 19, True, True, 20, is_atv=True)  # another ATV table
    assert not np.array_equal(t1, t2)
    call("tira_hip_release_workspace")
    a4 = api.tensorvote(f, 4.0, 1.0, 1, 25, True, True, 20)
    np.testing.assert_array_equal(a1, a4)
    b2 = api.tensorvote(f, 3.0, 0.0, 1, 19, True, True, 20)
    np.testing.assert_array_equal(b, b2)
    np.testing.assert_array_equal(t1, api.tensorvote(f, 2.0, 0.0, 1, 13, True, True, 20, is_atv=True))
    np.testing.assert_array_equal(t2, api.tensorvote(f, 3.0, 0.0, 1, 19, True, True, 20, is_atv=True))


def _search_params(k, seed):
    """Parameter sets as a differential-evolution population draws them over
    optimize_tk's bounds (optimize.py:258-259), sigma1 capped for test time."""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(1.0, 6.0)), float(rng.uniform(0.0, 4.0)), int(round(rng.uniform(1, 6))))
            for _ in range(k)]


def test_vote2_session_equals_module_calls(hip, oracle):
    """Session votes (one graph replay each, cached tables) against the module
    functions on the same field, with other library work and waits between
    replays: bit for bit with the caller's table, within 1e-6 of the largest
    component with the table built on the device; repeated parameter sets hit
    the table cache."""
    f = synth.bezier_field(seed=synth.SEED_CFG1)
    other = _fields64(3)["mixed"]
    params = _search_params(12, 42)
    params += params[:4]  # repeats: served from the cache
    with tvk.Session(f) as s:
        for i, (s1, s2, p) in enumerate(params):
            w = 2 * int(3 * max(s1, s2)) + 1
            nb = oracle.build_neighbors2d(w, s1, s2, p, 20)
            got = s.tensorvote_tk_prebuilt(nb.tobytes(), nb.size, w, s1, s2, p, plate=True, samples=20)
            ref = tvk.tensorvote_tk_prebuilt(f, nb.tobytes(), nb.size, w, s1, s2, p, plate=True, samples=20)
            np.testing.assert_array_equal(got, ref, err_msg=f"replay {i} ({s1}, {s2}, {p})")
            got2 = s.tensorvote_tk(s1, s2, p, stick=(i % 3 != 0), plate=(i % 3 != 1))
            ref2 = tvk.tensorvote_tk(f, s1, s2, p, stick=(i % 3 != 0), plate=(i % 3 != 1))
            # table built on the device: c1, c2 may sit one ulp from the host's
            assert np.abs(got2 - ref2).max() <= 1e-6 * np.abs(ref2).max(), f"replay {i} built table"
            api.tensorvote(other, 2.0, 0.0, 1, 0, True, True, 20)  # eager work between replays
        replays, hits, builds = s.stats()
    assert replays == 2 * len(params)
    assert hits >= 8 and builds <= 2 * 12
    # the session's own result against the oracle on one parameter set
    s1, s2, p = params[0]
    w = 2 * int(3 * max(s1, s2)) + 1
    with tvk.Session(f) as s:
        got = s.tensorvote_tk(s1, s2, p)
    assert rel_frob(got, oracle.tensorvote2(f, s1, s2, p, w, True, True, 20, nthreads=NT)) <= 1e-4


def test_vote2_session_device_output_and_eviction(hip, oracle):
    """Device output through the C ABI, more distinct tables than the cache
    holds (the first is rebuilt after eviction), and odd field shapes."""
    from tiralib_amd._lib import DeviceArray, call
    import ctypes as C
    f = _fields64(4)["mixed"][:37, :53].copy()
    h = C.c_void_p()
    call("tira_vote2_session_create", C.byref(h), f.ctypes.data, 37, 53)
    out = DeviceArray((37, 53, 4))
    try:
        params = [(1.0 + 0.1 * k, 0.5, 1 + k % 3) for k in range(70)] + [(1.0, 0.5, 1)]
        for k, (s1, s2, p) in enumerate(params):
            w = 2 * int(3 * max(s1, s2)) + 1
            call("tira_vote2_session_vote", h, out.ptr, None, 0, s1, s2, p, w, 1, 1, 20)
            ref = tvk.tensorvote_tk(f, s1, s2, p, w=w)
            got = out.to_host()
            assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max(), f"call {k}"
            if k % 10 == 0:  # a host output in between rebuilds the graph both ways
                host = np.empty((37, 53, 4), np.float32)
                call("tira_vote2_session_vote", h, host.ctypes.data, None, 0, s1, s2, p, w, 1, 1, 20)
                np.testing.assert_array_equal(host, got)
        r, hits, builds = (C.c_ulonglong(), C.c_ulonglong(), C.c_ulonglong())
        call("tira_vote2_session_stats", h, C.byref(r), C.byref(hits), C.byref(builds))
        assert builds.value == 71 and hits.value == 8  # 64 tables kept: the first is rebuilt
    finally:
        out.free()
        call("tira_vote2_session_destroy", h)


def test_vote2_session_errors(hip):
    f = _fields64(5)["stick"]
    s = tvk.Session(f)
    nb = np.zeros(3, tvk.NEIGHBOR2D_DTYPE)
    with pytest.raises(RuntimeError, match="nb_buf byte size"):
        s.tensorvote_tk_prebuilt(nb.tobytes(), 4, 3, 1.0)
    with pytest.raises(RuntimeError, match="sigma1 must be > 0"):
        s.tensorvote_tk(0.0)
    s.close()
    with pytest.raises(RuntimeError, match="closed"):
        s.tensorvote_tk(1.0)


def test_build_neighbors2d_device_equals_host(hip):
    """The device table build against the host build (the reference's
    arithmetic): bit-identical entries, c1 and c2 within one ulp (glibc's expf
    against the correctly rounded exponential), over a search population."""
    from tiralib_amd._lib import DeviceArray, call
    params = _search_params(40, 7) + [(3.0, 0.0, 1), (5.0, 2.0, 3), (0.5, 0.0, 1), (30.0, 20.0, 30)]
    ulps = 0
    for s1, s2, p in params:
        for stick, plate, N in ((1, 1, 20), (1, 0, 20), (0, 1, 7)):
            w = 2 * int(3 * max(s1, s2)) + 1
            host = api.build_neighbors2d(w, s1, s2, p, N, stick, plate)
            d = DeviceArray((w * w * 12,))
            try:
                call("tira_build_neighbors2d_async", w, s1, s2, p, N, stick, plate, d.ptr)
                call("tira_hip_synchronize")
                dev = d.to_host().view(tvk.NEIGHBOR2D_DTYPE).reshape(-1)
            finally:
                d.free()
            for k in ("du", "dv", "dx", "dy", "Ra", "Rb", "Rc", "Va", "Vb", "Vc"):
                np.testing.assert_array_equal(dev[k], host[k], err_msg=f"{k} at {(s1, s2, p, stick, plate)}")
            for k in ("c1", "c2"):
                diff = np.abs(dev[k].view(np.int32).astype(np.int64) - host[k].view(np.int32).astype(np.int64))
                assert diff.max() <= 1, (k, s1, s2, p)
                ulps += int((diff > 0).sum())
    print(f"c1/c2 entries one ulp apart: {ulps}")


def _rng42_fields(H, W):
    """The two fields of the reference's CUDA-vs-Python script
    (test_cuda_vs_python.py:48-85): mixed stick/ball and pure stick,
    default_rng(42), orientations and eigenvalues uniform."""
    out = {}
    for kind in ("mixed", "stick"):
        rng = np.random.default_rng(42)
        ang = rng.uniform(0, np.pi, (H, W))
        l1 = rng.uniform(0.5, 1.0, (H, W))
        l0 = rng.uniform(0.0, 0.5, (H, W)) if kind == "mixed" else np.zeros((H, W))
        vx, vy = np.cos(ang), np.sin(ang)
        T = np.zeros((H, W, 4))
        T[..., 0] = l0 + (l1 - l0) * vx * vx
        T[..., 1] = T[..., 2] = (l1 - l0) * vx * vy
        T[..., 3] = l0 + (l1 - l0) * vy * vy
        out[kind] = T.astype(np.float32)
    return out


@pytest.mark.parametrize("H", [16, 32, 64, 128])
def test_vote2_reference_script_sweep(hip, oracle, H):
    """Sections 2 and 3 of test_cuda_vs_python.py (sizes 16..128, sigma1 5,
    sigma2 2, p 1, N 20; stick-only on the pure-stick field, stick + plate on
    the mixed one) through the pybind-module mirror, against the oracle at
    1e-4 relative Frobenius (the script's own bar is 1%)."""
    f = _rng42_fields(H, H)
    w = 2 * int(3 * 5.0) + 1
    for kind, plate in (("stick", False), ("mixed", True)):
        gpu = tvk.tensorvote_tk(f[kind], 5.0, 2.0, 1, 0, True, plate, 20)
        ref = oracle.tensorvote2(f[kind], 5.0, 2.0, 1, w, True, plate, 20, nthreads=NT)
        e = rel_frob(gpu, ref)
        print(f"{kind} {H}x{H}: rel Frobenius {e:.2e}")
        assert e <= 1e-4
