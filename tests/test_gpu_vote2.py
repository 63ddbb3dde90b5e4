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
