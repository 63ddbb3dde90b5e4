"""Pins the oracle's analytical tensor voting (ATV, tensorvote.h:23-760 and the
window drivers 255-306, 1298-1316, 1443-1495) without importing the reference.

What fixes the formula independently of its algebra: every ATV vote is the
sum over the voter's eigen-subspaces of k times the mean power-1 stick vote
over the unit sphere of that k-dimensional subspace,

    V = c * sum_k s_k * k * E_{q in S^{k-1}(T_k)} [ (1 - (q.v)^2) (R q)(R q)^T ],

(T_k the subspace projectors, s_k the eigenvalue gaps, R = I - 2 v v^T), because
E[q q^T] = T/k and E[(q.v)^2 q q^T] = (v^T T v T + 2 T v v^T T) / (k (k + 2)).
The tests evaluate the right-hand side by exact quadrature (two points for
k = 1, a uniform circle rule for k = 2, Gauss-Legendre x uniform azimuth for
k = 3); a dropped term, a wrong 1/(k+2) or a lost factor 2 fails them.
Further checks: the vote at the voter itself is c times the voter tensor,
d -> -d symmetry, the window indexing (single-voter fields), the FP32
restatement against the FP64 route, and the Python module's reference test
field (test_cuda_vs_python.py:48-66).
"""
import numpy as np
import pytest

from tests.conftest import rel_frob


def _stick(q, v, c):
    t = q @ v
    r = q - 2.0 * t * v
    return c * (1.0 - t * t) * np.outer(r, r)


def _circle_mean(u, w, v, c, n=16):
    b = np.arange(n) * 2.0 * np.pi / n
    return sum(_stick(np.cos(x) * u + np.sin(x) * w, v, c) for x in b) / n


def _sphere_mean(v, c):
    z, wz = np.polynomial.legendre.leggauss(6)
    m = np.zeros((3, 3))
    for zi, wi in zip(z, wz):
        s = np.sqrt(1.0 - zi * zi)
        for ph in np.arange(12) * 2.0 * np.pi / 12:
            m += wi / 2.0 / 12 * _stick(np.array([s * np.cos(ph), s * np.sin(ph), zi]), v, c)
    return m


def _up6(m):  # (00, 01, 11, 02, 12, 22), the reference's a..f order
    return np.array([m[0, 0], m[0, 1], m[1, 1], m[0, 2], m[1, 2], m[2, 2]])


def _frame(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q[:, 0], q[:, 1], q[:, 2]


@pytest.mark.parametrize("sigma", [2.5, 0.0])
def test_atv3_pair_is_subspace_mean_of_stick_votes(oracle, sigma):
    rng = np.random.default_rng(7)
    for trial in range(60):
        e0, e1, e2 = _frame(rng)
        lam = np.sort(rng.normal(size=3))
        r = rng.integers(-4, 5, size=3).astype(float)
        if trial % 10 == 0:
            r[:] = 0.0
        dist = np.linalg.norm(r)
        if sigma == 0.0:
            c = 1.0 if dist <= 1e-12 else 0.0
        else:
            c = np.exp(-dist * dist / sigma ** 2)
        if dist > 0:
            v = r / dist
            M = ((lam[2] - lam[1]) * _stick(e2, v, c) + (lam[1] - lam[0]) * 2.0 * _circle_mean(e1, e2, v, c)
                 + lam[0] * 3.0 * _sphere_mean(v, c))
        else:
            M = c * (lam[0] * np.outer(e0, e0) + lam[1] * np.outer(e1, e1) + lam[2] * np.outer(e2, e2))
        got = oracle.atv3_pair(lam, e1, e2, r, sigma, f64=True)
        ref = _up6(M)
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() <= 1e-13 * max(scale, 1.0), (trial, got, ref)
        g32 = oracle.atv3_pair(lam, e1, e2, r, sigma)
        assert np.abs(g32 - got).max() <= 2e-6 * max(np.abs(got).max(), 1e-30) + 1e-30


def test_atv2_pair_is_subspace_mean_of_stick_votes(oracle):
    rng = np.random.default_rng(8)
    for trial in range(60):
        th = rng.uniform(-np.pi, np.pi)
        e1 = np.array([np.cos(th), np.sin(th)])
        e0 = np.array([-e1[1], e1[0]])
        lam = np.sort(rng.normal(size=2))
        r = rng.integers(-5, 6, size=2).astype(float)
        if trial % 10 == 0:
            r[:] = 0.0
        dist = np.linalg.norm(r)
        c = np.exp(-dist * dist / 9.0)
        if dist > 0:
            v = r / dist
            circ = sum(_stick(np.array([np.cos(b), np.sin(b)]), v, c) for b in np.arange(16) * np.pi / 8) / 16
            M = (lam[1] - lam[0]) * _stick(e1, v, c) + lam[0] * 2.0 * circ
        else:
            M = c * (lam[0] * np.outer(e0, e0) + lam[1] * np.outer(e1, e1))
        got = oracle.atv2_pair(lam[0], lam[1], e0, e1, r, 3.0, f64=True)
        ref = np.array([M[0, 0], M[0, 1], M[1, 1]])
        assert np.abs(got - ref).max() <= 1e-13 * max(np.abs(ref).max(), 1.0), (trial, got, ref)
        g32 = oracle.atv2_pair(lam[0], lam[1], e0, e1, r, 3.0)
        assert np.abs(g32 - got).max() <= 2e-6 * max(np.abs(got).max(), 1e-30) + 1e-30


def test_atv3_pair_sign_symmetry_and_stick_limit(oracle):
    """d -> -d leaves the vote unchanged; a pure stick voter (l = 0, 0, 1) votes
    c (1 - (e2.v)^2) (R e2)(R e2)^T, the power-1 stick vote without its norm."""
    rng = np.random.default_rng(9)
    for _ in range(20):
        e0, e1, e2 = _frame(rng)
        lam = np.sort(rng.normal(size=3))
        r = rng.integers(-3, 4, size=3).astype(float)
        if not r.any():
            r[0] = 1.0
        a = oracle.atv3_pair(lam, e1, e2, r, 2.0, f64=True)
        b = oracle.atv3_pair(lam, e1, e2, -r, 2.0, f64=True)
        assert np.allclose(a, b, rtol=0, atol=1e-14)
        dist = np.linalg.norm(r)
        st = oracle.atv3_pair([0.0, 0.0, 1.0], e1, e2, r, 2.0, f64=True)
        assert np.allclose(st, _up6(_stick(e2, r / dist, np.exp(-dist ** 2 / 4.0))), rtol=0, atol=1e-14)


def test_atv3_window_single_voter(oracle):
    """One nonzero voter: every receiver in its window gets the pair vote at the
    receiver-to-voter offset, every other receiver zero (window hw = w // 2)."""
    s0, s1, s2 = 7, 8, 9
    n = s0 * s1 * s2
    rng = np.random.default_rng(10)
    e0, e1, e2 = _frame(rng)
    lam = np.array([0.2, 0.5, 1.3])
    L = np.zeros((n, 3), np.float32)
    E1 = np.zeros((n, 3), np.float32)
    E2 = np.zeros((n, 3), np.float32)
    p = (3, 2, 5)
    pi = (p[0] * s1 + p[1]) * s2 + p[2]
    L[pi], E1[pi], E2[pi] = lam, e1, e2
    for w in (5, 6):
        out = oracle.tensorvote3_atv_eig(L, E1, E2, (s0, s1, s2), 1.7, w).reshape(s0, s1, s2, 9)
        hw = w // 2
        for z in range(s0):
            for y in range(s1):
                for x in range(s2):
                    d = (p[2] - x, p[1] - y, p[0] - z)
                    if max(abs(d[0]), abs(d[1]), abs(d[2])) > hw:
                        assert not out[z, y, x].any()
                        continue
                    V = oracle.atv3_pair(L[pi], E1[pi], E2[pi], d, 1.7)
                    m = out[z, y, x]
                    assert np.allclose(m[[0, 1, 4, 2, 5, 8]], V, rtol=0, atol=1e-7)
                    assert np.array_equal(m.reshape(3, 3), m.reshape(3, 3).T)


def test_atv3_f32_route_vs_f64(oracle):
    from tiralib_amd import synth
    rng = np.random.default_rng(11)
    shape = (8, 9, 10)
    t6, _ = synth.eig_tensors(int(np.prod(shape)), seed=11, lo=0.1, hi=10.0, degenerate_frac=0.0)
    f = synth.mat9_from_packed6(t6).reshape(*shape, 9)
    for sigma, w in ((1.5, 5), (2.0, 7)):
        o32 = oracle.tensorvote3_atv(f, sigma, w, nthreads=4)
        o64 = oracle.tensorvote3_atv_f64(f, sigma, w, nthreads=4)
        e = rel_frob(o32.reshape(-1, 9), o64.reshape(-1, 9))
        assert e <= 1e-5, e
        recv = rng.integers(0, int(np.prod(shape)), size=50)
        sub = oracle.tensorvote3_atv(f, sigma, w, recv=recv)
        assert np.array_equal(sub, o32.reshape(-1, 9)[recv])


def test_atv3_sigma0_returns_voter_tensor(oracle):
    from tiralib_amd import synth
    shape = (4, 5, 6)
    t6, _ = synth.eig_tensors(int(np.prod(shape)), seed=12, lo=0.1, hi=10.0, degenerate_frac=0.0)
    f = synth.mat9_from_packed6(t6).reshape(*shape, 9)
    o = oracle.tensorvote3_atv_f64(f, 0.0, 7)
    assert np.allclose(o, f, rtol=0, atol=1e-5 * np.abs(f).max())
    o32 = oracle.tensorvote3_atv(f, 0.0, 7)
    assert np.allclose(o32, f, rtol=0, atol=1e-5 * np.abs(f).max())


def _mixed_field(H, W, seed=42):
    """make_mixed_field of the reference's test_cuda_vs_python.py:48-66."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, np.pi, (H, W))
    l1 = rng.uniform(0.5, 1.0, (H, W))
    l0 = rng.uniform(0.0, 0.5, (H, W))
    vx, vy = np.cos(ang), np.sin(ang)
    T = np.zeros((H, W, 4))
    T[..., 0] = l0 + (l1 - l0) * vx * vx
    T[..., 1] = T[..., 2] = (l1 - l0) * vx * vy
    T[..., 3] = l0 + (l1 - l0) * vy * vy
    return T.astype(np.float32)


@pytest.mark.parametrize("H", [16, 32])
def test_atv2_mixed_field_f32_vs_f64(oracle, H):
    f = _mixed_field(H, H)
    w = 2 * int(3 * 5.0) + 1
    o32 = oracle.tensorvote2_atv(f, 5.0, w, nthreads=4)
    o64 = oracle.tensorvote2_atv_f64(f, 5.0, w, nthreads=4)
    # the reference's own bar for this field is 1% of max |Python| (test_cuda_vs_python.py:141-146)
    assert np.abs(o32 - o64).max() <= 1e-5 * np.abs(o64).max()
    assert np.array_equal(o32[..., 1], o32[..., 2])


def test_atv2_window_single_voter(oracle):
    H, W = 9, 11
    lam = np.zeros((H, W, 2), np.float32)
    th = np.zeros((H, W, 2), np.float32)
    p = (4, 6)
    lam[p] = (0.3, 1.1)
    th[p] = (0.4 + np.pi / 2 - np.pi, 0.4)
    out = oracle.tensorvote2_atv_eig(lam, th, 2.0, 5)
    for x0 in range(H):
        for x1 in range(W):
            d = (p[1] - x1, p[0] - x0)
            if max(abs(d[0]), abs(d[1])) > 2:
                assert not out[x0, x1].any()
                continue
            V = oracle.atv2_pair(0.3, 1.1, (np.cos(th[p][0]), np.sin(th[p][0])), (np.cos(0.4), np.sin(0.4)), d, 2.0)
            assert np.allclose(out[x0, x1], [V[0], V[1], V[1], V[2]], rtol=0, atol=1e-7)
