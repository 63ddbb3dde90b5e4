"""Pins the oracle's 3D stick / plate vote without retyping its formulas:
lone-voter fields against the single-pair kernels (driver indexing, window,
normalisation, eigenvalue scales), the normalisation against its closed-form
integral, geometric properties of the stick kernel, three independent
formulations of the integrated plate, and symmetry / axis-permutation
equivariance of the whole vote."""
import math

import numpy as np
import pytest

from tests.conftest import rel_frob

SLOTS = [0, 1, 2, 4, 5, 8]


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_stick_pair_geometry(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        q, d = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        c1, c2 = rng.uniform(0, 1, 2)
        V = oracle.stick3_pair_f64(d, c1, c2, 1, q)
        M = np.array([[V[0], V[1], V[2]], [V[1], V[3], V[4]], [V[2], V[4], V[5]]])
        w, U = np.linalg.eigh(M)
        r = U[:, 2]
        t = q @ d
        # rank one along the reflection of q in the plane orthogonal to d
        assert abs(w[0]) < 1e-12 and abs(w[1]) < 1e-12
        assert abs(abs(r @ d) - abs(t)) < 1e-9
        assert abs(w[2] - (c1 * (1 - t * t) + c2 * t * t)) < 1e-12
    q = _unit([1, 2, 3])
    d_perp = _unit(np.cross(q, [0, 0, 1]))
    V = oracle.stick3_pair_f64(d_perp, 0.7, 0.2, 3, q)
    np.testing.assert_allclose(V, 0.7 * np.outer(q, q)[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]], atol=1e-15)
    V = oracle.stick3_pair_f64(q, 0.7, 0.2, 3, q)
    np.testing.assert_allclose(V, 0.2 * np.outer(q, q)[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]], atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_sticknorm3_is_the_vote_integral(oracle, p):
    """1/sticknorm3 normalises the stick field (tensorvote.h:1162-1172): the
    integral over R^3 of trace(vote) of a unit stick equals sticknorm3. The
    lattice sum (origin excluded) converges as 1 + a / sigma^3; Richardson
    extrapolation over sigma and 2 sigma removes that term."""
    q = _unit([0.3, -0.5, 0.8])

    def lattice_sum(s1, s2):
        R = int(6.5 * s1)
        g = np.arange(-R, R + 1, dtype=np.float64)
        Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
        L2 = X * X + Y * Y + Z * Z
        L = np.sqrt(L2)
        t = np.where(L > 0, (X * q[0] + Y * q[1] + Z * q[2]) / np.where(L > 0, L, 1), 0.0)
        tr = np.exp(-L2 / s1 ** 2) * (1 - t * t) ** p + np.exp(-L2 / s2 ** 2) * (t * t) ** p
        tr[R, R, R] = 0.0
        return tr.sum() / oracle.sticknorm3(s1, s2, p)

    a, b = lattice_sum(3.0, 1.5), lattice_sum(6.0, 3.0)
    assert abs((8 * b - a) / 7 - 1) < 2e-4


def test_lone_stick_voter_field(oracle):
    """Single stick voter s q q^T: every receiver gets scale/sticknorm * the
    pair kernel at its offset; receivers outside the window get nothing."""
    n, w, s1 = 15, 9, 1.7
    q = _unit([0.2, 0.9, -0.4])
    vol = np.zeros((n, n, n, 9), np.float32)
    c = n // 2
    vol[c, c, c] = (2.5 * np.outer(q, q)).reshape(9)
    out = oracle.tensorvote3(vol, s1, 0.0, 1, w, stick=True, plate=False).reshape(n, n, n, 9)
    sn = oracle.sticknorm3(s1, 0.0, 1)
    ev = oracle.evals3(vol[c, c, c].reshape(1, 9))[0]
    scale = abs(ev[2]) - abs(ev[1])
    r = w // 2
    for z in range(n):
        for y in range(n):
            for x in range(n):
                o = np.array([c - x, c - y, c - z], np.float64)  # voter - receiver, in (x, y, z)
                got = out[z, y, x][SLOTS]
                if np.max(np.abs(o)) > r:
                    assert np.all(got == 0)
                    continue
                L = np.linalg.norm(o)
                d = o / L if L > 0 else np.zeros(3)
                exp = scale / sn * oracle.stick3_pair_f64(d, math.exp(-L * L / s1 ** 2), 1.0 if L == 0 else 0.0, 1, q)
                np.testing.assert_allclose(got, exp, rtol=2e-5, atol=2e-7)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_plate_three_formulations(oracle, p):
    """FP64 literal Riemann sum, the reference's analytic hypergeometric form
    (restated) and this project's moment closed form agree once N >= p+2;
    below that the Riemann sum genuinely differs."""
    rng = np.random.default_rng(p)

    def gk(k):
        return math.comb(2 * k, k) / 4 ** k

    def closed(d, c1, c2, n):
        e = [c1 * math.comb(p, j) * (-1) ** j + (c2 if j == p else 0) for j in range(p + 1)]
        s = n @ d
        x = 1 - s * s
        A = sum(math.pi * e[j] * (gk(j) - gk(j + 1)) * x ** j for j in range(p + 1))
        B = sum(math.pi * e[j] * (2 * gk(j + 1) - gk(j)) * x ** (j - 1) for j in range(1, p + 1))
        C = sum(-2 * math.pi * e[j] * gk(j + 1) * x ** j for j in range(p + 1))
        D = sum(4 * math.pi * e[j] * gk(j + 1) * x ** (j + 1) for j in range(p + 1))
        w = d - s * n
        M = A * (np.eye(3) - np.outer(n, n)) + B * np.outer(w, w) + C * (np.outer(w, d) + np.outer(d, w)) + D * np.outer(d, d)
        return M[[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]

    for _ in range(100):
        d, n = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        c1, c2 = rng.uniform(0, 1, 2)
        s64 = oracle.plate3_sum_f64(d, c1, c2, p, n, 97)
        for N in (p + 2, 20):
            np.testing.assert_allclose(oracle.plate3_sum_f64(d, c1, c2, p, n, N), s64, atol=1e-13)
        np.testing.assert_allclose(oracle.plate3_analytic_f64(d, c1, c2, p, n), s64, atol=1e-13)
        np.testing.assert_allclose(closed(d, c1, c2, n), s64, atol=1e-13)
    d, n = _unit([0.3, 0.5, 0.81]), _unit([0.1, 0.9, 0.2])
    assert np.abs(oracle.plate3_sum_f64(d, 0.7, 0.2, p, n, p + 1) - oracle.plate3_sum_f64(d, 0.7, 0.2, p, n, 97)).max() > 1e-4


@pytest.mark.parametrize("p", [1, 2, 3, 5, 8])
def test_plate_analytic_special_functions_vs_scipy(oracle, p):
    """The Boost.Math terms of the reference's analytic plate (tensorvote.h:1589-1592,
    1722-1723): J0 = pi/2 2F1(-p, 3/2; 2; a^2), J1 = pi 2F1(-p, 1/2; 1; a^2),
    K0 = B(1/2, p + 3/2), K1 = B(1/2, p + 1/2), here from scipy.special (hyp2f1 and
    beta, implementations independent of the oracle's terminating series and
    log-Gamma form). For a unit offset d and the normal n = z the reference's local
    matrices are rank-one blocks plus the y term: c1 A has the eigenvalues
    {J0, J1 - J0, 0} and c2 B {a^2p K0, a^2p (K1 - K0), 0}, since
    (1 - 2 a^2)^2 + 4 a^2 d_z^2 = 1 when |d| = 1."""
    from scipy import special
    n = np.array([0.0, 0.0, 1.0])
    rng = np.random.default_rng(40 + p)
    for _ in range(20):
        d = _unit(rng.normal(size=3))
        a2 = 1.0 - d[2] ** 2
        J0 = 0.5 * np.pi * special.hyp2f1(-p, 1.5, 2.0, a2)
        J1 = np.pi * special.hyp2f1(-p, 0.5, 1.0, a2)
        K0, K1 = special.beta(0.5, p + 1.5), special.beta(0.5, p + 0.5)
        for c1, c2, want in ((1.0, 0.0, [J0, J1 - J0, 0.0]),
                             (0.0, 1.0, [a2 ** p * K0, a2 ** p * (K1 - K0), 0.0])):
            u = oracle.plate3_analytic_f64(d, c1, c2, p, n)
            M = np.array([[u[0], u[1], u[2]], [u[1], u[3], u[4]], [u[2], u[4], u[5]]])
            np.testing.assert_allclose(np.linalg.eigvalsh(M), np.sort(want), rtol=0, atol=1e-11)


def test_plate_fp32_formulation_matches_fp64(oracle):
    rng = np.random.default_rng(9)
    for _ in range(200):
        d, n = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        c1 = float(rng.uniform(0, 1))
        v32 = oracle.plate3_numerical_f32(d.astype(np.float32), c1, 0.0, 1, n.astype(np.float32), 20)
        v64 = oracle.plate3_sum_f64(d, c1, 0.0, 1, n, 20)
        assert np.abs(v32 - v64).max() <= 2e-6 * max(np.abs(v64).max(), 1e-6)


def _field(shape, seed):
    from tiralib_amd import synth
    t6, _ = synth.eig_tensors(int(np.prod(shape)), seed=seed, lo=0.1, hi=10.0, degenerate_frac=0.0)
    return synth.mat9_from_packed6(t6).reshape(*shape, 9)


def test_fp32_route_matches_fp64_route(oracle):
    f = _field((9, 10, 11), 5)
    a = oracle.tensorvote3(f, 1.5, 0.4, 1, 7, samples=20, nthreads=4).reshape(-1, 9)[:, SLOTS]
    b = oracle.tensorvote3_f64(f, 1.5, 0.4, 1, 7, samples=20, nthreads=4).reshape(-1, 9)[:, SLOTS]
    assert rel_frob(a, b) < 1e-4


def test_point_reflection_symmetry(oracle):
    """Votes depend on the offset only through d d^T: reflecting the field
    through the volume centre reflects the result."""
    f = _field((8, 9, 10), 6)
    a = oracle.tensorvote3_f64(f, 1.3, 0.0, 1, 5, samples=20, nthreads=4)
    b = oracle.tensorvote3_f64(np.ascontiguousarray(f[::-1, ::-1, ::-1]), 1.3, 0.0, 1, 5, samples=20, nthreads=4)
    np.testing.assert_allclose(b[::-1, ::-1, ::-1], a, atol=1e-12 * np.abs(a).max())


def test_axis_permutation_equivariance(oracle):
    """Swapping the slowest and fastest volume axes together with the x and z
    tensor components permutes the vote the same way (catches transposed
    offsets / components)."""
    f = _field((7, 8, 9), 7)
    perm = [2, 1, 0]

    def swap_xz(t9):
        M = t9.reshape(*t9.shape[:-1], 3, 3)
        return np.ascontiguousarray(M[..., perm, :][..., :, perm].reshape(*t9.shape[:-1], 9))

    a = oracle.tensorvote3_f64(f, 1.2, 0.5, 2, 5, samples=20, nthreads=4)
    fp = swap_xz(np.ascontiguousarray(np.transpose(f, (2, 1, 0, 3))))
    b = oracle.tensorvote3_f64(fp, 1.2, 0.5, 2, 5, samples=20, nthreads=4)
    back = swap_xz(np.ascontiguousarray(np.transpose(b, (2, 1, 0, 3))))
    np.testing.assert_allclose(back, a, atol=1e-11 * np.abs(a).max())
