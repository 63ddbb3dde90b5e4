"""CPU checks that pin the NumPy restatement of the plate-vote network
(oracle/mlp_ref.py) before it judges the MFMA kernels: BF16 rounding against
hand-worked values, backward pass against central finite differences (FP64,
no rounding), the Adam step against its closed form on the first step, and
the plate-frame target against the FP64 plate integral it rotates."""
import numpy as np

from oracle import mlp_ref as M


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    got = M.bf16(x)
    want = np.array([1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5, np.float32(3.0e38), 0.0], np.float64)
    # 3e38 rounds to the nearest BF16 (0x7F62 -> 3.0019e38); 1e-40 is a subnormal that BF16 keeps to 7 bits
    want[5] = float(np.array([0x7F620000], np.uint32).view(np.float32)[0])
    want[6] = float(M.bf16_bits_to_f32(np.array([(np.array([1e-40], np.float32).view(np.uint32)[0] + 0x7FFF) >> 16]))[0])
    assert np.array_equal(got.astype(np.float64), want)


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.NPARAM).astype(np.float32)
    return p


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = _random_params(5).astype(np.float64)
    X = rng.normal(0, 1, (40, 32))
    T = rng.normal(0, 1, (40, 8))
    g = M.gradients(p, X, T, exact=True)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(M.forward(pp, X, True)[0], T) - M.loss(M.forward(pm, X, True)[0], T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(6)
    p = _random_params(6).astype(np.float64)
    g = rng.normal(0, 1, M.NPARAM)
    lr = 1e-3
    p1, m1, v1 = M.adam_step(p, g, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 1, lr)
    msk = M.mask()
    geff = g.copy()
    geff[0:2048].reshape(64, 32)[:, :5] += g[0:2048].reshape(64, 32)[:, 5:10]
    # t = 1: m = (1-b1) g, v = (1-b2) g^2 -> step = lr * g / (|g| + eps), ~ lr * sign(g)
    step = (p - p1)[msk]
    ge = geff[msk]
    assert np.allclose(step, lr * ge / (np.abs(ge) + 1e-8), rtol=1e-9, atol=1e-15)
    assert np.mean(np.isclose(step, lr * np.sign(ge), rtol=1e-3)) > 0.999
    assert np.array_equal(p1[~msk & ~_residual_cols()], p[~msk & ~_residual_cols()])
    W1 = p1[0:2048].reshape(64, 32)
    assert np.array_equal(W1[:, 5:10], W1[:, :5])


def _residual_cols():
    r = np.zeros(M.NPARAM, bool)
    r[0:2048].reshape(64, 32)[:, 5:10] = True
    return r


def test_target_rotates_back_to_the_integral(oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        d = rng.uniform(-6, 6, 3)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        o, c1 = M.target(d, n, 2.0, 0.0, 1, 64, oracle.plate3_sum_f64)
        want = oracle.plate3_sum_f64(d / np.linalg.norm(d), c1, 0.0, 1, n, 64)
        got = M.to_field(o, n, d) * c1
        # in the aligned frame the offset has no v component and the vote none of uv, vn
        assert abs(o[1]) <= 1e-12 * np.abs(o).max() and abs(o[4]) <= 1e-12 * np.abs(o).max()
        assert np.allclose(got, want, rtol=1e-12, atol=1e-14 * max(1, np.abs(want).max()))
        # the p = 1 plate integral is exact for N >= 3: the analytic form agrees
        ana = oracle.plate3_analytic_f64(d / np.linalg.norm(d), c1, 0.0, 1, n)
        assert np.allclose(got, ana, rtol=1e-9, atol=1e-12)


def test_features_and_input_row():
    n = np.array([0.0, 0.6, 0.8])
    d = np.array([1.0, 2.0, -2.0])
    f = M.features(d, n, 2.0)
    assert np.isclose(f[0] ** 2 + f[1] ** 2 + f[2] ** 2, 1.0)
    assert f[0] >= 0 and abs(f[1]) < 1e-12
    assert np.isclose(f[3], 1.5) and np.isclose(f[4], 2 * f[2] ** 2 - 1)
    row = M.input_row(f)
    assert np.allclose(row[:5] + row[5:10], f, rtol=0, atol=2e-5)
    assert row[10] == 1.0 and not row[11:].any()


def _exact_row(oracle, power=1, rho=1.0, M_=M.TAB_M):
    """Table row built from the FP64 integral (64 samples) in place of the
    network: the plate-frame targets at the grid's features."""
    Y = []
    for f in M.table_features(rho, M_):
        sn, _, s = f[0], f[1], f[2]
        d = np.array([sn, 0.0, s]) * (rho * 2.0)  # |d| = rho sigma1 with sigma1 = 2
        t, _ = M.target(d, np.array([0.0, 0.0, 1.0]), 2.0, 0.0, power, 64, oracle.plate3_sum_f64)
        Y.append(t)
    return M.table_row(np.array(Y), M_)


def _random_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def test_table_conversion_exact_at_grid_points(oracle):
    """At the cell centres the tabulated form A I + B nn + s g (nd + dn) + D dd
    rebuilt from the exact integral's plate-frame components equals the FP64
    integral in any orientation: a dropped term, a wrong sign of s or of the
    odd part, or a lost factor 2 in D/2 fails."""
    rng = np.random.default_rng(3)
    for power in (1, 2):
        row = _exact_row(oracle, power)
        x = M.table_grid()
        for j in (0, 1, 17, 40, 62, 63):
            for sg in (1.0, -1.0):
                s = sg * np.sqrt(1.0 - x[j])
                n = _random_unit(rng)
                w = _random_unit(rng)
                w -= (w @ n) * n
                w /= np.linalg.norm(w)
                d = s * n + np.sqrt(x[j]) * w
                got = M.table_vote(row[j], n, d)
                want = oracle.plate3_sum_f64(d, 1.0, 0.0, power, n, 64)
                assert np.allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max()), (power, j, sg)


def test_table_interpolation_error_bound(oracle):
    """Between and beyond the cell centres (linear in x, extrapolated over the
    end half cells) the tabulated exact integral stays within the
    interpolation bound: RMS 2e-4 and at most 1e-3 relative per pair for
    p = 1 at 64 cells (x-polynomials of degree <= 2), at s = 0 and s = +-1 too."""
    rng = np.random.default_rng(4)
    row = _exact_row(oracle, 1)
    num = den = 0.0
    worst = 0.0
    for s in np.concatenate([rng.uniform(-1, 1, 400), [0.0, 1.0, -1.0, 1e-4, -0.9999]]):
        n = _random_unit(rng)
        w = _random_unit(rng)
        w -= (w @ n) * n
        w /= np.linalg.norm(w)
        d = s * n + np.sqrt(max(0.0, 1.0 - s * s)) * w
        d /= np.linalg.norm(d)
        got = M.table_vote(M.table_lookup(row, 1.0 - (n @ d) ** 2), n, d)
        want = oracle.plate3_sum_f64(d, 1.0, 0.0, 1, n, 64)
        e = M.COMP_W @ (got - want) ** 2
        t = M.COMP_W @ want ** 2
        num += e
        den += t
        worst = max(worst, np.sqrt(e / t))
    assert np.sqrt(num / den) <= 2e-4 and worst <= 1e-3, (np.sqrt(num / den), worst)


def test_table_shells_are_sums_of_three_squares():
    """Rows 1..K: the distinct squared lattice distances inside 3 sigma1,
    checked against Legendre's three-square theorem (n is a sum of three
    squares unless n = 4^a (8b + 7))."""
    for sigma in (1.0, 1.5, 2.0, 4.0):
        got = M.shells(sigma)
        want = []
        for l2 in range(1, int(9 * sigma * sigma) + 2):
            q = l2
            while q % 4 == 0:
                q //= 4
            if q % 8 != 7 and np.sqrt(np.float32(l2)) / np.float32(sigma) < np.float32(3.0):
                want.append(l2)
        assert np.array_equal(got, want), sigma
