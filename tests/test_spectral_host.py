"""Host-side parts of the spectral 3D vote (no GPU): the offset kernels and
the FFT lengths.

The spectral vote evaluates sum_d K(d) F(x + d) with features F = scale x
(quartic monomial of the voter vector) and offset kernels K fitted by the
library in FP64. These tests pin the fit against the oracle's own FP64 pair
votes, which restate the reference formulas independently of the library:
  * stick: oracle_stick3_pair_f64 (tensorvote.h:1183-1213) / sticknorm3
    (tensorvote.h:1162-1172);
  * plate: oracle_plate3_analytic_f64 (the reference's hypergeometric closed
    form, tensorvote.h:1564-1635) and oracle_plate3_sum_f64 with 20 samples
    (the numerical integral, tensorvote.h:1651-1695), both / pi.
A wrong attenuation, a dropped 1 = |q|^2 homogenisation, a transposed
component or a wrong normalisation fails the pair comparison at random unit
vectors; the symmetry tests pin what the FFT stage relies on (even kernels,
definite parity along each axis, so real spectra and one stored octant)."""
import itertools

import numpy as np
import pytest

from tiralib_amd import api

MONO = [(a, b, 4 - a - b) for a in range(4, -1, -1) for b in range(4 - a, -1, -1)]
COMP_N = np.array([[2, 0, 0], [1, 1, 0], [1, 0, 1], [0, 2, 0], [0, 1, 1], [0, 0, 2]])  # xx xy xz yy yz zz


def monos(q):
    return np.array([q[0] ** a * q[1] ** b * q[2] ** c for a, b, c in MONO])


def _atten(l2, s):
    return np.exp(-l2 / s ** 2) if s > 0 else (0.0 if l2 > 0 else 1.0)


@pytest.mark.parametrize("w,s1,s2", [(7, 1.5, 0.0), (9, 2.0, 1.2), (6, 1.0, 0.0), (5, 0.4, 0.0)])
def test_offset_kernels_reproduce_pair_votes(oracle, w, s1, s2):
    K = api.spectral_offset_kernels(w, s1, s2)
    s1, s2 = float(np.float32(s1)), float(np.float32(s2))  # the C ABI takes float sigmas
    r = w // 2
    sn = oracle.sticknorm3(s1, s2, 1)
    rng = np.random.default_rng(w)
    Q = rng.normal(size=(6, 3))
    Q /= np.linalg.norm(Q, axis=1, keepdims=True)
    worst = 0.0
    for dw, dv, du in itertools.product(range(-r, r + 1), repeat=3):
        l2 = du * du + dv * dv + dw * dw
        d = np.array([du, dv, dw], float)
        d = d / np.sqrt(l2) if l2 else d
        c1, c2 = _atten(l2, s1), _atten(l2, s2)
        k = K[:, :, dw + r, dv + r, du + r]  # (6, 30)
        for q in Q:
            m = monos(q)
            st = oracle.stick3_pair_f64(d, c1, c2, 1, q) / sn
            pa = oracle.plate3_analytic_f64(d, c1, c2, 1, q) / np.pi
            ps = oracle.plate3_sum_f64(d, c1, c2, 1, q, 20) / np.pi
            assert np.allclose(pa, ps, atol=1e-13)
            worst = max(worst, np.abs(k[:, :15] @ m - st).max(), np.abs(k[:, 15:] @ m - pa).max())
    assert worst < 1e-12, worst


def test_offset_kernels_stick_plate_subsets():
    both = api.spectral_offset_kernels(7, 1.5, 0.3)
    st = api.spectral_offset_kernels(7, 1.5, 0.3, plate=False)
    pl = api.spectral_offset_kernels(7, 1.5, 0.3, stick=False)
    assert st.shape[1] == 15 and pl.shape[1] == 15
    np.testing.assert_array_equal(st, both[:, :15])
    np.testing.assert_array_equal(pl, both[:, 15:])


def test_offset_kernel_symmetries():
    """Even under d -> -d, and of parity (-1)^(n_axis(c) + exponent_axis(f))
    under the reflection of each axis: the FFT stage stores one octant of a
    real spectrum and signs the other seven from these parities."""
    K = api.spectral_offset_kernels(9, 2.0, 0.7)
    np.testing.assert_allclose(K, K[:, :, ::-1, ::-1, ::-1], atol=1e-13)
    exps = np.array(MONO * 2)  # (30, 3): x, y, z exponents
    for axis, vol_axis in ((0, 4), (1, 3), (2, 2)):  # x = du, y = dv, z = dw
        par = (COMP_N[:, axis][:, None] + exps[:, axis][None, :]) % 2  # (6, 30)
        sign = np.where(par == 1, -1.0, 1.0)[:, :, None, None, None]
        np.testing.assert_allclose(np.flip(K, axis=vol_axis), sign * K, atol=1e-13)


def test_offset_kernels_reject_other_powers():
    from tiralib_amd._lib import TiraInvalidArgument, call
    out = np.empty(10, np.float64)
    with pytest.raises(TiraInvalidArgument):
        call("tira_spectral_offset_kernels", 7, 1.5, 0.0, 2, 1, 1, out.ctypes.data, None)


@pytest.mark.parametrize("n,w,L", [(128, 13, 135), (256, 25, 270), (512, 25, 525), (384, 25, 400), (1, 13, 7),
                                   (20, 7, 24), (1013, 25, 0)])
def test_spectral_lengths(n, w, L):
    got = api.spectral_length(n, w)
    assert got == L
    if L:
        assert L >= n + w // 2
        m = L
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        assert m == 1


@pytest.mark.parametrize("w,sigma", [(7, 1.5), (9, 2.5), (6, 1.0)])
def test_atv_offset_kernels_reproduce_pair_votes(oracle, w, sigma):
    """Analytical tensor voting: sum_f K[c,f](d) F_f with F = (l2 - l1) m(e2),
    (l1 - l0) m(e0), l0 equals the oracle's literal FP64 atv() (tensorvote.h:688-760)
    for random voters at every offset of the window, the voter itself included."""
    K = api.spectral_offset_kernels_atv(w, sigma)
    sigma = float(np.float32(sigma))
    r = w // 2
    rng = np.random.default_rng(w)
    worst, scale = 0.0, 0.0
    for _ in range(4):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = np.sort(rng.uniform(-1.0, 3.0, 3))
        e0, e1, e2 = Q[:, 0], Q[:, 1], Q[:, 2]
        F = np.concatenate([(lam[2] - lam[1]) * monos(e2), (lam[1] - lam[0]) * monos(e0), [lam[0]]])
        for dw, dv, du in itertools.product(range(-r, r + 1), repeat=3):
            got = K[:, :, dw + r, dv + r, du + r] @ F
            ref = oracle.atv3_pair(lam, e1, e2, (du, dv, dw), sigma, f64=True)[[0, 1, 3, 2, 4, 5]]
            worst = max(worst, np.abs(got - ref).max())
            scale = max(scale, np.abs(ref).max())
    assert worst <= 1e-12 * max(scale, 1.0), (worst, scale)


def _circle_average_map():
    """A[f, g] with mean over the great circle perpendicular to n of monomial f
    = sum_g A[f, g] n^g: a 16-point rule on the circle (exact for the degree-4
    trigonometric polynomials), projected onto the monomials at random unit n."""
    rng = np.random.default_rng(0)
    ns = rng.normal(size=(64, 3))
    ns /= np.linalg.norm(ns, axis=1, keepdims=True)
    avg, mn = [], []
    for n in ns:
        a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        u = np.cross(n, a)
        u /= np.linalg.norm(u)
        v = np.cross(n, u)
        b = np.arange(16) * 2.0 * np.pi / 16
        avg.append(np.mean([monos(np.cos(t) * u + np.sin(t) * v) for t in b], axis=0))
        mn.append(monos(n))
    A = np.linalg.lstsq(np.array(mn), np.array(avg), rcond=None)[0].T
    assert np.abs(np.array(mn) @ A.T - np.array(avg)).max() < 1e-13
    return A


def _flat(K, f0, f1):  # rows (component, offset), columns features f0..f1-1
    return K[:, f0:f1].transpose(0, 2, 3, 4, 1).reshape(-1, f1 - f0)


@pytest.mark.parametrize("w,s1,s2", [(13, 2.0, 0.0), (9, 1.5, 0.7), (5, 0.4, 0.0)])
def test_plate_kernels_are_circle_averages_of_stick_kernels(oracle, w, s1, s2):
    """The identity the 15-feature spectral vote rests on: the plate vote of
    normal n integrates the stick vote over the great circle perpendicular to n
    (tensorvote.h:1651-1695), so the plate offset kernels are the stick ones
    times sticknorm3 A, A the circle average of the quartic monomials. A is
    built here from its definition, independently of the library; it has 42
    nonzero entries, all multiples of 1/8 (the table the x pass applies)."""
    A = _circle_average_map()
    assert np.allclose(A * 8, np.round(A * 8), atol=1e-12) and np.count_nonzero(np.round(A * 8)) == 42
    K = api.spectral_offset_kernels(w, s1, s2)
    sn = oracle.sticknorm3(float(np.float32(s1)), float(np.float32(s2)), 1)
    Ks, Kp = _flat(K, 0, 15), _flat(K, 15, 30)
    assert np.abs(Ks @ (sn * A) - Kp).max() <= 1e-12 * np.abs(Kp).max()


@pytest.mark.parametrize("w,sigma", [(13, 2.0), (6, 1.0)])
def test_atv_parts_fold_into_the_e2_part(w, sigma):
    """Analytical voting: the (l1 - l0) e0 part is twice the circle average of
    the (l2 - l1) e2 part and the l0 part three times its sphere average
    (x^4: 1/5, x^2 y^2: 1/15), so all 31 features fold into 15."""
    A = _circle_average_map()
    mu = np.array([1 / 5 if 4 in (a, b, c) else (1 / 15 if sorted((a, b, c)) == [0, 2, 2] else 0.0)
                   for a, b, c in MONO])
    K = api.spectral_offset_kernels_atv(w, sigma)
    K1, K2, K3 = _flat(K, 0, 15), _flat(K, 15, 30), _flat(K, 30, 31)[:, 0]
    scale = np.abs(K).max()
    assert np.abs(K1 @ (2 * A) - K2).max() <= 1e-12 * scale
    assert np.abs(K1 @ (3 * mu) - K3).max() <= 1e-12 * scale
