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
