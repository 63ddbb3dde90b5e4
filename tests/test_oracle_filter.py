"""Pins the oracle's Gaussian kernel and separable valid-mode convolution:
bit-exact against filter.h (golden vectors / live build), and against
scipy.ndimage.correlate1d and the normal density."""
import os

import numpy as np
import pytest
from scipy import ndimage

GOLD = os.path.join(os.path.dirname(__file__), "golden", "eigen_filter_ref.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_gaussian1d_golden(oracle, gold):
    for i, s in enumerate(gold["sigmas"]):
        k = oracle.gaussian_ksize(float(s))
        np.testing.assert_array_equal(oracle.gaussian1d(k, 0.0, float(s), 1.0), gold[f"g1d_{i}"])


def test_gaussian_convolve3_golden(oracle, gold):
    sw, sh, sd = (float(x) for x in gold["gconv_sigmas"])
    np.testing.assert_array_equal(oracle.gaussian_convolve3(gold["vol"], sw, sh, sd), gold["gconv"])


def test_ksize_rule(oracle):
    # ceil(6 sigma) forced odd (filter.h:178-183): sigma 1.5 -> 9 taps, 2 -> 13
    assert oracle.gaussian_ksize(1.5) == 9
    assert oracle.gaussian_ksize(2.0) == 13
    assert oracle.gaussian_ksize(0.5) == 3


def test_gaussian1d_matches_normal_density(oracle):
    for s in (0.8, 1.5, 2.0, 4.0):
        k = 11 if s < 2 else oracle.gaussian_ksize(s)
        g = oracle.gaussian1d(k, 0.0, s, 1.0).astype(np.float64)
        x = np.arange(k) - (k - 1) / 2.0
        ref = np.exp(-x * x / (2 * s * s))
        np.testing.assert_allclose(g, ref / ref.sum(), rtol=2e-6, atol=1e-9)
        assert abs(g.sum() - 1) < 1e-6
        np.testing.assert_allclose(g, g[::-1], rtol=1e-6)


def test_convolve3_vs_scipy(oracle):
    rng = np.random.default_rng(0)
    vol = rng.normal(size=(12, 13, 17)).astype(np.float32)
    g = oracle.gaussian1d(7, 0.0, 1.2, 1.0)
    out = oracle.convolve3(vol, g, (7, 1, 1))        # along x (fastest)
    ref = ndimage.correlate1d(vol.astype(np.float64), g.astype(np.float64), axis=2, mode="constant")[:, :, 3:-3]
    np.testing.assert_allclose(out, ref, atol=1e-5)
    out = oracle.convolve3(vol, g, (1, 1, 7))        # along z (slowest)
    ref = ndimage.correlate1d(vol.astype(np.float64), g.astype(np.float64), axis=0, mode="constant")[3:-3]
    np.testing.assert_allclose(out, ref, atol=1e-5)


def test_gaussian_convolve3_constant_field(oracle):
    vol = np.full((20, 21, 22), 3.25, np.float32)
    out = oracle.gaussian_convolve3(vol, 1.5, 1.0, 0.8)
    np.testing.assert_allclose(out, 3.25, rtol=2e-6)


def test_same_packed6_vs_scipy(oracle):
    rng = np.random.default_rng(1)
    t6 = rng.normal(size=(9, 10, 11, 6)).astype(np.float32)
    g = oracle.gaussian1d(11, 0.0, 1.5, 1.0)
    out = oracle.gaussian3_same_packed6_f64(t6, g)
    ref = t6.astype(np.float64)
    for ax in (2, 1, 0):
        ref = ndimage.correlate1d(ref, g.astype(np.float64), axis=ax, mode="constant")
    np.testing.assert_allclose(out, ref, atol=1e-12)
