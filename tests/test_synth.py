"""Seeded synthetic workloads are deterministic and have the stated content."""
import numpy as np

from tiralib_amd import synth


def test_bezier_field_deterministic_and_unit_sticks():
    a = synth.bezier_field(64, 64, n_curves=3, seed=1)
    b = synth.bezier_field(64, 64, n_curves=3, seed=1)
    np.testing.assert_array_equal(a, b)
    tr = a[..., 0] + a[..., 3]
    on = tr > 0.5
    assert on.sum() > 100
    np.testing.assert_allclose(tr[on], 1.0, atol=1e-5)          # curve pixels: unit sticks
    np.testing.assert_allclose(a[..., 1], a[..., 2])             # symmetric [a b b c]
    det = a[..., 0] * a[..., 3] - a[..., 1] ** 2
    assert np.abs(det).max() < 1e-5                              # rank one everywhere


def test_tubes_sheets_volume_content():
    v = synth.tubes_sheets_volume((40, 40, 40), n_straight=2, n_helix=2, n_sheets=2, seed=3)
    np.testing.assert_array_equal(v, synth.tubes_sheets_volume((40, 40, 40), n_straight=2, n_helix=2, n_sheets=2, seed=3))
    M = v.reshape(-1, 3, 3).astype(np.float64)
    np.testing.assert_allclose(M, np.transpose(M, (0, 2, 1)))
    tr = np.trace(M, axis1=1, axis2=2)
    ev = np.linalg.eigvalsh(M)
    stick = np.isclose(tr, 1.0, atol=1e-5) & np.isclose(ev[:, 1], 0, atol=1e-5)
    plate = np.isclose(tr, 2.0, atol=1e-5) & np.isclose(ev[:, 0], 0, atol=1e-5)
    iso = (tr > 0) & np.isclose(ev[:, 0], ev[:, 2], atol=1e-6)
    assert stick.sum() > 0 and plate.sum() > 0 and iso.sum() > 0
    assert np.all(stick | plate | iso | (tr == 0))


def test_eig_tensors_spectrum():
    t6, lam = synth.eig_tensors(20000, seed=4)
    A = synth.mat9_from_packed6(t6).reshape(-1, 3, 3).astype(np.float64)
    ev = np.linalg.eigvalsh(A)
    assert np.all(np.abs(ev - lam) <= 1e-5 * lam.max(1, keepdims=True))
    gap = np.minimum(lam[:, 1] - lam[:, 0], lam[:, 2] - lam[:, 1]) / lam[:, 2]
    frac = np.mean(gap < 1e-4)
    assert 0.08 < frac < 0.2
    assert lam.min() >= 1e-3 * 0.999 and lam.max() <= 1e2 * 1.001
    np.testing.assert_array_equal(synth.packed6_from_mat9(synth.mat9_from_packed6(t6)), t6)
