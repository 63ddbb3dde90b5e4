"""Host-side 2D logic of the library (no GPU): the Neighbor2D table builder
against the oracle, and the reference module's argument validation."""
import numpy as np
import pytest


@pytest.mark.parametrize("w,s1,s2,p,n", [(19, 3.0, 0.0, 1, 20), (11, 2.0, 1.0, 3, 10), (9, 1.5, 0.5, 2, 0)])
def test_build_neighbors2d_matches_oracle(oracle, w, s1, s2, p, n):
    from tiralib_amd import api
    a = api.build_neighbors2d(w, s1, s2, p, n)
    b = oracle.build_neighbors2d(w, s1, s2, p, n)
    assert a.dtype.itemsize == 48 and len(a) == w * w
    for f in ("du", "dv"):
        np.testing.assert_array_equal(a[f], b[f])
    for f in ("dx", "dy", "c1", "c2", "Ra", "Rb", "Rc", "Va", "Vb", "Vc"):
        np.testing.assert_allclose(a[f], b[f], rtol=1e-6, atol=1e-9)


def test_tensorvote_tk_module_validation():
    from tiralib_amd import tensorvote_tk as tvk
    with pytest.raises(RuntimeError, match=r"shape \(H, W, 4\)"):
        tvk.tensorvote_tk(np.zeros((4, 4), np.float32), 1.0)
    with pytest.raises(RuntimeError, match="sigma1 must be > 0"):
        tvk.tensorvote_tk(np.zeros((4, 4, 4), np.float32), 0.0)
    with pytest.raises(RuntimeError, match="positive odd"):
        tvk.tensorvote_tk_prebuilt(np.zeros((4, 4, 4), np.float32), b"", 0, 4, 1.0)
    with pytest.raises(RuntimeError, match="48"):
        tvk.tensorvote_tk_prebuilt(np.zeros((4, 4, 4), np.float32), b"\0" * 47, 1, 3, 1.0)
