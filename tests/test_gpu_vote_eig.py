"""tira_tensorvote3_eig_async against tira_tensorvote3_async followed by
tira_eig3_async on its output, and against the FP64 oracle's eigenvalues of
the voted tensors (<= 1e-6 normwise with the FP64 re-solve and
scale-invariant thresholds): the same field
bit for bit, and eigenvalues of the same solver to FP32 rounding (the solver inlined into another kernel may
contract a few operations differently: 0.6% of values differ, by one or two
ulps), with the FP64 re-solve taken for the same tensors up to the borderline
ones, for both output layouts and the eigen flags the pipelines use. Sizes from 32^3 voxels take the spectral vote,
whose last pass writes the eigenvalues; the small volume takes the gather and
a separate eig3."""
import numpy as np
import pytest

from tiralib_amd import _lib
from tiralib_amd._lib import DeviceArray, call

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    """Sticks of random direction and strength, an isotropic block and an
    empty slab."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    q = rng.normal(size=(n, 3))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = np.einsum("ni,nj->nij", q, q) * rng.uniform(0.2, 1.0, (n, 1, 1))
    t = t.reshape(*shape, 3, 3)
    s0 = shape[0]
    t[: s0 // 4] = np.eye(3)
    t[s0 // 2: s0 // 2 + 3] = 0.0
    return np.ascontiguousarray(t.reshape(n, 9), np.float32)


@pytest.mark.parametrize("shape,sigma,w", [((40, 36, 48), 2.0, 13), ((12, 10, 14), 1.5, 7)])
@pytest.mark.parametrize("out_layout", [_lib.LAYOUT_PACKED6, _lib.LAYOUT_MAT9])
@pytest.mark.parametrize("flags", [_lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT, 0, _lib.EIG_REFINE_FP64,
                                   _lib.EIG_REFINE_FP64 | _lib.EIG_SALIENCY | _lib.EIG_MAGNITUDE])
def test_vote_eig_equals_vote_then_eig3(hip, oracle, shape, sigma, w, out_layout, flags):
    n = int(np.prod(shape))
    print(f"vote+eig shape={shape} out_layout={out_layout} flags={flags}", flush=True)
    comps = 9 if out_layout == _lib.LAYOUT_MAT9 else 6
    d_in = DeviceArray.from_host(_field(shape, 7))
    bufs = [DeviceArray((n, comps)), DeviceArray((n, 3)), DeviceArray((n, comps)), DeviceArray((n, 3)),
            DeviceArray((2,), np.uint64)]
    f_ref, e_ref, f_fus, e_fus, cnt = bufs
    try:
        cnt.zero()
        args = (d_in.ptr, _lib.LAYOUT_MAT9)
        tail = (*shape, sigma, 0.0, w, 1, 1, 1, 20, _lib.PLATE_AUTO)
        call("tira_tensorvote3_async", *args, f_ref.ptr, out_layout, *tail)
        call("tira_eig3_async", f_ref.ptr, n, out_layout, e_ref.ptr, None, None, flags, cnt.ptr)
        call("tira_tensorvote3_eig_async", *args, f_fus.ptr, out_layout, *tail, e_fus.ptr, flags, cnt.ptr + 8)
        call("tira_hip_synchronize")
        fr, er, ff, ef, c = (b.to_host() for b in bufs)
    finally:
        d_in.free()
        for b in bufs:
            b.free()
    np.testing.assert_array_equal(ff, fr)
    # FP32 eigen-solver differences scale with the tensor (saliency values are
    # differences of eigenvalues and can be far smaller than the tensor)
    scale = float(np.abs(fr).max())
    assert np.abs(ef - er).max() <= 4e-6 * scale, np.abs(ef - er).max() / scale
    assert abs(int(c[0]) - int(c[1])) <= max(2, int(c[0]) // 100)
    assert np.isfinite(ef).all()
    if flags == _lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT:
        # against the FP64 oracle: eigenvalues of the same voted tensors in
        # double (with the reference's absolute thresholds, flags without
        # SCALE_INVARIANT, tiny tensors may take its diagonal shortcut)
        t6 = ff if out_layout == _lib.LAYOUT_PACKED6 else ff[:, [0, 1, 2, 4, 5, 8]]
        own = oracle.eig3_packed6_f64(t6)
        scale = np.maximum(np.abs(own).max(axis=1), 1e-30)
        live = np.abs(t6).max(axis=1) > 0
        err = (np.abs(ef.astype(np.float64) - own).max(axis=1) / scale)[live]
        print(f"fused eigenvalues vs FP64: normwise max {err.max():.2e}")
        assert err.max() <= 1e-6
