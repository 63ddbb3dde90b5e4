"""Python mirror of TIRA's GPU entry points for the tensor-voting path.

Names, argument order and meaning follow the reference's tira::cuda functions
(eigen.h:768-1012, tensorvote.h:2816-2873, filter.h:400-534); every call goes
through libtira_hip.so. Inputs may be numpy arrays (host; results come back
as numpy arrays) or DeviceArray objects (device; results stay on the device).
Argument checks raise ValueError like the reference's pybind layer
(tensorvote_tk.cu:73-83, 162-178).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, call, ptr


def _f32(a):
    if isinstance(a, DeviceArray):
        if a.dtype != np.float32:
            raise ValueError("device arrays must be float32")
        return a
    return np.ascontiguousarray(a, dtype=np.float32)


def _like(src, shape):
    if isinstance(src, DeviceArray):
        return DeviceArray(shape, np.float32)
    return np.empty(shape, np.float32)


def _count(a, per):
    n = int(np.prod(a.shape))
    if n % per:
        raise ValueError(f"array of {n} floats is not a whole number of {per}-float tensors")
    return n // per


# ------------------------------------------------------------------ eigen
def evals2_symmetric(mats, device: int = 0):
    """tira::cuda::evals2_symmetric: n x [a b b c] -> n x 2 ascending eigenvalues."""
    m = _f32(mats)
    n = _count(m, 4)
    out = _like(m, (n, 2))
    call("tira_evals2_symmetric", ptr(m), n, ptr(out), int(device))
    return out


def evecs2polar_symmetric(mats, evals, device: int = 0):
    """tira::cuda::evecs2polar_symmetric: eigenvector angles (theta0, theta1)."""
    m, e = _f32(mats), _f32(evals)
    n = _count(m, 4)
    out = _like(m, (n, 2))
    call("tira_evecs2polar_symmetric", ptr(m), ptr(e), n, ptr(out), int(device))
    return out


def evals3_symmetric(mats, device: int = 0):
    """tira::cuda::evals3_symmetric: n x 9 (column-major) -> n x 3 ascending."""
    m = _f32(mats)
    n = _count(m, 9)
    out = _like(m, (n, 3))
    call("tira_evals3_symmetric", ptr(m), n, ptr(out), int(device))
    return out


def evecs3spherical_symmetric(mats, lambdas, device: int = 0):
    """tira::cuda::evecs3spherical_symmetric: n x 6 = (theta, phi) of e0, e1, e2."""
    m, l = _f32(mats), _f32(lambdas)
    n = _count(m, 9)
    out = _like(m, (n, 6))
    call("tira_evecs3spherical_symmetric", ptr(m), ptr(l), n, ptr(out), int(device))
    return out


def eig3(tensors, layout=_lib.LAYOUT_PACKED6, vectors=False,
         flags=_lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT):
    """Native batched eig3 on device arrays: returns (evals, evec0, evec2)."""
    if not isinstance(tensors, DeviceArray):
        raise ValueError("eig3 works on DeviceArray inputs")
    n = _count(tensors, 9 if layout == _lib.LAYOUT_MAT9 else 6)
    ev = DeviceArray((n, 3))
    e0 = DeviceArray((n, 3)) if vectors else None
    e2 = DeviceArray((n, 3)) if vectors else None
    call("tira_eig3_async", tensors.ptr, n, layout, ev.ptr, e0.ptr if e0 else None, e2.ptr if e2 else None,
         int(flags), None)
    call("tira_hip_synchronize")
    return ev, e0, e2


# ------------------------------------------------------------------ 3D voting
def _vol_shape(field):
    sh = tuple(field.shape)
    if len(sh) == 5 and sh[3:] == (3, 3):
        return sh[:3]
    if len(sh) == 4 and sh[3] == 9:
        return sh[:3]
    raise ValueError("3D field must have shape (s0, s1, s2, 9) or (s0, s1, s2, 3, 3)")


def tensorvote3(input_field, sigma: float, sigma2: float = 0.0, w: int = 0, power: int = 1, device: int = 0,
                stick: bool = True, plate: bool = True, samples: int = 20, is_atv: bool = False):
    """tira::cuda::tensorvote3 (tensorvote.h:2816-2873) on a (s0, s1, s2, 9) field.

    w = 0 picks the window like the reference's Python layer: 2*floor(3*max(sigma, sigma2)) + 1.
    """
    f = _f32(input_field)
    s0, s1, s2 = _vol_shape(f)
    if sigma <= 0:
        raise ValueError("sigma must be > 0")
    if w <= 0:
        w = 2 * int(3.0 * max(sigma, sigma2)) + 1
    out = _like(f, (s0, s1, s2, 9))
    call("tira_tensorvote3", ptr(f), ptr(out), s0, s1, s2, float(sigma), float(sigma2), int(w), int(power),
         int(device), int(stick), int(plate), int(samples), int(is_atv))
    return out


def tensorvote3_tk(lambdas, evec2, evec0, shape, sigma1, sigma2, power, w, stick=True, plate=True, samples=20):
    """tira::cuda::_tensorvote3_tk (tensorvote.h:2741-2809) from precomputed eigenpairs."""
    s0, s1, s2 = shape
    l = _f32(lambdas)
    e2 = _f32(evec2) if evec2 is not None else None
    e0 = _f32(evec0) if evec0 is not None else None
    out = _like(l, (s0, s1, s2, 9))
    call("tira_tensorvote3_tk", ptr(out), ptr(l), ptr(e2) if e2 is not None else None,
         ptr(e0) if e0 is not None else None, s0, s1, s2, float(sigma1), float(sigma2), int(power), int(w),
         int(stick), int(plate), int(samples))
    return out


def tensorvote3_async(field, out, shape, sigma1, sigma2, w, power, stick=True, plate=True, samples=20,
                      plate_mode=_lib.PLATE_AUTO, in_layout=_lib.LAYOUT_MAT9, out_layout=_lib.LAYOUT_MAT9):
    """Native non-blocking 3D vote on device arrays (library stream)."""
    s0, s1, s2 = shape
    call("tira_tensorvote3_async", field.ptr, int(in_layout), out.ptr, int(out_layout), s0, s1, s2,
         float(sigma1), float(sigma2), int(w), int(power), int(stick), int(plate), int(samples), int(plate_mode))
    return out


# ------------------------------------------------------------------ 2D voting
NEIGHBOR2D_DTYPE = np.dtype([
    ("du", np.int32), ("dv", np.int32),
    ("dx", np.float32), ("dy", np.float32), ("c1", np.float32), ("c2", np.float32),
    ("Ra", np.float32), ("Rb", np.float32), ("Rc", np.float32),
    ("Va", np.float32), ("Vb", np.float32), ("Vc", np.float32),
])


def build_neighbors2d(w, sigma1, sigma2, power, samples, stick=True, plate=True):
    """tira::cpu::build_neighbors2d (tensorvote.h:1223-1280): structured array of
    48-byte Neighbor2D records (dv outer, du inner)."""
    cnt = C.c_int()
    call("tira_build_neighbors2d", int(w), float(sigma1), float(sigma2), int(power), int(samples), int(stick),
         int(plate), None, 0, C.byref(cnt))
    nb = np.zeros(cnt.value, NEIGHBOR2D_DTYPE)
    call("tira_build_neighbors2d", int(w), float(sigma1), float(sigma2), int(power), int(samples), int(stick),
         int(plate), nb.ctypes.data if cnt.value else None, cnt.value, C.byref(cnt))
    return nb


def _field2_shape(field):
    if field.ndim != 3 or field.shape[2] != 4:
        raise ValueError("2D field must have shape (H, W, 4)")
    return field.shape[0], field.shape[1]


def tensorvote(t_in, sigma1, sigma2=0.0, power=1, w=0, stick=True, plate=True, samples=10, is_atv=False):
    """tira::cuda::tensorvote (2D, tensorvote.h:2200-2226) on an (H, W, 4) field."""
    f = _f32(t_in)
    H, W = _field2_shape(f)
    if w <= 0:
        w = 2 * int(3.0 * max(sigma1, sigma2)) + 1
    out = _like(f, (H, W, 4))
    call("tira_tensorvote2", ptr(out), ptr(f), float(sigma1), float(sigma2), int(power), int(w), H, W, int(stick),
         int(plate), int(samples), int(is_atv))
    return out


def tensorvote_prebuilt(t_in, nb, sigma1, sigma2, power, w, stick=True, plate=True, samples=10):
    """tira::cuda::tensorvote_prebuilt (tensorvote.h:2236-2305) with a host Neighbor2D table."""
    f = _f32(t_in)
    H, W = _field2_shape(f)
    nb = np.ascontiguousarray(nb)
    out = _like(f, (H, W, 4))
    call("tira_tensorvote2_prebuilt", ptr(out), ptr(f), nb.ctypes.data if nb.size else None, int(nb.size),
         float(sigma1), float(sigma2), int(power), int(w), H, W, int(stick), int(plate), int(samples))
    return out


def tensorvote_tk(lambdas, thetas, shape, sigma1, sigma2, power, w, stick=True, plate=True, samples=10):
    """tira::cuda::tensorvote_tk (2D, tensorvote.h:2125-2179) from eigenpairs (H, W, 2) each."""
    H, W = shape
    l, t = _f32(lambdas), _f32(thetas)
    out = _like(l, (H, W, 4))
    call("tira_tensorvote2_tk", ptr(out), ptr(l), ptr(t), float(sigma1), float(sigma2), int(power), int(w), H, W,
         int(stick), int(plate), int(samples))
    return out
