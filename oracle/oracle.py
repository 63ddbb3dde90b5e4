"""ctypes binding of the CPU oracle (oracle/libtira_oracle.so) and of the
reference's own eigen.h / filter.h built from source (oracle/_ref/libtira_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "libtira_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libtira_ref.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_sz = C.c_size_t
_u = C.c_uint
_i = C.c_int
_f = C.c_float
_d = C.c_double

NEIGHBOR2D_DTYPE = np.dtype([
    ("du", np.int32), ("dv", np.int32),
    ("dx", np.float32), ("dy", np.float32), ("c1", np.float32), ("c2", np.float32),
    ("Ra", np.float32), ("Rb", np.float32), ("Rc", np.float32),
    ("Va", np.float32), ("Vb", np.float32), ("Vc", np.float32),
])
NEIGHBOR3D_DTYPE = np.dtype([
    ("du", np.int32), ("dv", np.int32), ("dw", np.int32),
    ("d", np.float32, (3,)), ("c1", np.float32), ("c2", np.float32),
])
assert NEIGHBOR2D_DTYPE.itemsize == 48 and NEIGHBOR3D_DTYPE.itemsize == 32

_lib = None
_ref = None


def _sig(lib, name, args, res=None):
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = res
    return fn


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(ORACLE_SO)
        _sig(L, "oracle_evals2_symmetric", [_f32p, _sz, _f32p])
        _sig(L, "oracle_evecs2polar_symmetric", [_f32p, _f32p, _sz, _f32p])
        _sig(L, "oracle_evals3_symmetric", [_f32p, _sz, _f32p])
        _sig(L, "oracle_evecs3spherical_symmetric", [_f32p, _f32p, _sz, _f32p])
        _sig(L, "oracle_spherical_to_cartesian", [_f, _f, _f32p])
        _sig(L, "oracle_eig3_jacobi_f64", [_f64p, _f64p, _f64p])
        _sig(L, "oracle_eig3_f64", [_f32p, _sz, _f64p, _f64p, _f64p])
        _sig(L, "oracle_eig3_packed6_f64", [_f32p, _sz, _f64p])
        _sig(L, "oracle_build_neighbors2d", [_i, _f, _f, _u, _u, _i, _i, C.c_void_p], _i)
        _sig(L, "oracle_tk_stick_norm_2d", [_f, _f, _u], _f)
        _sig(L, "oracle_tk_stickvote_2d", [_f, _f, _f, _f, _u, _f, C.POINTER(_f), C.POINTER(_f), C.POINTER(_f)])
        _sig(L, "oracle_tensorvote2", [_f32p, _f32p, _f, _f, _i, _u, _u, _u, _i, _i, _u, _i])
        _sig(L, "oracle_tensorvote2_tk", [_f32p, _f32p, _f32p, _f, _f, _u, _u, _u, _u, _i, _i, _u,
                                          C.c_void_p, _i, _i])
        _sig(L, "oracle_tensorvote2_f64", [_f64p, _f32p, _d, _d, _i, _u, _u, _u, _i, _i, _u, _i])
        _sig(L, "oracle_tk_plate_numerical_2d", [_f, _f, _f, _f, _u, C.POINTER(_f), C.POINTER(_f),
                                                 C.POINTER(_f), _u])
        _sig(L, "oracle_build_neighbors3d", [_i, _f, _f, C.c_void_p], _i)
        _sig(L, "oracle_sticknorm3", [_f, _f, _u], _f)
        _sig(L, "oracle_tensorvote3_f32", [_f32p, _f32p, _u, _u, _u, _f, _f, _u, _u, _i, _i, _u,
                                           C.c_void_p, _sz, _i])
        _sig(L, "oracle_tensorvote3_tk_f32", [_f32p, _f32p, _f32p, _f32p, _u, _u, _u, _f, _f, _u, _u,
                                              _i, _i, _u, C.c_void_p, _sz, _i])
        _sig(L, "oracle_tensorvote3_f64", [_f32p, _f64p, _u, _u, _u, _d, _d, _u, _u, _i, _i, _u,
                                           C.c_void_p, _sz, _i])
        _sig(L, "oracle_stick3_pair_f64", [_f64p, _d, _d, _u, _f64p, _f64p])
        _sig(L, "oracle_plate3_sum_f64", [_f64p, _d, _d, _u, _f64p, _u, _f64p])
        _sig(L, "oracle_plate3_analytic_f64", [_f64p, _d, _d, _u, _f64p, _f64p])
        _sig(L, "oracle_plate3_numerical_f32", [_f32p, _f, _f, _u, _f32p, _u, _f32p])
        _sig(L, "oracle_gaussian_ksize", [_f], _u)
        _sig(L, "oracle_gaussian1d", [_u, _f, _f, _f, _f32p])
        _sig(L, "oracle_convolve3", [_f32p, _u, _u, _u, _f32p, _u, _u, _u, _f32p,
                                     C.POINTER(_u), C.POINTER(_u), C.POINTER(_u)])
        _sig(L, "oracle_gaussian_convolve3", [_f32p, _u, _u, _u, _f, _f, _f, _f32p,
                                              C.POINTER(_u), C.POINTER(_u), C.POINTER(_u)])
        _sig(L, "oracle_gaussian3_same_packed6_f64", [_f32p, _u, _u, _u, _f32p, _u, _f64p, _i])
        _sig(L, "oracle_atv2_f32", [_f] * 9 + [_f32p])
        _sig(L, "oracle_atv2_f64", [_d] * 9 + [_f64p])
        _sig(L, "oracle_atv3_f32", [_f, _f, _f, _f32p, _f32p, _f, _f, _f, _f, _f32p])
        _sig(L, "oracle_atv3_f64", [_d, _d, _d, _f64p, _f64p, _d, _d, _d, _d, _f64p])
        _sig(L, "oracle_tensorvote2_atv_f32", [_f32p, _f32p, _f32p, _f, _u, _u, _u, _i])
        _sig(L, "oracle_tensorvote2_atv_full_f32", [_f32p, _f32p, _f, _u, _u, _u, _i])
        _sig(L, "oracle_tensorvote2_atv_f64", [_f64p, _f32p, _d, _u, _u, _u, _i])
        _sig(L, "oracle_tensorvote3_atv_eig_f32", [_f32p, _f32p, _f32p, _f32p, _u, _u, _u, _f, _u,
                                                   C.c_void_p, _sz, _i])
        _sig(L, "oracle_tensorvote3_atv_f32", [_f32p, _f32p, _u, _u, _u, _f, _u, C.c_void_p, _sz, _i])
        _sig(L, "oracle_tensorvote3_atv_f64", [_f32p, _f64p, _u, _u, _u, _d, _u, C.c_void_p, _sz, _i])
        _sig(L, "oracle_otsu_threshold_f32", [_f32p, _sz], _f)
        _sig(L, "oracle_thin_euler_lut_sum", [_u], _i)
        _sig(L, "oracle_thin_delta_euler", [_u], _i)
        _sig(L, "oracle_thin_simple_octree", [_u], _i)
        _sig(L, "oracle_thin_simple_bfs", [_u], _i)
        _sig(L, "oracle_thin_exhaustive", [C.POINTER(C.c_ulonglong), _i])
        _sig(L, "oracle_lee_thin3", [C.c_void_p, _u, _u, _u], _i)
        _sig(L, "oracle_thin_classify", [_u, _u, C.c_void_p, _i])
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's eigen.h / filter.h compiled from source, or None."""
    global _ref
    if _ref is None and ref_available():
        R = C.CDLL(REF_SO)
        _sig(R, "ref_evals2_symmetric", [_f32p, _sz, _f32p])
        _sig(R, "ref_evecs2polar_symmetric", [_f32p, _f32p, _sz, _f32p])
        _sig(R, "ref_evals3_symmetric", [_f32p, _sz, _f32p])
        _sig(R, "ref_evecs3spherical_symmetric", [_f32p, _f32p, _sz, _f32p])
        _sig(R, "ref_gaussian1d", [_u, _f, _f, _f, _f32p])
        _sig(R, "ref_convolve3", [_f32p, _u, _u, _u, _f32p, _u, _u, _u, _f32p,
                                  C.POINTER(_u), C.POINTER(_u), C.POINTER(_u)])
        _sig(R, "ref_gaussian_convolve3", [_f32p, _u, _u, _u, _f, _f, _f, _f32p,
                                           C.POINTER(_u), C.POINTER(_u), C.POINTER(_u)])
        _ref = R
    return _ref


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- eigen
def evals2(mats4, impl="oracle"):
    m = _c32(mats4).reshape(-1, 4)
    out = np.empty((m.shape[0], 2), np.float32)
    L = lib() if impl == "oracle" else ref()
    getattr(L, ("oracle_" if impl == "oracle" else "ref_") + "evals2_symmetric")(m, m.shape[0], out)
    return out


def evecs2polar(mats4, evals, impl="oracle"):
    m = _c32(mats4).reshape(-1, 4)
    e = _c32(evals).reshape(-1, 2)
    out = np.empty((m.shape[0], 2), np.float32)
    L = lib() if impl == "oracle" else ref()
    getattr(L, ("oracle_" if impl == "oracle" else "ref_") + "evecs2polar_symmetric")(m, e, m.shape[0], out)
    return out


def evals3(mats9, impl="oracle"):
    m = _c32(mats9).reshape(-1, 9)
    out = np.empty((m.shape[0], 3), np.float32)
    L = lib() if impl == "oracle" else ref()
    getattr(L, ("oracle_" if impl == "oracle" else "ref_") + "evals3_symmetric")(m, m.shape[0], out)
    return out


def evecs3spherical(mats9, evals, impl="oracle"):
    m = _c32(mats9).reshape(-1, 9)
    e = _c32(evals).reshape(-1, 3)
    out = np.empty((m.shape[0], 6), np.float32)
    L = lib() if impl == "oracle" else ref()
    getattr(L, ("oracle_" if impl == "oracle" else "ref_") + "evecs3spherical_symmetric")(m, e, m.shape[0], out)
    return out


def eig3_f64(mats9):
    m = _c32(mats9).reshape(-1, 9)
    n = m.shape[0]
    ev = np.empty((n, 3)); e0 = np.empty((n, 3)); e2 = np.empty((n, 3))
    lib().oracle_eig3_f64(m, n, ev, e0, e2)
    return ev, e0, e2


def eig3_packed6_f64(t6):
    t = _c32(t6).reshape(-1, 6)
    ev = np.empty((t.shape[0], 3))
    lib().oracle_eig3_packed6_f64(t, t.shape[0], ev)
    return ev


def jacobi_f64(a6):
    a = np.ascontiguousarray(a6, np.float64)
    w = np.empty(3); V = np.empty(9)
    lib().oracle_eig3_jacobi_f64(a, w, V)
    return w, V.reshape(3, 3)


def spherical_to_cartesian(theta, phi):
    out = np.empty(3, np.float32)
    lib().oracle_spherical_to_cartesian(float(theta), float(phi), out)
    return out


# ---------------------------------------------------------------- 2D voting
def build_neighbors2d(w, sigma1, sigma2, power, samples, stick=True, plate=True):
    nb = np.zeros(w * w, NEIGHBOR2D_DTYPE)
    k = lib().oracle_build_neighbors2d(int(w), float(sigma1), float(sigma2), int(power), int(samples),
                                       int(stick), int(plate), nb.ctypes.data)
    return nb[:k]


def tensorvote2(field, sigma1, sigma2, power, w, stick=True, plate=True, samples=10, nthreads=1):
    f = _c32(field)
    H, W = f.shape[0], f.shape[1]
    out = np.zeros((H, W, 4), np.float32)
    lib().oracle_tensorvote2(out, f.reshape(H, W, 4), float(sigma1), float(sigma2), int(power), int(w), H, W,
                             int(stick), int(plate), int(samples), int(nthreads))
    return out


def tensorvote2_f64(field, sigma1, sigma2, power, w, stick=True, plate=True, samples=10, nthreads=1):
    f = _c32(field)
    H, W = f.shape[0], f.shape[1]
    out = np.zeros((H, W, 4), np.float64)
    lib().oracle_tensorvote2_f64(out, f.reshape(H, W, 4), float(sigma1), float(sigma2), int(power), int(w), H, W,
                                 int(stick), int(plate), int(samples), int(nthreads))
    return out


def tk_stickvote_2d(rx, ry, sigma1, sigma2, power, theta):
    a, b, c = _f(), _f(), _f()
    lib().oracle_tk_stickvote_2d(float(rx), float(ry), float(sigma1), float(sigma2), int(power), float(theta),
                                 C.byref(a), C.byref(b), C.byref(c))
    return np.array([a.value, b.value, c.value], np.float32)


def tk_plate_numerical_2d(rx, ry, sigma1, sigma2, power, n):
    a, b, c = _f(), _f(), _f()
    lib().oracle_tk_plate_numerical_2d(float(rx), float(ry), float(sigma1), float(sigma2), int(power),
                                       C.byref(a), C.byref(b), C.byref(c), int(n))
    return np.array([a.value, b.value, c.value], np.float32)


# ---------------------------------------------------------------- 3D voting
def build_neighbors3d(w, sigma1, sigma2):
    nb = np.zeros(w * w * w, NEIGHBOR3D_DTYPE)
    k = lib().oracle_build_neighbors3d(int(w), float(sigma1), float(sigma2), nb.ctypes.data)
    return nb[:k]


def sticknorm3(sigma1, sigma2, p):
    return lib().oracle_sticknorm3(float(sigma1), float(sigma2), int(p))


def _recv(recv):
    if recv is None:
        return None, 0, None
    r = np.ascontiguousarray(recv, np.int64)
    return r.ctypes.data, r.shape[0], r


def tensorvote3(vol9, sigma1, sigma2, power, w, stick=True, plate=True, samples=20, recv=None, nthreads=1):
    """Reference-faithful FP32 3D vote (cpu::tensorvote3). vol9: (s0, s1, s2, 9)."""
    v = _c32(vol9)
    s0, s1, s2 = v.shape[:3]
    rp, nr, keep = _recv(recv)
    out = np.zeros(((nr if recv is not None else s0 * s1 * s2), 9), np.float32)
    lib().oracle_tensorvote3_f32(v.reshape(-1), out, s0, s1, s2, float(sigma1), float(sigma2), int(power), int(w),
                                 int(stick), int(plate), int(samples), rp, nr, int(nthreads))
    return out if recv is not None else out.reshape(s0, s1, s2, 9)


def tensorvote3_tk(lambdas, evec2, evec0, shape, sigma1, sigma2, power, w, stick=True, plate=True, samples=20,
                   recv=None, nthreads=1):
    s0, s1, s2 = shape
    rp, nr, keep = _recv(recv)
    out = np.zeros(((nr if recv is not None else s0 * s1 * s2), 9), np.float32)
    lib().oracle_tensorvote3_tk_f32(_c32(lambdas).reshape(-1), _c32(evec2).reshape(-1), _c32(evec0).reshape(-1),
                                    out, s0, s1, s2, float(sigma1), float(sigma2), int(power), int(w), int(stick),
                                    int(plate), int(samples), rp, nr, int(nthreads))
    return out if recv is not None else out.reshape(s0, s1, s2, 9)


def tensorvote3_f64(vol9, sigma1, sigma2, power, w, stick=True, plate=True, samples=20, recv=None, nthreads=1):
    v = _c32(vol9)
    s0, s1, s2 = v.shape[:3]
    rp, nr, keep = _recv(recv)
    out = np.zeros(((nr if recv is not None else s0 * s1 * s2), 9), np.float64)
    lib().oracle_tensorvote3_f64(v.reshape(-1), out, s0, s1, s2, float(sigma1), float(sigma2), int(power), int(w),
                                 int(stick), int(plate), int(samples), rp, nr, int(nthreads))
    return out if recv is not None else out.reshape(s0, s1, s2, 9)


def stick3_pair_f64(d, c1, c2, power, q):
    o = np.empty(6)
    lib().oracle_stick3_pair_f64(np.asarray(d, np.float64), float(c1), float(c2), int(power),
                                 np.asarray(q, np.float64), o)
    return o


def plate3_sum_f64(d, c1, c2, power, n, samples):
    o = np.empty(6)
    lib().oracle_plate3_sum_f64(np.asarray(d, np.float64), float(c1), float(c2), int(power),
                                np.asarray(n, np.float64), int(samples), o)
    return o


def plate3_analytic_f64(d, c1, c2, power, n):
    o = np.empty(6)
    lib().oracle_plate3_analytic_f64(np.asarray(d, np.float64), float(c1), float(c2), int(power),
                                     np.asarray(n, np.float64), o)
    return o


def plate3_numerical_f32(d, c1, c2, power, n, samples):
    o = np.empty(6, np.float32)
    lib().oracle_plate3_numerical_f32(_c32(d), float(c1), float(c2), int(power), _c32(n), int(samples), o)
    return o


# ---------------------------------------------------------------- analytical tensor voting
def atv2_pair(l0, l1, ev0, ev1, r, sigma, f64=False):
    """One 2D ATV vote (tensorvote.h:224-253): (a, b, c) at offset r = (rx, ry)."""
    if f64:
        o = np.empty(3)
        lib().oracle_atv2_f64(float(l0), float(l1), float(ev0[0]), float(ev0[1]), float(ev1[0]), float(ev1[1]),
                              float(r[0]), float(r[1]), float(sigma), o)
    else:
        o = np.empty(3, np.float32)
        lib().oracle_atv2_f32(float(l0), float(l1), float(ev0[0]), float(ev0[1]), float(ev1[0]), float(ev1[1]),
                              float(r[0]), float(r[1]), float(sigma), o)
    return o


def atv3_pair(lam, e1, e2, r, sigma, f64=False):
    """One 3D ATV vote (tensorvote.h:688-760): (00, 01, 11, 02, 12, 22)."""
    if f64:
        o = np.empty(6)
        lib().oracle_atv3_f64(float(lam[0]), float(lam[1]), float(lam[2]), np.ascontiguousarray(e1, np.float64),
                              np.ascontiguousarray(e2, np.float64), float(r[0]), float(r[1]), float(r[2]), float(sigma), o)
    else:
        o = np.empty(6, np.float32)
        lib().oracle_atv3_f32(float(lam[0]), float(lam[1]), float(lam[2]), _c32(e1), _c32(e2), float(r[0]),
                              float(r[1]), float(r[2]), float(sigma), o)
    return o


def tensorvote2_atv(field, sigma, w, nthreads=1):
    """cpu::tensorvote with is_atv (FP32, reference-faithful). field: (H, W, 4)."""
    f = _c32(field)
    H, W = f.shape[0], f.shape[1]
    out = np.zeros((H, W, 4), np.float32)
    lib().oracle_tensorvote2_atv_full_f32(out, f.reshape(H, W, 4), float(sigma), int(w), H, W, int(nthreads))
    return out


def tensorvote2_atv_eig(lambdas, thetas, sigma, w, nthreads=1):
    """ATV window from eigenpairs (lambdas, thetas: (H, W, 2))."""
    lam = _c32(lambdas)
    H, W = lam.shape[0], lam.shape[1]
    out = np.zeros((H, W, 4), np.float32)
    lib().oracle_tensorvote2_atv_f32(out, lam.reshape(-1), _c32(thetas).reshape(-1), float(sigma), int(w), H, W,
                                     int(nthreads))
    return out


def tensorvote2_atv_f64(field, sigma, w, nthreads=1):
    f = _c32(field)
    H, W = f.shape[0], f.shape[1]
    out = np.zeros((H, W, 4), np.float64)
    lib().oracle_tensorvote2_atv_f64(out, f.reshape(H, W, 4), float(sigma), int(w), H, W, int(nthreads))
    return out


def tensorvote3_atv(vol9, sigma, w, recv=None, nthreads=1):
    """cpu::tensorvote3 with is_atv (FP32, reference-faithful). vol9: (s0, s1, s2, 9)."""
    v = _c32(vol9)
    s0, s1, s2 = v.shape[:3]
    rp, nr, keep = _recv(recv)
    out = np.zeros(((nr if recv is not None else s0 * s1 * s2), 9), np.float32)
    lib().oracle_tensorvote3_atv_f32(v.reshape(-1), out, s0, s1, s2, float(sigma), int(w), rp, nr, int(nthreads))
    return out if recv is not None else out.reshape(s0, s1, s2, 9)


def tensorvote3_atv_eig(lambdas, evec1, evec2, shape, sigma, w, recv=None, nthreads=1):
    """ATV window from eigenpairs (the inputs of _tensorvote3_atv)."""
    s0, s1, s2 = shape
    rp, nr, keep = _recv(recv)
    out = np.zeros(((nr if recv is not None else s0 * s1 * s2), 9), np.float32)
    lib().oracle_tensorvote3_atv_eig_f32(_c32(lambdas).reshape(-1), _c32(evec1).reshape(-1),
                                         _c32(evec2).reshape(-1), out, s0, s1, s2, float(sigma), int(w), rp, nr,
                                         int(nthreads))
    return out if recv is not None else out.reshape(s0, s1, s2, 9)


def tensorvote3_atv_f64(vol9, sigma, w, recv=None, nthreads=1):
    v = _c32(vol9)
    s0, s1, s2 = v.shape[:3]
    rp, nr, keep = _recv(recv)
    out = np.zeros(((nr if recv is not None else s0 * s1 * s2), 9), np.float64)
    lib().oracle_tensorvote3_atv_f64(v.reshape(-1), out, s0, s1, s2, float(sigma), int(w), rp, nr, int(nthreads))
    return out if recv is not None else out.reshape(s0, s1, s2, 9)


# ---------------------------------------------------------------- saliency threshold, Lee thinning
def otsu_threshold(values):
    """Otsu threshold of float32 values (reconstruct.py:37-63 on a float32 array)."""
    v = _c32(values).reshape(-1)
    return lib().oracle_otsu_threshold_f32(v, v.size)


def fill_enclosed(boundary, close_radius):
    """_fill_enclosed (reconstruct.py:98-116) exactly as the reference calls it:
    scipy.ndimage.binary_closing with a (2r+1)^2 square (r > 0), then
    binary_fill_holes (scipy 1.15.3, the image's version)."""
    from scipy.ndimage import binary_closing, binary_fill_holes
    b = np.asarray(boundary).astype(bool)
    if close_radius and close_radius > 0:
        size = 2 * int(close_radius) + 1
        b = binary_closing(b, structure=np.ones((size, size), dtype=bool))
    return binary_fill_holes(b)


def fill_enclosed_bruteforce(boundary, close_radius):
    """The same from the definitions, by loops (small masks only): dilation and
    erosion with the square, pixels outside the image counting as 0; then a
    breadth-first flood of the background from the border (4-connected); the
    result is every pixel the flood did not reach."""
    b = np.asarray(boundary).astype(bool)
    H, W = b.shape
    r = int(close_radius) if close_radius and close_radius > 0 else 0
    if r:
        def morph(m, dil):
            o = np.zeros_like(m)
            for y in range(H):
                for x in range(W):
                    vals = [m[y + j, x + i] if 0 <= y + j < H and 0 <= x + i < W else False
                            for j in range(-r, r + 1) for i in range(-r, r + 1)]
                    o[y, x] = any(vals) if dil else all(vals)
            return o
        b = morph(morph(b, True), False)
    reach = np.zeros_like(b)
    queue = [(y, x) for y in range(H) for x in range(W)
             if (y in (0, H - 1) or x in (0, W - 1)) and not b[y, x]]
    for y, x in queue:
        reach[y, x] = True
    while queue:
        y, x = queue.pop()
        for j, i in ((y - 1, x), (y + 1, x), (y, x - 1), (y, x + 1)):
            if 0 <= j < H and 0 <= i < W and not b[j, i] and not reach[j, i]:
                reach[j, i] = True
                queue.append((j, i))
    return ~reach


def lee_thin3(vol):
    """computeThinImage (lee_thinning.h:600-680) of a (z, y, x) volume; nonzero = object.
    Returns (uint8 0/1 volume, outer iterations)."""
    v = np.ascontiguousarray(np.asarray(vol) != 0, dtype=np.uint8)
    s0, s1, s2 = v.shape
    it = lib().oracle_lee_thin3(v.ctypes.data, s0, s1, s2)
    return v, it


def thin_exhaustive(nthreads=8):
    out = (C.c_ulonglong * 2)()
    lib().oracle_thin_exhaustive(out, int(nthreads))
    return int(out[0]), int(out[1])


def thin_classify(first, count, nthreads=8):
    f = np.empty(count, np.uint8)
    lib().oracle_thin_classify(int(first), int(count), f.ctypes.data, int(nthreads))
    return f


# ---------------------------------------------------------------- Gaussian
def gaussian_ksize(sigma):
    return int(lib().oracle_gaussian_ksize(float(sigma)))


def gaussian1d(k, mu, sigma, dx, impl="oracle"):
    out = np.empty(k, np.float32)
    if impl == "oracle":
        lib().oracle_gaussian1d(int(k), float(mu), float(sigma), float(dx), out)
    else:
        ref().ref_gaussian1d(int(k), float(mu), float(sigma), float(dx), out)
    return out


def convolve3(vol, kernel, kshape, impl="oracle"):
    """vol: (s2, s1, s0) array with s0 fastest (filter.h index order)."""
    v = _c32(vol)
    i2, i1, i0 = v.shape
    k0, k1, k2 = kshape
    o0, o1, o2 = _u(), _u(), _u()
    out = np.empty((i2 - k2 + 1) * (i1 - k1 + 1) * (i0 - k0 + 1), np.float32)
    fn = lib().oracle_convolve3 if impl == "oracle" else ref().ref_convolve3
    fn(v.reshape(-1), i0, i1, i2, _c32(kernel).reshape(-1), k0, k1, k2, out, C.byref(o0), C.byref(o1), C.byref(o2))
    return out.reshape(o2.value, o1.value, o0.value)


def gaussian_convolve3(vol, sw, sh, sd, impl="oracle"):
    v = _c32(vol)
    sz, sy, sx = v.shape
    kw, kh, kd = gaussian_ksize(sw), gaussian_ksize(sh), gaussian_ksize(sd)
    out = np.empty((sz - kd + 1) * (sy - kh + 1) * (sx - kw + 1), np.float32)
    ox, oy, oz = _u(), _u(), _u()
    fn = lib().oracle_gaussian_convolve3 if impl == "oracle" else ref().ref_gaussian_convolve3
    fn(v.reshape(-1), sx, sy, sz, float(sw), float(sh), float(sd), out, C.byref(ox), C.byref(oy), C.byref(oz))
    return out.reshape(oz.value, oy.value, ox.value)


def gaussian3_same_packed6_f64(t6, kernel, nthreads=1):
    t = _c32(t6)
    s0, s1, s2 = t.shape[:3]
    out = np.empty((s0, s1, s2, 6), np.float64)
    k = _c32(kernel)
    lib().oracle_gaussian3_same_packed6_f64(t.reshape(-1), s0, s1, s2, k, k.shape[0], out, int(nthreads))
    return out


def gradient3(vol, spacing=(1.0, 1.0, 1.0)):
    """tira::volume::gradient (volume.h:974-1021) along x, y, z of a
    (s0, s1, s2) float32 volume (x = axis 2 fastest): forward difference at the
    first voxel of an axis, backward at the last, central inside, each
    (float difference) / spacing evaluated in double and stored as float32."""
    v = np.asarray(vol, np.float32)
    out = []
    for axis, d in zip((2, 1, 0), spacing):
        g = np.empty(v.shape, np.float64)
        a = np.moveaxis(v, axis, 0)
        gg = np.moveaxis(g, axis, 0)
        gg[0] = (a[1] - a[0]).astype(np.float64) / d
        gg[-1] = (a[-1] - a[-2]).astype(np.float64) / d
        gg[1:-1] = (a[2:] - a[:-2]).astype(np.float64) / (2.0 * d)
        out.append(g.astype(np.float32))
    return out


def structure_tensor3(vol, spacing=(1.0, 1.0, 1.0)):
    """g g^T of gradient3 as packed-6 (xx, xy, xz, yy, yz, zz), float32."""
    gx, gy, gz = gradient3(vol, spacing)
    return np.stack([gx * gx, gx * gy, gx * gz, gy * gy, gy * gz, gz * gz], axis=-1)
