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
    if sigma < 0 or (sigma == 0 and not is_atv):
        raise ValueError("sigma must be > 0 (ATV also takes 0: each voxel keeps its own tensor)")
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


def spectral_offset_kernels(w, sigma1, sigma2=0.0, stick=True, plate=True):
    """Offset kernels of the spectral vote, (6, nf, W, W, W) float64 indexed
    [c, f, dw, dv, du] with W = 2 (w // 2) + 1 (host computation, no GPU)."""
    W = 2 * (int(w) // 2) + 1
    nf = 15 * (int(bool(stick)) + int(bool(plate)))
    out = np.empty((6, nf, W, W, W), np.float64)
    got = C.c_uint()
    call("tira_spectral_offset_kernels", int(w), float(sigma1), float(sigma2), 1, int(stick), int(plate),
         out.ctypes.data, C.byref(got))
    assert got.value == nf
    return out


def spectral_offset_kernels_atv(w, sigma):
    """Offset kernels of the spectral analytical vote, (6, 31, W, W, W) float64 [c, f, dw, dv, du]."""
    W = 2 * (int(w) // 2) + 1
    out = np.empty((6, 31, W, W, W), np.float64)
    got = C.c_uint()
    call("tira_spectral_offset_kernels_atv", int(w), float(sigma), out.ctypes.data, C.byref(got))
    assert got.value == 31
    return out


def spectral_length(n, w):
    """FFT length of an axis of n voxels with window w (0 when above 1024)."""
    return int(_lib.load().tira_spectral_length(int(n), int(w)))


def fft3(data, inverse=False):
    """In-place 3D FFT of a DeviceArray of complex volumes (nchan, L0, L1, L2, 2) float32 (blocking)."""
    nchan, L0, L1, L2 = data.shape[:4]
    call("tira_fft3_async", data.ptr, nchan, L0, L1, L2, int(bool(inverse)))
    call("tira_hip_synchronize")
    return data


def tensorvote3_atv(lambdas, evec1, evec2, shape, sigma, w):
    """tira::cuda::_tensorvote3_atv (tensorvote.h:2695-2739): analytical tensor voting
    from eigenpairs (ascending eigenvalues, middle and largest eigenvectors)."""
    s0, s1, s2 = shape
    l, e1, e2 = _f32(lambdas), _f32(evec1), _f32(evec2)
    out = _like(l, (s0, s1, s2, 9))
    call("tira_tensorvote3_atv", ptr(out), ptr(l), ptr(e1), ptr(e2), s0, s1, s2, float(sigma), int(w))
    return out


def tensorvote3_atv_async(field, out, shape, sigma, w, in_layout=_lib.LAYOUT_MAT9, out_layout=_lib.LAYOUT_MAT9):
    """Native non-blocking 3D analytical tensor voting on device arrays (library stream)."""
    s0, s1, s2 = shape
    call("tira_tensorvote3_atv_async", field.ptr, int(in_layout), out.ptr, int(out_layout), s0, s1, s2,
         float(sigma), int(w))
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


# ------------------------------------------------------------------ Gaussian filter
def gaussian_taps(ksize, sigma):
    """tira::cpu::gaussian1d(ksize, 0, sigma, 1) (filter.h:37-64)."""
    out = np.empty(int(ksize), np.float32)
    call("tira_gaussian_taps", int(ksize), float(sigma), out.ctypes.data)
    return out


def gaussian_filter3d(src, sigma_w, sigma_h, sigma_d):
    """tira::cuda::gaussian_filter3d (filter.h:400-534): valid-mode, src (depth, height, width)."""
    s = _f32(src)
    if len(s.shape) != 3:
        raise ValueError("volume must be 3D (depth, height, width)")
    d, h, w = s.shape
    ow, oh, od = C.c_uint(), C.c_uint(), C.c_uint()
    call("tira_gaussian_filter3d", ptr(s), w, h, d, float(sigma_w), float(sigma_h), float(sigma_d), None,
         C.byref(ow), C.byref(oh), C.byref(od))
    out = _like(s, (od.value, oh.value, ow.value))
    call("tira_gaussian_filter3d", ptr(s), w, h, d, float(sigma_w), float(sigma_h), float(sigma_d), ptr(out),
         C.byref(ow), C.byref(oh), C.byref(od))
    return out


def gaussian_convolve2(src, sigma1, sigma2):
    """tira::cuda::gaussian_convolve2 (filter.h:304-397): valid-mode, src (height, width)."""
    s = _f32(src)
    if len(s.shape) != 2:
        raise ValueError("image must be 2D (height, width)")
    h, w = s.shape
    ow, oh = C.c_uint(), C.c_uint()
    call("tira_gaussian_convolve2", ptr(s), w, h, float(sigma1), float(sigma2), None, C.byref(ow), C.byref(oh))
    out = _like(s, (oh.value, ow.value))
    call("tira_gaussian_convolve2", ptr(s), w, h, float(sigma1), float(sigma2), ptr(out), C.byref(ow), C.byref(oh))
    return out


def gaussian_filter3_tensor(field, taps, same=True):
    """Native fused separable filter of a (s0, s1, s2, C) device field, C in {1, 6}."""
    if not isinstance(field, DeviceArray):
        raise ValueError("gaussian_filter3_tensor works on DeviceArray inputs")
    s0, s1, s2 = field.shape[:3]
    comp = field.shape[3] if len(field.shape) == 4 else 1
    k = np.ascontiguousarray(taps, np.float32)
    if same:
        out = DeviceArray(field.shape)
    else:
        K = k.shape[0]
        out = DeviceArray((s0 - K + 1, s1 - K + 1, s2 - K + 1) + ((comp,) if comp > 1 else ()))
    call("tira_gaussian_filter3_tensor_async", field.ptr, out.ptr, s0, s1, s2, comp, k.ctypes.data, k.shape[0],
         int(same))
    call("tira_hip_synchronize")
    return out


def eig3_batch(tensors, counts=None, layout=_lib.LAYOUT_PACKED6, evals=None, flags=None):
    """Eigenvalues of many device tensor arrays in one launch
    (tira_eig3_batch_async). `tensors`: DeviceArrays; returns (and fills) one
    n x 3 DeviceArray per input. Non-blocking on the current stream."""
    per = 9 if layout == _lib.LAYOUT_MAT9 else 6
    if counts is None:
        counts = [_count(t, per) for t in tensors]
    if evals is None:
        evals = [DeviceArray((c, 3)) for c in counts]
    if flags is None:
        flags = _lib.EIG_REFINE_FP64 | _lib.EIG_SCALE_INVARIANT
    n = len(tensors)
    P = C.c_void_p * max(n, 1)
    tin = P(*[t.ptr for t in tensors])
    tout = P(*[e.ptr for e in evals])
    cnt = np.ascontiguousarray(counts, np.uint64)
    call("tira_eig3_batch_async", C.cast(tin, C.c_void_p), cnt.ctypes.data, n, layout, C.cast(tout, C.c_void_p),
         None, None, flags, None)
    return evals


# ------------------------------------------------------------------ batched volumes
class BatchGaussEig3:
    """Persistent, graph-captured Gaussian smoothing + eig3 of many uneven
    6-component device volumes (tira_batch_gauss_eig3_*)."""

    def __init__(self, volumes, taps, autotune=True):
        self.volumes = list(volumes)
        self.out = [DeviceArray(v.shape) for v in self.volumes]
        self.evals = [DeviceArray(v.shape[:3] + (3,)) for v in self.volumes]
        n = len(self.volumes)
        P = C.c_void_p * n
        self._in = P(*[v.ptr for v in self.volumes])
        self._out = P(*[o.ptr for o in self.out])
        self._ev = P(*[e.ptr for e in self.evals])
        self._shapes = np.ascontiguousarray([v.shape[:3] for v in self.volumes], np.uint32)
        self._taps = np.ascontiguousarray(taps, np.float32)
        h = C.c_void_p()
        call("tira_batch_gauss_eig3_create", C.byref(h), n, C.cast(self._in, C.c_void_p), C.cast(self._out, C.c_void_p),
             C.cast(self._ev, C.c_void_p), self._shapes.ctypes.data, self._taps.ctypes.data, self._taps.shape[0],
             int(autotune))
        self.handle = h.value

    def run(self):
        call("tira_batch_gauss_eig3_run", self.handle)

    def plan(self):
        zc, occ, ms, items = C.c_int(), C.c_int(), C.c_float(), C.c_uint()
        call("tira_batch_gauss_eig3_plan", self.handle, C.byref(zc), C.byref(occ), C.byref(ms), C.byref(items))
        tx, ty = C.c_int(), C.c_int()
        call("tira_batch_gauss_eig3_tile", self.handle, C.byref(tx), C.byref(ty))
        return {"tile": [tx.value, ty.value], "zchunk": zc.value, "blocks_per_cu": occ.value, "tuned_ms": ms.value,
                "items": items.value}

    def set_plan(self, tile_y, zchunk, blocks_per_cu):
        """Fix the launch plan (tile height 8/16/32, z-chunk, blocks per CU) and re-capture."""
        call("tira_batch_gauss_eig3_set_plan", self.handle, int(tile_y), int(zchunk), int(blocks_per_cu))

    def close(self):
        if self.handle:
            call("tira_batch_gauss_eig3_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ learned plate vote
class PlateMLP:
    """BF16 plate-vote surrogate (tira_plate_mlp_*): trained on the device
    against the FP64 integrator; votes through tira_tensorvote3_mlp."""

    NPARAM = _lib.MLP_NPARAM
    LAYERS = (("W1", 0, 64, 32), ("W2", 2048, 64, 64), ("W3", 6144, 64, 64), ("W4", 10240, 32, 64))

    def __init__(self, sigma1, sigma2=0.0, power=1, integrator_samples=64, seed=1004):
        h = C.c_void_p()
        call("tira_plate_mlp_create", C.byref(h), float(sigma1), float(sigma2), int(power), int(integrator_samples),
             int(seed))
        self.handle = h.value
        self.sigma1, self.sigma2, self.power = float(sigma1), float(sigma2), int(power)

    def train(self, nsamples=1 << 22, epochs=100, batch=4096, lr=3e-3, lr_final=1e-6, seed=1004):
        tr, va = C.c_float(), C.c_float()
        call("tira_plate_mlp_train", self.handle, int(nsamples), int(epochs), int(batch), float(lr), float(lr_final),
             int(seed), C.byref(tr), C.byref(va))
        return tr.value, va.value

    def params(self):
        p = np.empty(self.NPARAM, np.float32)
        call("tira_plate_mlp_params", self.handle, p.ctypes.data, None)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (self.NPARAM,):
            raise ValueError(f"expected {self.NPARAM} parameters")
        call("tira_plate_mlp_params", self.handle, None, p.ctypes.data)

    def affine(self):
        a = np.empty(12, np.float32)
        call("tira_plate_mlp_affine", self.handle, a.ctypes.data, None)
        return a[:6].copy(), a[6:].copy()

    def set_affine(self, mu, sd):
        a = np.ascontiguousarray(np.concatenate([mu, sd]), np.float32)
        call("tira_plate_mlp_affine", self.handle, None, a.ctypes.data)

    def samples(self, n, seed, raw=False):
        X = DeviceArray((n, 32), np.uint16)
        T = DeviceArray((n, 8))
        R = DeviceArray((n, 8)) if raw else None
        call("tira_plate_mlp_samples", self.handle, int(n), int(seed), X.ptr, T.ptr, R.ptr if R else None)
        return X, T, R

    def forward(self, X):
        n = X.shape[0]
        Y = DeviceArray((n, 32))
        call("tira_plate_mlp_forward", self.handle, X.ptr, n, Y.ptr)
        return Y

    def gradients(self, X, T):
        g = np.empty(self.NPARAM, np.float32)
        sums = np.empty(2, np.float32)
        call("tira_plate_mlp_gradients", self.handle, X.ptr, T.ptr, X.shape[0], g.ctypes.data, sums.ctypes.data)
        return g, sums

    def step(self, X, T, lr):
        call("tira_plate_mlp_step", self.handle, X.ptr, T.ptr, X.shape[0], float(lr))

    def bin_errors(self):
        e = np.empty(_lib.MLP_NBINS, np.float32)
        call("tira_plate_mlp_bin_errors", self.handle, e.ctypes.data)
        return e

    def table_errors(self):
        e = np.empty(_lib.MLP_NBINS, np.float32)
        call("tira_plate_mlp_table_errors", self.handle, e.ctypes.data)
        return e

    def refresh_bins(self):
        call("tira_plate_mlp_refresh_bins", self.handle)

    def vote_table(self):
        """(table (rows, cells, 8) float32, shells (squared distances of rows 1..K))
        of the tabulated network the vote reads (tira_plate_mlp_vote_table)."""
        rows, cells = C.c_int(), C.c_int()
        call("tira_plate_mlp_vote_table", self.handle, None, 0, C.byref(rows), C.byref(cells), None)
        tab = np.empty((rows.value, cells.value, 8), np.float32)
        shells = np.zeros(max(rows.value - 2, 1), np.int32)
        call("tira_plate_mlp_vote_table", self.handle, tab.ctypes.data, tab.size, None, None, shells.ctypes.data)
        return tab, shells[:rows.value - 2].copy()

    def set_vote_table(self, tab):
        """Replaces the vote table (rows, cells, 8) until the next weight change (test hook)."""
        tab = np.ascontiguousarray(tab, np.float32)
        call("tira_plate_mlp_set_vote_table", self.handle, tab.ctypes.data, tab.size)

    def vote(self, field, out, shape, w, stick=True, tol=1e-3, in_layout=_lib.LAYOUT_MAT9,
             out_layout=_lib.LAYOUT_PACKED6, count_fallback=False, method=_lib.MLP_TABLE):
        """Stick + learned plate vote of a device field; returns the number of
        integrator-fallback voters when count_fallback (synchronises).
        method: _lib.MLP_TABLE (tabulated network, default) or _lib.MLP_PER_PAIR."""
        nf = C.c_ulonglong(0)
        call("tira_tensorvote3_mlp_method", self.handle, field.ptr, in_layout, out.ptr, out_layout,
             *[int(s) for s in shape], int(w), int(stick), float(tol), C.byref(nf) if count_fallback else None,
             int(method))
        return nf.value if count_fallback else None

    def close(self):
        if self.handle:
            call("tira_plate_mlp_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ front end and pipeline
def structure_tensor3(intensity, spacing=(1.0, 1.0, 1.0), out=None):
    """Gradient (volume.h:974-1021 edge rules) and structure tensor g g^T of a
    device (s0, s1, s2) volume: returns a (s0, s1, s2, 6) packed DeviceArray."""
    shape = intensity.shape
    if out is None:
        out = DeviceArray(tuple(shape) + (6,))
    call("tira_structure_tensor3_async", intensity.ptr, *[int(s) for s in shape], *[float(x) for x in spacing],
         out.ptr)
    return out


def pipeline3(intensity, smooth_sigma=2.0, vote_sigma=4.0, w=None, power=1, stick=True, plate=True, samples=20,
              plate_mode=_lib.PLATE_AUTO, mlp=None, mlp_tol=4e-3, spacing=(1.0, 1.0, 1.0), saliency=None,
              flags=_lib.EIG_SCALE_INVARIANT):
    """Structure tensor -> Gaussian -> stick + plate vote -> eig3 saliency
    (l2 - l1, l1 - l0, l0) of a device intensity volume (non-blocking)."""
    shape = intensity.shape
    if w is None:
        w = 2 * int(round(3 * vote_sigma)) + 1
    if saliency is None:
        saliency = DeviceArray(tuple(shape) + (3,))
    call("tira_tensorvote3_pipeline_async", intensity.ptr, *[int(s) for s in shape], *[float(x) for x in spacing],
         float(smooth_sigma), float(vote_sigma), 0.0, int(w), int(power), int(stick), int(plate), int(samples),
         int(plate_mode), mlp.handle if mlp is not None else None, float(mlp_tol), saliency.ptr, int(flags))
    return saliency


def tensorvote_atv(lambdas, thetas, shape, sigma, w):
    """tira::cuda::tensorvote_atv (2D, tensorvote.h:2029-2066) from eigenpairs (H, W, 2) each."""
    H, W = shape
    l, t = _f32(lambdas), _f32(thetas)
    out = _like(l, (H, W, 4))
    call("tira_tensorvote2_atv", ptr(out), ptr(l), ptr(t), float(sigma), int(w), H, W)
    return out


# ------------------------------------------------------------------ saliency, threshold, thinning
def _dev(a):
    """(device array, owned) for a host array or an existing DeviceArray."""
    if isinstance(a, DeviceArray):
        return a, False
    return DeviceArray.from_host(np.ascontiguousarray(a)), True


def stick_saliency2(field):
    """Stick saliency |l|max - |l|min of an (H, W, 4) or (H, W, 2, 2) field (reconstruct.py:14-34), float32."""
    f = _f32(field)
    H, W = f.shape[0], f.shape[1]
    f = f.reshape(H, W, 4)
    d, o = DeviceArray.from_host(f), DeviceArray((H, W))
    try:
        call("tira_stick_saliency2_async", d.ptr, H * W, o.ptr)
        call("tira_hip_synchronize")
        return o.to_host()
    finally:
        d.free(); o.free()


def otsu_threshold(values, stride=1, n=None):
    """Otsu threshold (reconstruct.py:37-63) of values[::stride] (host array or DeviceArray)."""
    d, own = _dev(np.asarray(values, np.float32) if not isinstance(values, DeviceArray) else values)
    if n is None:
        n = int(np.prod(d.shape)) // int(stride)
    t = DeviceArray((1,))
    try:
        call("tira_otsu_threshold_async", d.ptr, int(n), int(stride), t.ptr)
        call("tira_hip_synchronize")
        return float(t.to_host()[0])
    finally:
        t.free()
        if own:
            d.free()


def saliency_threshold2(T, threshold="otsu", fill=False, close_radius=0):
    """reconstruct.saliency_threshold(T, fill, threshold, close_radius) on the GPU: bool mask (H, W)."""
    s = stick_saliency2(T)
    thr = otsu_threshold(s) if threshold == "otsu" else float(threshold)
    binary = s > thr
    return fill_holes2(binary, close_radius) if fill else binary


def fill_holes2(boundary, close_radius=0, return_sweeps=False):
    """_fill_enclosed(boundary, close_radius) (reconstruct.py:98-116) on the GPU:
    closing with a (2r+1)^2 square when close_radius > 0, then hole filling.
    bool (H, W) in and out."""
    b = np.ascontiguousarray(np.asarray(boundary).astype(bool), dtype=np.uint8)
    if b.ndim != 2:
        raise ValueError(f"boundary must be 2D, got shape {b.shape}")
    r = int(close_radius) if close_radius and close_radius > 0 else 0
    out = np.empty_like(b)
    sweeps = C.c_uint()
    call("tira_fill_holes2", ptr(b), b.shape[0], b.shape[1], r, ptr(out), C.byref(sweeps))
    out = out.astype(bool)
    return (out, int(sweeps.value)) if return_sweeps else out


def lee_thin3(volume):
    """tira::lee / computeThinImage (lee_thinning.h:600-706) of a (z, y, x) volume.
    Returns (int32 0/1 volume, outer iterations)."""
    v = np.ascontiguousarray(volume, dtype=np.int32)
    s0, s1, s2 = v.shape
    out = np.empty_like(v)
    it = C.c_uint()
    call("tira_lee_thin3", ptr(v), ptr(out), s2, s1, s0, C.byref(it))
    return out, int(it.value)


def skeleton3(voted, shape, layout=_lib.LAYOUT_PACKED6):
    """Voted field (DeviceArray) -> magnitude-ranked saliency, Otsu threshold of the
    stick saliency, Lee-thinned mask. Returns dict of DeviceArrays and numbers."""
    s0, s1, s2 = shape
    n = s0 * s1 * s2
    sal, thr, sk = DeviceArray((n, 3)), DeviceArray((1,)), DeviceArray((s0, s1, s2), np.uint8)
    it = C.c_uint()
    call("tira_skeleton3", voted.ptr, int(layout), s0, s1, s2, sal.ptr, thr.ptr, sk.ptr, C.byref(it))
    return {"saliency": sal, "threshold": float(thr.to_host()[0]), "skeleton": sk, "iterations": int(it.value)}
