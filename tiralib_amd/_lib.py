"""ctypes loader for libtira_hip.so (the C ABI declared in include/tira_hip.h).

The library is the product path: there is no CPU fallback. If the shared
object is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtira_hip.so")

TIRA_OK = 0
TIRA_ERR_INVALID_ARG = 1
TIRA_ERR_HIP = 2
TIRA_ERR_OUT_OF_MEMORY = 3
TIRA_ERR_UNSUPPORTED = 4
TIRA_ERR_NO_DEVICE = 5

LAYOUT_MAT9 = 0
LAYOUT_PACKED6 = 1

PLATE_AUTO = 0
PLATE_SUM = 1
PLATE_CLOSED = 2
PLATE_MLP = 3
VOTE_SPECTRAL = 4

EIG_REFINE_FP64 = 1
EIG_SCALE_INVARIANT = 2
EIG_SALIENCY = 4
EIG_MAGNITUDE = 8

MLP_NPARAM = 12288
MLP_NBINS = 2048
MLP_TABLE = 0      # TIRA_MLP_TABLE: network tabulated over (distance row, x = 1 - (n.d)^2)
MLP_PER_PAIR = 1   # TIRA_MLP_PER_PAIR: one network evaluation per pair


class TiraError(RuntimeError):
    def __init__(self, status, fn, msg):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


class TiraInvalidArgument(TiraError, ValueError):
    pass


class TiraUnsupported(TiraError, NotImplementedError):
    pass


_vp = C.c_void_p
_sz = C.c_size_t
_u = C.c_uint
_i = C.c_int
_f = C.c_float

# name -> argtypes (all return int status)
SIGNATURES = {
    "tira_hip_device_count": [C.POINTER(_i)],
    "tira_hip_set_device": [_i],
    "tira_hip_malloc": [C.POINTER(_vp), _sz],
    "tira_hip_pointer_is_device": [_vp, C.POINTER(_i)],
    "tira_hip_free": [_vp],
    "tira_hip_memcpy": [_vp, _vp, _sz],
    "tira_hip_memset": [_vp, _i, _sz],
    "tira_hip_synchronize": [],
    "tira_hip_stream_create": [C.POINTER(_vp)],
    "tira_hip_stream_destroy": [_vp],
    "tira_hip_set_stream": [_vp],
    "tira_hip_stream_synchronize": [_vp],
    "tira_hip_event_create": [C.POINTER(_vp)],
    "tira_hip_event_destroy": [_vp],
    "tira_hip_event_record": [_vp, _vp],
    "tira_hip_event_elapsed_ms": [_vp, _vp, C.POINTER(_f)],
    "tira_hip_release_workspace": [],
    "tira_hip_copy_async": [_vp, _vp, _sz],
    "tira_hip_fma_loop_async": [_vp, _u, _i],
    "tira_hip_profile_begin": [],
    "tira_hip_profile_end": [],
    "tira_hip_profile_get": [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)],
    "tira_evals2_symmetric": [_vp, _sz, _vp, _i],
    "tira_evecs2polar_symmetric": [_vp, _vp, _sz, _vp, _i],
    "tira_evals3_symmetric": [_vp, _sz, _vp, _i],
    "tira_evecs3spherical_symmetric": [_vp, _vp, _sz, _vp, _i],
    "tira_eig3_async": [_vp, _sz, _i, _vp, _vp, _vp, _u, _vp],
    "tira_tensorvote3_eig_async": [_vp, _i, _vp, _i, _u, _u, _u, _f, _f, _u, _u, _i, _i, _u, _i, _vp, _u, _vp],
    "tira_eig3_batch_async": [_vp, _vp, _i, _i, _vp, _vp, _vp, _u, _vp],
    "tira_build_neighbors2d": [_i, _f, _f, _u, _u, _i, _i, _vp, _i, C.POINTER(_i)],
    "tira_tensorvote2": [_vp, _vp, _f, _f, _i, _u, _u, _u, _i, _i, _u, _i],
    "tira_tensorvote2_prebuilt": [_vp, _vp, _vp, _i, _f, _f, _i, _u, _u, _u, _i, _i, _u],
    "tira_tensorvote2_prebuilt_noalloc": [_vp, _vp, _vp, _i, _f, _f, _i, _u, _u, _u, _i, _i, _u],
    "tira_build_neighbors2d_async": [_i, _f, _f, _u, _u, _i, _i, _vp],
    "tira_vote2_session_create": [C.POINTER(_vp), _vp, _u, _u],
    "tira_vote2_session_vote": [_vp, _vp, _vp, _i, _f, _f, _i, _u, _i, _i, _u],
    "tira_vote2_session_stats": [_vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)],
    "tira_vote2_session_destroy": [_vp],
    "tira_tensorvote2_tk": [_vp, _vp, _vp, _f, _f, _u, _u, _u, _u, _i, _i, _u],
    "tira_tensorvote3": [_vp, _vp, _u, _u, _u, _f, _f, _u, _u, _i, _i, _i, _u, _i],
    "tira_tensorvote3_tk": [_vp, _vp, _vp, _vp, _u, _u, _u, _f, _f, _u, _u, _i, _i, _u],
    "tira_tensorvote3_async": [_vp, _i, _vp, _i, _u, _u, _u, _f, _f, _u, _u, _i, _i, _u, _i],
    "tira_tensorvote2_atv": [_vp, _vp, _vp, _f, _u, _u, _u],
    "tira_stick_saliency2_async": [_vp, _sz, _vp],
    "tira_spectral_offset_kernels": [_u, _f, _f, _u, _i, _i, _vp, _vp],
    "tira_spectral_offset_kernels_atv": [_u, _f, _vp, _vp],
    "tira_tensorvote3_atv_method_async": [_vp, _i, _vp, _i, _u, _u, _u, _f, _u, _i],
    "tira_fft3_async": [_vp, _u, _u, _u, _u, _i],
    "tira_spectral_length": [_u, _u],
    "tira_spectral_cache_info": [C.POINTER(_u), C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)],
    "tira_otsu_threshold_async": [_vp, _sz, _i, _vp],
    "tira_threshold_async": [_vp, _sz, _i, _vp, _vp],
    "tira_lee_thin3": [_vp, _vp, _u, _u, _u, _vp],
    "tira_lee_thin3_u8": [_vp, _u, _u, _u, _vp],
    "tira_thin3_classify_async": [_u, _u, _vp],
    "tira_skeleton3": [_vp, _i, _u, _u, _u, _vp, _vp, _vp, _vp],
    "tira_fill_holes2": [_vp, _u, _u, _u, _vp, _vp],
    "tira_tensorvote3_atv": [_vp, _vp, _vp, _vp, _u, _u, _u, _f, _u],
    "tira_tensorvote3_atv_async": [_vp, _i, _vp, _i, _u, _u, _u, _f, _u],
    "tira_gaussian_filter3d": [_vp, _u, _u, _u, _f, _f, _f, _vp, C.POINTER(_u), C.POINTER(_u), C.POINTER(_u)],
    "tira_gaussian_convolve2": [_vp, _u, _u, _f, _f, _vp, C.POINTER(_u), C.POINTER(_u)],
    "tira_gaussian_taps": [_u, _f, _vp],
    "tira_gaussian_filter3_tensor_async": [_vp, _vp, _u, _u, _u, _u, _vp, _u, _i],
    "tira_batch_gauss_eig3_create": [C.POINTER(_vp), _i, _vp, _vp, _vp, _vp, _vp, _u, _i],
    "tira_batch_gauss_eig3_run": [_vp],
    "tira_batch_gauss_eig3_plan": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_f), C.POINTER(_u)],
    "tira_batch_gauss_eig3_tile": [_vp, C.POINTER(_i), C.POINTER(_i)],
    "tira_batch_gauss_eig3_set_plan": [_vp, _i, _i, _i],
    "tira_batch_gauss_eig3_destroy": [_vp],
    "tira_synth_eig_tensors_async": [_vp, _sz, C.c_ulonglong, _f, _f, _f, _f, _f, _vp],
    "tira_synth_sheets_async": [_vp, _u, _u, _u, _i, _vp, _f, _f, C.c_ulonglong],
    "tira_synth_vessels_async": [_vp, _u, _u, _u, _vp, _i, _f, C.c_ulonglong],
    "tira_structure_tensor3_async": [_vp, _u, _u, _u, C.c_double, C.c_double, C.c_double, _vp],
    "tira_tensorvote3_pipeline_async": [_vp, _u, _u, _u, C.c_double, C.c_double, C.c_double, _f, _f, _f, _u, _u,
                                        _i, _i, _u, _i, _vp, _f, _vp, _u],
    "tira_plate_mlp_create": [C.POINTER(_vp), _f, _f, _u, _u, C.c_ulonglong],
    "tira_plate_mlp_destroy": [_vp],
    "tira_plate_mlp_train": [_vp, _sz, _u, _u, _f, _f, C.c_ulonglong, C.POINTER(_f), C.POINTER(_f)],
    "tira_plate_mlp_params": [_vp, _vp, _vp],
    "tira_plate_mlp_samples": [_vp, _sz, C.c_ulonglong, _vp, _vp, _vp],
    "tira_plate_mlp_forward": [_vp, _vp, _sz, _vp],
    "tira_plate_mlp_gradients": [_vp, _vp, _vp, _sz, _vp, _vp],
    "tira_plate_mlp_step": [_vp, _vp, _vp, _sz, _f],
    "tira_plate_mlp_affine": [_vp, _vp, _vp],
    "tira_plate_mlp_bin_errors": [_vp, _vp],
    "tira_plate_mlp_refresh_bins": [_vp],
    "tira_tensorvote3_mlp": [_vp, _vp, _i, _vp, _i, _u, _u, _u, _u, _i, _f, C.POINTER(C.c_ulonglong)],
    "tira_tensorvote3_mlp_method": [_vp, _vp, _i, _vp, _i, _u, _u, _u, _u, _i, _f, C.POINTER(C.c_ulonglong), _i],
    "tira_plate_mlp_table_errors": [_vp, _vp],
    "tira_plate_mlp_vote_table": [_vp, _vp, _sz, C.POINTER(_i), C.POINTER(_i), _vp],
    "tira_plate_mlp_set_vote_table": [_vp, _vp, _sz],
}

_lib = None
# Entry points that return a value rather than a status.
RESTYPES = {"tira_spectral_length": _u}

_lock = threading.Lock()


def load():
    """Load libtira_hip.so (raises if it is absent: no CPU fallback exists)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"libtira_hip.so not built at {LIB_PATH}; run `python -m tiralib_amd.build`")
            L = C.CDLL(LIB_PATH)
            L.tira_hip_last_error.restype = C.c_char_p
            L.tira_hip_last_error.argtypes = []
            L.tira_hip_abi_version.restype = _i
            for name, args in SIGNATURES.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = RESTYPES.get(name, _i)
            _lib = L
    return _lib


def exported_symbols():
    return ["tira_hip_abi_version", "tira_hip_last_error", *SIGNATURES.keys()]


def call(name, *args):
    L = load()
    st = getattr(L, name)(*args)
    if st != TIRA_OK:
        msg = (L.tira_hip_last_error() or b"").decode(errors="replace")
        if st == TIRA_ERR_INVALID_ARG:
            raise TiraInvalidArgument(st, name, msg)
        if st == TIRA_ERR_UNSUPPORTED:
            raise TiraUnsupported(st, name, msg)
        raise TiraError(st, name, msg)
    return st


def device_count() -> int:
    L = load()
    n = C.c_int(0)
    L.tira_hip_device_count(C.byref(n))
    return n.value


def ptr(a) -> int:
    """Address of a numpy array or DeviceArray."""
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data


def profile_get(name: str):
    """(average ms, launches) of a library kernel between profile_begin/end."""
    tot, cnt = C.c_double(), C.c_long()
    call("tira_hip_profile_get", name.encode(), C.byref(tot), C.byref(cnt))
    return tot.value / max(cnt.value, 1), cnt.value


class DeviceArray:
    """A device buffer with a numpy-like shape/dtype, owned by this object."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("tira_hip_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            call("tira_hip_memcpy", d.ptr, a.ctypes.data, a.nbytes)
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("tira_hip_memcpy", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def zero(self):
        call("tira_hip_memset", self.ptr, 0, self.nbytes)

    def free(self):
        if self.ptr:
            call("tira_hip_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.tira_hip_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        call("tira_hip_stream_create", C.byref(p))
        self.handle = p.value

    def use(self):
        call("tira_hip_set_stream", self.handle)
        return self

    def synchronize(self):
        call("tira_hip_stream_synchronize", self.handle)

    def destroy(self):
        if self.handle:
            call("tira_hip_stream_destroy", self.handle)
            self.handle = None


class Event:
    def __init__(self):
        p = C.c_void_p()
        call("tira_hip_event_create", C.byref(p))
        self.handle = p.value

    def record(self, stream=None):
        call("tira_hip_event_record", self.handle, stream.handle if stream else None)

    def elapsed_ms(self, end) -> float:
        ms = C.c_float()
        call("tira_hip_event_elapsed_ms", self.handle, end.handle, C.byref(ms))
        return ms.value
