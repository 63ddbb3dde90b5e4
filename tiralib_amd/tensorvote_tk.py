"""Drop-in for the reference's pybind11 module `tensorvote_tk`
(python/tira/cuda_ext/tensorvote_tk.cu:206-254), served by libtira_hip.so.

    import tiralib_amd.tensorvote_tk as tensorvote_tk
    out = tensorvote_tk.tensorvote_tk(field, sigma1, sigma2=0.0, power=1, w=0, stick=True, plate=True, samples=20)

Same argument names, defaults, validation messages and RuntimeError type as
the reference module (tensorvote_tk.cu:73-95, 161-178).

For a parameter search that votes one field many times (optimize_tk,
optimize.py:256-288), bind the field once:

    with tensorvote_tk.Session(field) as s:
        out = s.tensorvote_tk_prebuilt(nb_buf, nb_count, w, sigma1, sigma2, power)

The session keeps the field and its recent neighbour tables on the device and
runs each vote as one graph replay; results are identical to the module
functions.
"""
from __future__ import annotations

import numpy as np

import ctypes as C

from . import api
from ._lib import TiraError, call, ptr

NEIGHBOR2D_DTYPE = api.NEIGHBOR2D_DTYPE


def _field(field):
    f = np.ascontiguousarray(field, dtype=np.float32)
    if f.ndim != 3 or f.shape[2] != 4:
        raise RuntimeError("field must have shape (H, W, 4)")
    if f.shape[0] <= 0 or f.shape[1] <= 0:
        raise RuntimeError("field must have positive H and W")
    return f


def tensorvote_tk(field, sigma1, sigma2=0.0, power=1, w=0, stick=True, plate=True, samples=20):
    """CUDA-accelerated 2D TK voting in the reference; gfx950 here. Returns (H, W, 4) float32."""
    f = _field(field)
    if sigma1 <= 0.0:
        raise RuntimeError("sigma1 must be > 0")
    if power < 1:
        power = 1
    w = _window(sigma1, sigma2, w)
    try:
        return api.tensorvote(f, float(sigma1), float(sigma2), int(power), int(w), bool(stick), bool(plate),
                              int(samples))
    except TiraError as e:
        raise RuntimeError(str(e)) from e


def tensorvote_tk_prebuilt(field, nb_buf, nb_count, w, sigma1, sigma2=0.0, power=1, stick=True, plate=True,
                           samples=20):
    """TK voting with a caller-built Neighbor2D table (48 bytes per entry)."""
    f = _field(field)
    if sigma1 <= 0.0:
        raise RuntimeError("sigma1 must be > 0")
    if power < 1:
        power = 1
    if w <= 0 or w % 2 == 0:
        raise RuntimeError("w must be a positive odd integer")
    raw = np.frombuffer(memoryview(nb_buf).cast("B"), dtype=np.uint8)
    if raw.size != int(nb_count) * NEIGHBOR2D_DTYPE.itemsize:
        raise RuntimeError("nb_buf byte size does not match nb_count * sizeof(Neighbor2D) (48)")
    nb = raw.view(NEIGHBOR2D_DTYPE)
    try:
        return api.tensorvote_prebuilt(f, nb, float(sigma1), float(sigma2), int(power), int(w), bool(stick),
                                       bool(plate), int(samples))
    except TiraError as e:
        raise RuntimeError(str(e)) from e


def _window(sigma1, sigma2, w):
    """The window rule of tensorvote_tk (tensorvote_tk.cu:88-95)."""
    if w <= 0:
        sigmax = sigma2 if sigma2 > sigma1 else sigma1
        w = 2 * int(3.0 * np.float32(sigmax)) + 1
        if w < 3:
            w = 3
    if w % 2 == 0:
        w += 1
    return w


class Session:
    """A field bound once for repeated votes: the module's two functions with
    the field argument dropped. Holds device memory until close()."""

    def __init__(self, field):
        f = _field(field)
        self.shape = f.shape
        h = C.c_void_p()
        try:
            call("tira_vote2_session_create", C.byref(h), ptr(f), f.shape[0], f.shape[1])
        except TiraError as e:
            raise RuntimeError(str(e)) from e
        self._h = h

    def _vote(self, nb, nb_count, sigma1, sigma2, power, w, stick, plate, samples):
        if self._h is None:
            raise RuntimeError("session is closed")
        out = np.empty(self.shape, np.float32)
        try:
            call("tira_vote2_session_vote", self._h, ptr(out), nb, int(nb_count), float(sigma1), float(sigma2),
                 int(power), int(w), bool(stick), bool(plate), int(samples))
        except TiraError as e:
            raise RuntimeError(str(e)) from e
        return out

    def tensorvote_tk(self, sigma1, sigma2=0.0, power=1, w=0, stick=True, plate=True, samples=20):
        """tensorvote_tk(field, ...) on the bound field."""
        if sigma1 <= 0.0:
            raise RuntimeError("sigma1 must be > 0")
        if power < 1:
            power = 1
        return self._vote(None, 0, sigma1, sigma2, power, _window(sigma1, sigma2, w), stick, plate, samples)

    def tensorvote_tk_prebuilt(self, nb_buf, nb_count, w, sigma1, sigma2=0.0, power=1, stick=True, plate=True,
                               samples=20):
        """tensorvote_tk_prebuilt(field, nb_buf, ...) on the bound field."""
        if sigma1 <= 0.0:
            raise RuntimeError("sigma1 must be > 0")
        if power < 1:
            power = 1
        if w <= 0 or w % 2 == 0:
            raise RuntimeError("w must be a positive odd integer")
        raw = np.ascontiguousarray(np.frombuffer(memoryview(nb_buf).cast("B"), dtype=np.uint8))
        if raw.size != int(nb_count) * NEIGHBOR2D_DTYPE.itemsize:
            raise RuntimeError("nb_buf byte size does not match nb_count * sizeof(Neighbor2D) (48)")
        return self._vote(raw.ctypes.data if raw.size else None, nb_count, sigma1, sigma2, power, w, stick, plate,
                          samples)

    def stats(self):
        """(graph replays, table cache hits, table uploads) so far."""
        r, h, b = C.c_ulonglong(), C.c_ulonglong(), C.c_ulonglong()
        call("tira_vote2_session_stats", self._h, C.byref(r), C.byref(h), C.byref(b))
        return int(r.value), int(h.value), int(b.value)

    def close(self):
        if getattr(self, "_h", None) is not None:
            call("tira_vote2_session_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
