"""Drop-in for the reference's pybind11 module `tensorvote_tk`
(python/tira/cuda_ext/tensorvote_tk.cu:206-254), served by libtira_hip.so.

    import tiralib_amd.tensorvote_tk as tensorvote_tk
    out = tensorvote_tk.tensorvote_tk(field, sigma1, sigma2=0.0, power=1, w=0, stick=True, plate=True, samples=20)

Same argument names, defaults, validation messages and RuntimeError type as
the reference module (tensorvote_tk.cu:73-95, 161-178).
"""
from __future__ import annotations

import numpy as np

from . import api
from ._lib import TiraError

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
    if w <= 0:
        sigmax = sigma2 if sigma2 > sigma1 else sigma1
        w = 2 * int(3.0 * np.float32(sigmax)) + 1
        if w < 3:
            w = 3
    if w % 2 == 0:
        w += 1
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
