"""Drop-in for the reference's pybind11 module `tensorvote_atv`
(python/tira/cuda_ext/tensorvote_atv.cu:53-135), served by libtira_hip.so.

    import tiralib_amd.tensorvote_atv as tensorvote_atv
    out = tensorvote_atv.tensorvote_atv(field, sigma, w=0, stick=True, plate=True, samples=10)

Same argument names, defaults, window rule, validation messages and
RuntimeError type as the reference module (tensorvote_atv.cu:62-81).
stick, plate and samples are accepted and, as in the reference, do not
change the analytical vote.
"""
from __future__ import annotations

import numpy as np

from . import api
from ._lib import TiraError


def tensorvote_atv(field, sigma, w=0, stick=True, plate=True, samples=10):
    """2D analytical tensor voting of an (H, W, 4) float32 field [a, b, b, c]; returns (H, W, 4)."""
    f = np.ascontiguousarray(field, dtype=np.float32)
    if f.ndim != 3 or f.shape[2] != 4:
        raise RuntimeError("field must have shape (H, W, 4)")
    if f.shape[0] <= 0 or f.shape[1] <= 0:
        raise RuntimeError("field must have positive H and W")
    sigma = float(np.float32(sigma))
    if sigma <= 0.0:
        raise RuntimeError("sigma must be > 0")
    w = int(w)
    if w <= 0:
        w = 2 * int(np.float32(3.0) * np.float32(sigma)) + 1
        if w < 3:
            w = 3
    if w % 2 == 0:
        w += 1
    try:
        return api.tensorvote(f, sigma, 0.0, 1, w, bool(stick), bool(plate), int(samples), is_atv=True)
    except TiraError as e:
        raise RuntimeError(str(e)) from e
