"""Drop-in for the reference's saliency reconstruction
(python/tira/tensors_lib/reconstruct.py:14-97), served by libtira_hip.so.

    from tiralib_amd import reconstruct
    mask = reconstruct.saliency_threshold(T)            # (H, W, 2, 2) or (H, W, 4) field

Same names, arguments and results: stick saliency |l|max - |l|min, Otsu's
threshold over a 256-bin histogram (numpy.histogram's float32 binning
reproduced on the device), mask = saliency > threshold. The saliency is
computed in float32 on the GPU (the reference's numpy.linalg.eigh follows the
field's dtype). fill=True (scipy closing + hole filling on the host in the
reference) is not served.
"""
from __future__ import annotations

import numpy as np

from . import api


def stick_saliency(T):
    """Per-pixel stick saliency |lambda|max - |lambda|min of a 2D tensor field, float32 (H, W)."""
    return api.stick_saliency2(T)


def _otsu_threshold(values):
    """Otsu's threshold of nonnegative values (256-bin histogram on [0, max])."""
    return api.otsu_threshold(np.ascontiguousarray(values, np.float32).reshape(-1))


def saliency_threshold(T, fill=False, threshold="otsu", close_radius=0):
    """Binary mask of a voted 2D tensor field by thresholding its stick saliency."""
    if fill:
        raise NotImplementedError("fill=True (closing + hole filling) is not served by libtira_hip")
    s = stick_saliency(T)
    thr = _otsu_threshold(s.ravel()) if threshold == "otsu" else float(threshold)
    return s > thr
