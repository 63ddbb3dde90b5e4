"""Drop-in for the reference's saliency reconstruction
(python/tira/tensors_lib/reconstruct.py:14-116), served by libtira_hip.so.

    from tiralib_amd import reconstruct
    mask = reconstruct.saliency_threshold(T)            # (H, W, 2, 2) or (H, W, 4) field

Same names, arguments and results: stick saliency |l|max - |l|min, Otsu's
threshold over a 256-bin histogram (numpy.histogram's float32 binning
reproduced on the device), mask = saliency > threshold. The saliency is
computed in float32 on the GPU (the reference's numpy.linalg.eigh follows the
field's dtype). fill=True closes the thresholded mask with a (2r+1)^2 square
(close_radius r > 0) and fills the holes on the GPU, bit-identical to the
reference's scipy.ndimage binary_closing + binary_fill_holes.
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
    s = stick_saliency(T)
    thr = _otsu_threshold(s.ravel()) if threshold == "otsu" else float(threshold)
    binary = s > thr
    if not fill:
        return binary
    return _fill_enclosed(binary, close_radius)


def _fill_enclosed(boundary, close_radius):
    """Close gaps in a boundary mask (square of side 2r+1), then fill the
    background that cannot reach the image border through background."""
    return api.fill_holes2(boundary, close_radius)
