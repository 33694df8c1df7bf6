"""The uv field and masks the tracer consumes (input contract of the path).

Restates python/streamlines/preprocess.py without the loop/blockage
conditioning (which is outside the tracing path):
  compute_topo_gradient_field  :51-62   np.gradient / pixel_size (float32)
  compute_gradient_velocity_field :183-193  u,v = -grad; speed = hypot, non-finite
                                  or 0 -> 1; normalise
  conditioned_gradient_vector_field :584-596  re-normalise (do_normalize_speed),
                                  pad 'edge', stack float32
  mask_nan_uv                  :602-621  NaN uv -> masked, uv set to 0
and the DTM mask of geodata.py:262-273 (NaN -> masked, padded with True),
merged as state.merge_active_masks does (state.py:435-458).
"""
from __future__ import annotations

import numpy as np


def _speed(u, v):
    s = np.hypot(u, v).astype(np.float32)
    s[(~np.isfinite(s)) | (s == 0.0)] = 1.0
    return s


def gradient_uv(roi_array: np.ndarray, pixel_size: float, pad_width: int = 2,
                do_normalize_speed: bool = True):
    """Returns (uv float32 [nxp,nyp,2], uv_mask bool [nxp,nyp])."""
    roi = roi_array.astype(np.float32)
    ps = np.float32(pixel_size)
    gx = (np.gradient(roi, axis=0) / ps).astype(np.float32)
    gy = (np.gradient(roi, axis=1) / ps).astype(np.float32)
    u, v = -gx, -gy
    s = _speed(u, v)
    u, v = (u / s).astype(np.float32), (v / s).astype(np.float32)
    if do_normalize_speed:
        s = _speed(u, v)
        u, v = (u / s).astype(np.float32), (v / s).astype(np.float32)
    pad = int(pad_width)
    uv = np.stack((np.pad(u, (pad, pad), "edge"), np.pad(v, (pad, pad), "edge")), axis=2).astype(np.float32)
    uv_mask = np.isnan(uv[:, :, 0]) | np.isnan(uv[:, :, 1])
    uv[uv_mask] = [0.0, 0.0]
    return uv, uv_mask


def dtm_mask(roi_array: np.ndarray, pad_width: int = 2, h_min=None) -> np.ndarray:
    m = np.isnan(roi_array)
    if h_min is not None:
        with np.errstate(invalid="ignore"):
            m |= roi_array <= h_min
    pad = int(pad_width)
    return np.pad(m, (pad, pad), "constant", constant_values=(True, True))


def prepare(roi_array: np.ndarray, pixel_size: float, pad_width: int = 2):
    """DTM -> (merged mask, uv): what Trace hands the tracer."""
    uv, uv_mask = gradient_uv(roi_array, pixel_size, pad_width)
    mask = dtm_mask(roi_array, pad_width) | uv_mask
    return mask, uv
