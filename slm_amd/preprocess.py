"""The uv field and masks the tracer consumes (input contract of the path).

The numpy functions restate python/streamlines/preprocess.py without the
loop/blockage conditioning (they build the synthetic test inputs and check
slm_gradient_uv):
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


class Preprocess:
    """Preprocess (python/streamlines/preprocess.py:475-621) with the uv field
    computed on the device: do() runs the loop and blockage conditioning of
    conditioned_gradient_vector_field (slm_condition_uv) when
    state.do_condition is set (the reference's default), otherwise the same
    gradient field without the fixes (slm_gradient_uv), then mask_nan_uv.

    Attributes as the reference's: uv_array (float32 [nxp, nyp, 2], host),
    blockages_array / blocked_neighbors_array (uint8 [nx, ny]) and
    where_blockages_array / where_blocked_neighbors_array, where_looped_array
    and n_loops (the four corners of every looped 2x2 cell).  slope_array is
    not computed (the tracer does not read it).  The reference's raw path
    (raw_gradient_vector_field, :623-645) calls an undefined `derivatives`
    and pads twice; the path without fixes here is the conditioned one
    without its fixes.

    state: the JSON state keys (do_condition) and, optionally,
    add_active_mask({name: mask}); params: the JSON preprocess keys;
    geodata: roi_array (float32 [nx, ny]), roi_pixel_size, pad_width."""

    def __init__(self, state, params, geodata, device=None):
        self.state = state
        self.geodata = geodata
        self.device = device
        p = vars(params) if not isinstance(params, dict) else params
        self.do_normalize_speed = bool(p.get("do_normalize_speed", True))
        self.vecsum_threshold = float(p.get("vecsum_threshold", 0.95))
        self.divergence_threshold = float(p.get("divergence_threshold", -0.5))
        self.curl_threshold = float(p.get("curl_threshold", 0.0))
        self.uv_array = None

    def do(self):
        if getattr(self.state, "do_condition", True):
            self.conditioned_gradient_vector_field()
        else:
            self.raw_gradient_vector_field()
        self.mask_nan_uv()

    def _roi(self):
        import torch
        dev = self.device or torch.device("cuda", torch.cuda.current_device())
        return torch.from_numpy(np.ascontiguousarray(self.geodata.roi_array, dtype=np.float32)).to(dev)

    def conditioned_gradient_vector_field(self):
        from . import hip
        uv, mask, maps = hip.condition_uv(self._roi(), int(self.geodata.pad_width),
                                          float(self.geodata.roi_pixel_size), self.do_normalize_speed,
                                          self.vecsum_threshold, self.divergence_threshold, self.curl_threshold,
                                          maps=True)
        m = maps.cpu().numpy()
        self.blockages_array, self.blocked_neighbors_array = m[0], m[1]
        self.where_blockages_array = np.argwhere(m[0])
        self.where_blocked_neighbors_array = np.argwhere(m[1])
        cells = np.argwhere(m[2])
        self.where_looped_array = (cells[:, None, :] + np.array([[0, 0], [1, 0], [0, 1], [1, 1]])[None]).reshape(-1, 2)
        self.n_loops = self.where_looped_array.shape[0]
        self.uv_array = uv.cpu().numpy()
        self._mask = mask.cpu().numpy()

    def raw_gradient_vector_field(self):
        from . import hip
        uv, mask = hip.gradient_uv(self._roi(), int(self.geodata.pad_width), float(self.geodata.roi_pixel_size),
                                   self.do_normalize_speed)
        self.uv_array = uv.cpu().numpy()
        self._mask = mask.cpu().numpy()

    def mask_nan_uv(self):
        """The uv of masked pixels is 0 (done on the device); the mask joins the
        state's active masks as 'uv'.  It is the merged mask (NaN uv, NaN DTM,
        pad ring), which merges with the DTM mask to the same result."""
        add = getattr(self.state, "add_active_mask", None)
        if add is not None:
            add({"uv": self._mask})
