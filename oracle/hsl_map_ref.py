"""CPU restatement of Mapping.map_hsl and Mapping.map_aspect.

TEST INFRASTRUCTURE ONLY: the checker of slm_map_hsl / slm_map_aspect
(slm_amd/hsl.py).  It restates python/streamlines/mapping.py:
  map_hsl     :987-1020  uint16 rescaling of the HSL grid, grey_dilation
                         (dfw x dfw, scipy, into masked pixels only), mean
                         over disk(mdr) (skimage.filters.rank.mean), rescale
  map_aspect  :1022-1053 median over disk(r) (skimage.filters.rank.median)
                         of the uint8-quantised slope and of u, v; aspect =
                         arctan2(v, u) as float32; mask = gentle or channel
The dilation is scipy's own routine.  scikit-image is absent here: its rank
mean and median are restated from their definitions, with the footprint
disk(r) = {x^2 + y^2 <= r^2} and only pixels inside the image counted (pop):
  mean   = (sum of the values) / pop in float64, truncated to the input type;
  median = the smallest value v with 2 * #(values <= v) >= pop;
parity against skimage is unpinned.  Brute force over the footprint offsets.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def disk_offsets(r):
    r = int(r)
    return [(dx, dy) for dx in range(-r, r + 1) for dy in range(-r, r + 1) if dx * dx + dy * dy <= r * r]


def _stack(img, r):
    """values and in-image flags of every footprint offset: [k, nx, ny]."""
    nx, ny = img.shape
    vals, ok = [], []
    pad = int(r)
    big = np.zeros((nx + 2 * pad, ny + 2 * pad), img.dtype)
    big[pad:pad + nx, pad:pad + ny] = img
    inside = np.zeros(big.shape, bool)
    inside[pad:pad + nx, pad:pad + ny] = True
    for dx, dy in disk_offsets(r):
        vals.append(big[pad + dx:pad + dx + nx, pad + dy:pad + dy + ny])
        ok.append(inside[pad + dx:pad + dx + nx, pad + dy:pad + dy + ny])
    return np.stack(vals), np.stack(ok)


def rank_mean(img, r):
    v, ok = _stack(img, r)
    pop = ok.sum(0).astype(np.float64)
    s = np.where(ok, v.astype(np.float64), 0.0).sum(0)
    return np.where(pop > 0, s / np.maximum(pop, 1), 0).astype(img.dtype)


def rank_median(img, r):
    v, ok = _stack(img, r)
    pop = ok.sum(0)
    out = np.zeros(img.shape, img.dtype)
    done = np.zeros(img.shape, bool)
    for b in range(int(np.iinfo(img.dtype).max) + 1):
        c = (ok & (v <= b)).sum(0)
        hit = (~done) & (2 * c >= pop) & (pop > 0)
        out[hit] = b
        done |= hit
        if done.all():
            break
    return out


def map_hsl(hsl, mask, hsl_dilation_width, hsl_mean_radius, pixel_size):
    """mapping.py:987-1020: hsl float32 [nx, ny]; mask bool (True = masked)."""
    hsl = np.asarray(hsl, np.float32)
    hsl_min, hsl_max = np.min(hsl), np.max(hsl)
    hsl_clipped = (65535 * (hsl - hsl_min) / (hsl_max - hsl_min)).astype(np.uint16)
    dfw = int(hsl_dilation_width / pixel_size)
    mdr = int(hsl_mean_radius / pixel_size)
    if dfw == 0:
        hsl_dilated = hsl_clipped
    else:
        hsl_dilated = ndimage.grey_dilation(hsl_clipped, size=(dfw, dfw))
        hsl_dilated[~mask] = hsl_clipped[~mask]
    hsl_mean = rank_mean(hsl_dilated, mdr)
    return ((hsl_mean.astype(np.float32)) / 65535) * (hsl_max - hsl_min) + hsl_min


def map_aspect(slope, uv, mapping, is_channel, aspect_median_filter_radius, uv_median_radius, pixel_size,
               aspect_slope_threshold=15.0, do_aspect_median_filtering=True):
    """mapping.py:1022-1053: slope float32 [nx, ny] (degrees), uv float32
    [nx, ny, 2], mapping uint32.  Returns (aspect float32, mask bool)."""
    slope_array = np.asarray(slope, np.float32).copy()
    chan = (mapping & np.uint32(is_channel)) == np.uint32(is_channel)
    slope_array[chan] = 0.0
    median_radius = aspect_median_filter_radius / pixel_size
    if do_aspect_median_filtering:
        sf = np.max(slope_array) / 255.0
        slope_array = sf * rank_median(np.uint8(slope_array / sf), median_radius)
    r = uv_median_radius / pixel_size
    sf = 2.0 / 255.0
    uvx = sf * rank_median(np.uint8((uv[:, :, 0] + 1.0) / sf), r) - 1.0
    uvy = sf * rank_median(np.uint8((uv[:, :, 1] + 1.0) / sf), r) - 1.0
    mask = (slope_array < aspect_slope_threshold) | chan
    return np.arctan2(uvy, uvx).astype(np.float32), mask
