"""CPU restatement of Mapping.map_midslopes / map_ridges (mapping.py:779-817).

TEST INFRASTRUCTURE ONLY: the checker of slm_map_midslopes_ridges and
slm_gaussian_filter.  The smoothing, dilation and hole filling are scipy's
own ndimage routines, as the reference calls them; skeletonize (skimage,
absent here) is replaced by the Zhang-Suen thinning of the mapping oracle
(oracle/slm_oracle_map.c), the same stand-in the GPU uses, so the skeleton
step is parity unpinned against skimage.  One deliberate difference: the
sla angle is the correctly rounded float32 arctan2 (through float64);
numpy's float32 arctan2, which the reference calls, is a vectorised
approximation that differs from it by an ulp on a large share of inputs
(CPU and numpy version dependent).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

from oracle import oracle

IS_THINCHANNEL = 0x2
IS_CHANNEL = 0x1


def gaussian_weights(sigma, truncate=4.0):
    """scipy.ndimage._gaussian_kernel1d for order 0: (weights float64, radius)."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    w = np.exp(-0.5 / (float(sigma) ** 2) * x ** 2)
    return w / w.sum(), r


def sla_angle(sla):
    """arctan2(dsla, usla) - pi/4, float32 (sla [..., 2]: downstream, upstream)."""
    a = np.arctan2(sla[..., 0].astype(np.float64), sla[..., 1].astype(np.float64)).astype(np.float32)
    return a - np.float32(np.pi / 4)


def thin(img, p):
    """Zhang-Suen thinning of a bool image by the mapping oracle (p: any
    SlmParams; its grid fields are set here)."""
    p.nx_padded, p.ny_padded, p.nxy_padded = img.shape[0], img.shape[1], img.size
    m = np.where(img, np.uint32(IS_CHANNEL), np.uint32(0)).astype(np.uint32).reshape(-1)
    oracle.map_call("thin_channels", p, mapping=m)
    return ((m & IS_THINCHANNEL) != 0).reshape(img.shape)


def midslopes_ridges(sla, mask, p, midslope_filter_sigma=5.0, midslope_threshold=0.03, ridge_filter_sigma=2.0,
                     ridge_threshold=0.8):
    """(midslope skeleton, ridge skeleton, smoothed angle for midslopes,
    smoothed angle for ridges) on a grid; mask True = masked."""
    ang = sla_angle(sla)
    gm = ndimage.gaussian_filter(ang, midslope_filter_sigma)
    gr = ndimage.gaussian_filter(ang, ridge_filter_sigma)
    mid = (~mask) & (np.fabs(gm) <= midslope_threshold)
    ridge = (~mask) & (gr <= -(np.pi / 4) * ridge_threshold)
    fat = ndimage.binary_dilation(ridge, structure=ndimage.generate_binary_structure(2, 2), iterations=2)
    fat = ndimage.binary_fill_holes(fat)
    return thin(mid, p), thin(fat, p), gm, gr
