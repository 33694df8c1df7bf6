"""CPU restatement of the reference's FMM hillslope lengths.

TEST INFRASTRUCTURE ONLY: the checker of slm_fmm_pairs / slm_fmm_hsl /
slm_hsl_remap (slm_amd/hsl.py).  It restates python/streamlines/mapping.py:
  select_subsegments(do_restrict=False)  :819-839  labels of unmasked pixels
  fmm_lengths                            :841-900  per subsegment s: phi = 1 on
      the 4-dilation of |label| == |s|, -1 on its thin-channel pixels, masked
      elsewhere; distance = skfmm.distance(phi, dx); hsl = 2 * mean(distance),
      kept (as int32, truncated) when n_phi / n_channel > 4
  parse_hsl + remap_hsl(threshold 0)     :923-985  each label's value on the grid
with these stated points:
  * skfmm.distance (scikit-fmm, absent here) is restated from its published
    algorithm in oracle/slm_oracle_fmm.c: order 2 (scikit-fmm's default,
    fmm_lengths(order=2)) or order 1; parity with scikit-fmm itself is
    unpinned, the restatement is pinned by closed forms in tests/test_fmm.py;
    a negative discriminant, on which scikit-fmm raises, leaves the
    subsegment unmeasured as the reference's except clause does;
  * the mean distance of a subsegment is accumulated in fixed point
    (round(d * 2^24) summed in int64), so it does not depend on summation
    order; the reference's float64 mean differs by far less than the int32
    truncation step.  Points no channel pixel reaches are left out of the
    mean (scikit-fmm's handling of them is unpinned).
"""
from __future__ import annotations

import ctypes

import numpy as np
from scipy import ndimage

from oracle import oracle

FIX = float(1 << 24)   # fixed-point scale of the per-subsegment distance sums


def _lib():
    lib = oracle.load()
    lib.slmo_fmm_distance.restype = ctypes.c_int
    lib.slmo_fmm_distance.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_double, ctypes.c_void_p]
    lib.slmo_fmm_update.restype = ctypes.c_double
    lib.slmo_fmm_update.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.slmo_fmm_distance2.restype = ctypes.c_int
    lib.slmo_fmm_distance2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    return lib


def fmm_update(a, b, dx):
    return _lib().slmo_fmm_update(float(a), float(b), float(dx))


def fmm_distance(phi, dom, dx):
    """Signed first-order FMM distance (float64, NaN outside dom) of a binary
    phi (+1 / -1) on the grid dom (bool)."""
    phi = np.ascontiguousarray(np.where(np.asarray(phi) < 0, -1, 1).astype(np.int8))
    dom = np.ascontiguousarray(np.asarray(dom, bool).astype(np.uint8))
    nx, ny = phi.shape
    out = np.empty((nx, ny), np.float64)
    if _lib().slmo_fmm_distance(nx, ny, phi.ctypes.data, dom.ctypes.data, float(dx), out.ctypes.data):
        raise MemoryError("fmm oracle")
    return out


class NegativeDiscriminant(ArithmeticError):
    """The second-order update met a negative discriminant (scikit-fmm raises;
    the reference then leaves the subsegment's length unset)."""


def fmm_distance2(phi, dom, dx, order=2):
    """Signed FMM distance of order 1 or 2 (scikit-fmm's second-order update,
    points accepted in (value, index) order).  Raises NegativeDiscriminant
    where scikit-fmm raises."""
    phi = np.ascontiguousarray(np.where(np.asarray(phi) < 0, -1, 1).astype(np.int8))
    dom = np.ascontiguousarray(np.asarray(dom, bool).astype(np.uint8))
    nx, ny = phi.shape
    out = np.empty((nx, ny), np.float64)
    if order == 1:
        return fmm_distance(phi, dom, dx)
    rc = _lib().slmo_fmm_distance2(nx, ny, phi.ctypes.data, dom.ctypes.data, float(dx), 2, out.ctypes.data)
    if rc == -1:
        raise MemoryError("fmm oracle")
    if rc == -2:
        raise NegativeDiscriminant("negative discriminant in the second-order update")
    return out


def subsegment_problem(label, mapping, is_thinchannel, a):
    """(phi int8, domain bool, n_phi, n_channel) of subsegment |label| == a."""
    phi_bool = np.abs(label) == a
    n_phi = int(phi_bool.sum())
    dom = ndimage.binary_dilation(phi_bool, structure=ndimage.generate_binary_structure(2, 1), iterations=1)
    phi = np.ones(label.shape, np.int8)
    phi[((mapping & is_thinchannel) == is_thinchannel) & phi_bool] = -1
    n_ch = int(((phi == -1) & dom).sum())
    return phi, dom, n_phi, n_ch


def fmm_lengths(label, mapping, mask, is_thinchannel, pixel_size, order=2):
    """Returns (selected int32[], hsl int32[], stats): stats maps |label| to
    (fixed-point distance sum, domain size, n_phi, n_channel).  order 2 uses
    fmm_distance2; a subsegment whose second-order
    march meets a negative discriminant keeps hsl 0, as the reference's
    except clause leaves it (stats entry (None, ...))."""
    label = np.asarray(label).astype(np.int64)
    mapping = np.asarray(mapping).astype(np.uint32)
    sel = np.unique(label[~mask].astype(np.int32))
    sel = sel[sel != 0].astype(np.int32)
    hsl = np.zeros(sel.shape[0], np.int32)
    stats = {}
    for idx, s in enumerate(sel):
        a = abs(int(s))
        if a not in stats:
            phi, dom, n_phi, n_ch = subsegment_problem(label, mapping, np.uint32(is_thinchannel), a)
            if n_ch == 0:
                stats[a] = (0, int(dom.sum()), n_phi, 0)
            else:
                # points no channel reaches (a domain part without channel
                # pixels) stay infinite and are left out of the mean
                try:
                    d = fmm_distance2(phi, dom, pixel_size, order)[dom]
                except NegativeDiscriminant:
                    d = None
                if d is None:
                    stats[a] = (None, int(dom.sum()), n_phi, n_ch)
                else:
                    d = d[np.isfinite(d)]
                    q = int(np.rint(d * FIX).astype(np.int64).sum())
                    stats[a] = (q, int(d.size), n_phi, n_ch)
        q, cnt, n_phi, n_ch = stats[a]
        if n_ch == 0 or q is None:
            continue
        h = 2.0 * (q / FIX) / cnt
        if n_phi / n_ch > 4.0:
            hsl[idx] = np.int32(h)
    return sel, hsl, stats


def remap(label, sel, hsl):
    """remap_hsl with n_hsl_averaging_threshold = 0: each selected label's
    value on its pixels (float32), 0 elsewhere."""
    grid = np.zeros(label.shape, np.float32)
    for s, h in zip(sel, hsl):
        grid[label == s] = np.float32(h)
    return grid


def fmm_mean_hsl(hsl):
    """fmm_mean_hsl of fmm_lengths (mapping.py:892-896), or None."""
    pos = hsl[hsl > 0]
    return float(np.mean(pos)) if pos.size else None


def trace_mean_hsl(grid, threshold=5.0):
    """remap_hsl's mean of the grid values >= hsl_averaging_threshold
    (mapping.py:977-985), or None."""
    v = grid[~np.isnan(grid)]
    v = v[v >= threshold]
    return np.mean(v) if v.shape[0] else None
