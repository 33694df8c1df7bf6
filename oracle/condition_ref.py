"""CPU restatement of the reference's DTM conditioning (loops and blockages).

TEST INFRASTRUCTURE ONLY: the checker of slm_condition_uv.  It restates
python/streamlines/preprocess.py, Preprocess.conditioned_gradient_vector_field
(:531-600) and the functions it calls:

  find_blockages            :647-698  two scipy generic_filter passes (size 3,
                                      mode 'reflect') over roi.T
  has_one_diagonal_outflow  :390-432
  upstream_of_diagonal_outflow :434-473
  compute_topo_gradient_field / compute_gradient_velocity_field :51-62, :177-193
  check_has_loop            :229-261  (numba: the mean vector is float64)
  find_and_fix_loops        :303-350  + break_out_of_loop :263-301
  fix_blockages             :352-388
  get_flow_vector           :196-227
  normalize / unnormalize, pad, mask_nan_uv :119-175, :584-621

Two restatements of the blockage maps are given: `blockage_codes_filter` runs
scipy.ndimage.generic_filter with literal restatements of the reference's
callbacks (the reference's own mechanism), `blockage_codes` is a vectorised
numpy version; tests check them against each other and against hand-built
cases.

One deliberate difference: find_and_fix_loops evaluates check_has_loop at the
last row and column, where it reads u[x+1], v[y+1] past the array (numba does
no bounds check; the values are whatever memory follows).  Here those cells
are never loops.  Cells whose 2x2 lies inside the grid are restated exactly.
"""
from __future__ import annotations

import numpy as np

# get_flow_vector (:196-227): priority SW, SE, NW, NE, W, E, S, N
_FLOW = [(16, (-1, -1)), (32, (+1, -1)), (64, (-1, +1)), (128, (+1, +1)),
         (1, (-1, 0)), (2, (+1, 0)), (4, (0, -1)), (8, (0, +1))]


def flow_vector(nn: int):
    nn = int(nn)
    for bit, vec in _FLOW:
        if nn & bit:
            return vec
    return (0, 0)


# ------------------------------------------------------------ blockages

def _has_one_diagonal_outflow(w):
    """preprocess.py:390-432 on a flattened 3x3 window of roi.T."""
    h = w[4]
    nn = 0
    if w[0] < h:
        nn += 16   # SW
    if w[1] < h:
        nn += 4    # S
    if w[2] < h:
        nn += 32   # SE
    if w[3] < h:
        nn += 1    # W
    if w[5] < h:
        nn += 2    # E
    if w[6] < h:
        nn += 64   # NW
    if w[7] < h:
        nn += 8    # N
    if w[8] < h:
        nn += 128  # NE
    return nn if (nn & (16 + 32 + 64 + 128)) == nn else 0


def _upstream_of_diagonal_outflow(w):
    """preprocess.py:434-473 on a flattened 3x3 window of the blockage codes."""
    nn = 0
    if w[0] == 16:
        nn += 16
    if w[1] == 16:
        nn += 4
    if w[1] == 32:
        nn += 4
    if w[2] == 32:
        nn += 32
    if w[3] == 16:
        nn += 1
    if w[3] == 64:
        nn += 1
    if w[5] == 128:
        nn += 2
    if w[5] == 32:
        nn += 2
    if w[6] == 64:
        nn += 64
    if w[7] == 64:
        nn += 8
    if w[7] == 128:
        nn += 8
    if w[8] == 128:
        nn += 128
    return nn


def blockage_codes_filter(roi):
    """find_blockages (:647-698) with scipy's generic_filter, as the reference
    runs it: returns (blockages, blocked_neighbors), uint8 [nx, ny]."""
    from scipy.ndimage import generic_filter
    rt = np.ascontiguousarray(roi.T)
    b = np.zeros_like(rt, dtype=np.uint8)
    generic_filter(rt, _has_one_diagonal_outflow, output=b, size=3)
    n = np.zeros_like(b)
    generic_filter(b, _upstream_of_diagonal_outflow, output=n, size=3)
    return b.T.copy(), n.T.copy()


def _neighbours(a):
    """The 3x3 window of every pixel of a [nx, ny] array with scipy's 'reflect'
    border (numpy 'symmetric'), as w[k] of the reference's callbacks: the
    callbacks see roi.T, so w[3*(dy+1) + (dx+1)] is a[x+dx, y+dy]."""
    p = np.pad(a, 1, mode="symmetric")
    nx, ny = a.shape
    return [p[1 + dx:1 + dx + nx, 1 + dy:1 + dy + ny] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def blockage_codes(roi):
    """Vectorised restatement of blockage_codes_filter."""
    w = _neighbours(roi.astype(np.float64))
    h = w[4]
    with np.errstate(invalid="ignore"):
        nn = ((w[0] < h) * 16 + (w[1] < h) * 4 + (w[2] < h) * 32 + (w[3] < h) * 1 + (w[5] < h) * 2
              + (w[6] < h) * 64 + (w[7] < h) * 8 + (w[8] < h) * 128).astype(np.int64)
    b = np.where((nn & 240) == nn, nn, 0).astype(np.uint8)
    c = _neighbours(b)
    n = ((c[0] == 16) * 16 + (c[1] == 16) * 4 + (c[1] == 32) * 4 + (c[2] == 32) * 32 + (c[3] == 16) * 1
         + (c[3] == 64) * 1 + (c[5] == 128) * 2 + (c[5] == 32) * 2 + (c[6] == 64) * 64 + (c[7] == 64) * 8
         + (c[7] == 128) * 8 + (c[8] == 128) * 128).astype(np.uint8)
    return b, n


# ------------------------------------------------------------ loops

def _speed(u, v):
    s = np.hypot(u, v).astype(np.float32)
    s[(~np.isfinite(s)) | (s == 0.0)] = 1.0
    return s


def velocity_field(roi, pixel_size):
    """compute_topo_gradient_field + compute_gradient_velocity_field:
    normalised u, v and the raw speed (float32 [nx, ny])."""
    r = roi.astype(np.float32)
    ps = np.float32(pixel_size)
    gx = (np.gradient(r, axis=0) / ps).astype(np.float32)
    gy = (np.gradient(r, axis=1) / ps).astype(np.float32)
    u, v = -gx, -gy
    s = _speed(u, v)
    return (u / s).astype(np.float32), (v / s).astype(np.float32), s


def loop_cells(u, v, vecsum_threshold, divergence_threshold, curl_threshold):
    """check_has_loop (:229-261) for every cell whose 2x2 lies in the grid:
    bool [nx-1, ny-1].  The mean vector is float64 (numba types the float32
    array divided by 4.0 as float64); the dot products with +-1 vectors and
    their sums are float32."""
    f32 = np.float32
    u00, v00 = u[:-1, :-1], v[:-1, :-1]
    u10, v10 = u[1:, :-1], v[1:, :-1]
    u11, v11 = u[1:, 1:], v[1:, 1:]
    u01, v01 = u[:-1, 1:], v[:-1, 1:]
    mu = (((u00 + u10).astype(f32) + u11).astype(f32) + u01).astype(f32).astype(np.float64) / 4.0
    mv = (((v00 + v10).astype(f32) + v11).astype(f32) + v01).astype(f32).astype(np.float64) / 4.0
    speed = np.sqrt(mu * mu + mv * mv).astype(f32)
    div = ((((-u00 - v00).astype(f32) + (u10 - v10).astype(f32)).astype(f32) + (u11 + v11).astype(f32)).astype(f32)
           + (-u01 + v01).astype(f32)).astype(f32)
    curl = ((((u00 - v00).astype(f32) + (u10 + v10).astype(f32)).astype(f32) + (-u11 + v11).astype(f32)).astype(f32)
            + (-u01 - v01).astype(f32)).astype(f32)
    with np.errstate(invalid="ignore"):
        return (speed.astype(np.float64) < vecsum_threshold) & (div.astype(np.float64) <= divergence_threshold) \
            & (np.abs(curl).astype(np.float64) >= curl_threshold)


def break_out_of_loop(x, y, u, v, roi):
    """break_out_of_loop (:263-301): point the vector at the lowest of the 8
    neighbours (order W, E, S, N, SW, SE, NW, NE; strict <, so the first
    lowest wins and NaN never does); only well inside the grid."""
    nx, ny = roi.shape
    if not (1 <= x <= nx - 2 and 1 <= y <= ny - 2):
        return
    order = [flow_vector(b) for b in (1, 2, 4, 8, 16, 32, 64, 128)]
    h_min = np.finfo(np.float32).max
    best = 0
    for k, (dx, dy) in enumerate(order):
        h = roi[x + dx, y + dy]
        if h_min > h:
            h_min = h
            best = k
    dx, dy = order[best]
    n = np.hypot(float(dx), float(dy))
    u[x, y] = np.float32(dx / n)
    v[x, y] = np.float32(dy / n)


# ------------------------------------------------------------ the whole step

def conditioned_uv(roi, pixel_size, pad_width=2, do_normalize_speed=True, vecsum_threshold=0.95,
                   divergence_threshold=-0.5, curl_threshold=0.0, blockages=None):
    """conditioned_gradient_vector_field + mask_nan_uv: returns (uv float32
    [nx+2p, ny+2p, 2], uv mask bool, blockages uint8, blocked neighbours
    uint8, looped cells bool [nx-1, ny-1]).  blockages: optional precomputed
    (blockages, blocked_neighbors) pair."""
    roi = np.asarray(roi, np.float32)
    b, nb = blockage_codes(roi) if blockages is None else blockages
    u, v, raw = velocity_field(roi, pixel_size)
    loops = loop_cells(u, v, vecsum_threshold, divergence_threshold, curl_threshold)
    # find_and_fix_loops: the four corners of every looped cell, in scan order
    for xi, yi in zip(*np.nonzero(loops)):
        for x, y in ((xi, yi), (xi + 1, yi), (xi, yi + 1), (xi + 1, yi + 1)):
            break_out_of_loop(x, y, u, v, roi)
    # fix_blockages: blockages, then blocked neighbours
    r2 = np.sqrt(2.0)
    for codes in (b, nb):
        for x, y in zip(*np.nonzero(codes)):
            dx, dy = flow_vector(codes[x, y])
            u[x, y] = np.float32(np.float32(dx) / r2)
            v[x, y] = np.float32(np.float32(dy) / r2)
    if do_normalize_speed:
        s = _speed(u, v)
        u, v = (u / s).astype(np.float32), (v / s).astype(np.float32)
    else:
        u, v = (u * raw).astype(np.float32), (v * raw).astype(np.float32)
    p = int(pad_width)
    uv = np.stack((np.pad(u, (p, p), "edge"), np.pad(v, (p, p), "edge")), axis=2).astype(np.float32)
    uv_mask = np.isnan(uv[:, :, 0]) | np.isnan(uv[:, :, 1])
    uv[uv_mask] = [0.0, 0.0]
    return uv, uv_mask, b, nb, loops
