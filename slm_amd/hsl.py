"""FMM hillslope lengths on the GPU: the reference's live HSL path.

Mirrors python/streamlines/mapping.py:
  select_subsegments(do_restrict=False)  :819-839
  fmm_lengths / measure_hsl_fmm          :841-908
  parse_hsl                              :923-950
  remap_hsl(n_hsl_averaging_threshold=0) :953-985
with the per-subsegment skfmm.distance loop replaced by one colouring of the
subsegments and one solve per colour on the GPU (slm_fmm_pairs, slm_fmm_hsl,
slm_hsl_remap in hsl.hip).  The distance is the first-order fast-marching
solution (scikit-fmm, absent here, defaults to second order: parity with the
reference unpinned); each subsegment's mean distance is summed in fixed
point (2^-24), so the result does not depend on summation order.

Typical use, after the segmentation chain has labelled the grid:
    res = fmm_lengths(label, mapping, mask, info.is_thinchannel, pixel_size)
    grid = remap_hsl(label, res)                 # float32 device grid
    stats = parse_hsl(res, coarse_label)         # pandas, as the reference
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import hip

FIX = float(1 << 24)


@dataclass
class FmmResult:
    selected: np.ndarray          # int32 signed subsegment labels (select_subsegments)
    hsl: np.ndarray               # int32 per selected label (0 where not kept)
    fmm_mean_hsl: float | None    # mean of the positive hsl values
    n_colours: int
    sums: np.ndarray              # int64 per |label|: fixed-point distance sums
    cnt: np.ndarray               # uint32 per |label|: points a channel reaches
    nch: np.ndarray               # uint32 per |label|: channel pixels
    nphi: np.ndarray              # uint32 per |label|: own pixels


def select_subsegments(label, mask):
    """Signed labels of the unmasked pixels, without 0 (torch int32, sorted)."""
    import torch
    sel = torch.unique(label[~mask.bool()])
    return sel[sel != 0].to(torch.int32)


def colour_labels(n_labels, problems, pairs):
    """Greedy colouring (slm_greedy_colour, libslm_host.so): labels joined by
    a pair get different colours.  problems: |label| values to colour; pairs:
    unique int64 lo << 32 | hi.  Returns (colour uint8 [n_labels], 0xff for
    non-problems; number of colours)."""
    from .synth import _hostlib
    colour = np.full(n_labels, 0xFF, np.uint8)
    colour[np.asarray(problems, np.int64)] = 0
    pairs = np.ascontiguousarray(np.asarray(pairs, np.int64).view(np.uint64))
    lib = _hostlib()
    fn = lib.slm_greedy_colour
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    n_col = fn(n_labels, colour.ctypes.data, pairs.ctypes.data if pairs.size else None, pairs.size)
    if n_col < 0:
        raise RuntimeError("subsegment colouring failed (%d)" % n_col)
    return colour, n_col


def fmm_lengths(label, mapping, mask, is_thinchannel, pixel_size, stream=None):
    """fmm_lengths (mapping.py:841-900) on device tensors: label int32
    [nx, ny] (signed flank labels), mapping int32 [nx, ny] (uint32 bits),
    mask bool [nx, ny].  Returns an FmmResult (host arrays)."""
    import torch
    hip._require_cuda(label, mapping, mask)
    dev = label.device
    nx, ny = label.shape
    lib = hip.load()
    sel = select_subsegments(label, mask)
    sel_np = sel.cpu().numpy().astype(np.int32)
    absl = label.to(torch.int64).abs()
    n_labels = int(absl.max().item()) + 1 if label.numel() else 1
    problems = np.unique(np.abs(sel_np.astype(np.int64)))
    flag = np.full(n_labels, 0xFF, np.uint8)
    flag[problems] = 0
    flag_d = torch.from_numpy(flag).to(dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    cap = max(1 << 16, 2 * (nx + ny))
    while True:
        pairs = torch.empty(cap, dtype=torch.int64, device=dev)
        hip.check(lib.slm_fmm_pairs(hip._ptr(label), nx, ny, hip._ptr(flag_d), n_labels, hip._ptr(pairs), cap,
                                    hip._ptr(count), hip._stream(stream), hip.SLM_FLAG_SYNC))
        found = int(count.item())
        if found <= cap:
            break
        cap = found
    uniq = torch.unique(pairs[:found]).cpu().numpy() if found else np.zeros(0, np.int64)
    colour, n_col = colour_labels(n_labels, problems, uniq)
    colour_d = torch.from_numpy(colour).to(dev)
    sums = torch.empty(n_labels, dtype=torch.int64, device=dev)
    cnt = torch.empty(n_labels, dtype=torch.int32, device=dev)
    nch = torch.empty(n_labels, dtype=torch.int32, device=dev)
    nphi = torch.empty(n_labels, dtype=torch.int32, device=dev)
    ws_bytes = lib.slm_fmm_workspace_size(nx, ny)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    hip.check(lib.slm_fmm_hsl(hip._ptr(label), hip._ptr(mapping), int(is_thinchannel), hip._ptr(colour_d), n_labels,
                              n_col, nx, ny, ctypes.c_double(float(pixel_size)), hip._ptr(sums), hip._ptr(cnt),
                              hip._ptr(nch), hip._ptr(nphi), hip._ptr(ws), ws_bytes, hip._stream(stream),
                              hip.SLM_FLAG_SYNC))
    s_np, c_np = sums.cpu().numpy(), cnt.cpu().numpy().view(np.uint32)
    ch_np, ph_np = nch.cpu().numpy().view(np.uint32), nphi.cpu().numpy().view(np.uint32)
    # per selected label, in the reference's order of operations
    # (2 * mean distance; kept when n_phi / n_channel > 4; int32 truncation)
    a = np.abs(sel_np.astype(np.int64))
    q, c = s_np[a].astype(np.float64), c_np[a].astype(np.int64)
    ch, ph = ch_np[a].astype(np.int64), ph_np[a].astype(np.int64)
    ok = (ch > 0) & (c > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        h = 2.0 * (q / FIX) / c
        keep = ok & (ph / np.where(ch > 0, ch, 1) > 4.0)
    hsl = np.where(keep, h, 0.0).astype(np.int32)
    pos = hsl[hsl > 0]
    mean = float(np.mean(pos)) if pos.size else None
    return FmmResult(sel_np, hsl, mean, n_col, s_np, c_np, ch_np, ph_np)


def remap_hsl(label, res: FmmResult, stream=None):
    """remap_hsl with n_hsl_averaging_threshold = 0 (the FMM workflow,
    mapping.py:490-494): each selected label's HSL on its pixels, float32
    device grid of label's shape, 0 elsewhere."""
    import torch
    hip._require_cuda(label)
    dev = label.device
    n_labels = int(label.to(torch.int64).abs().max().item()) + 1 if label.numel() else 1
    hsl_abs = np.zeros(n_labels, np.int32)
    sel_bits = np.zeros(n_labels, np.uint8)
    a = np.abs(res.selected.astype(np.int64))
    hsl_abs[a] = res.hsl          # +s and -s share one subsegment problem, hence one value
    np.bitwise_or.at(sel_bits, a, np.where(res.selected > 0, 1, 2).astype(np.uint8))
    hsl_d = torch.from_numpy(hsl_abs).to(dev)
    sel_d = torch.from_numpy(sel_bits).to(dev)
    grid = torch.empty(label.shape, dtype=torch.float32, device=dev)
    hip.check(hip.load().slm_hsl_remap(hip._ptr(label), label.numel(), hip._ptr(hsl_d), hip._ptr(sel_d), n_labels,
                                       hip._ptr(grid), hip._stream(stream), hip.SLM_FLAG_SYNC))
    return grid


def parse_hsl(res: FmmResult, coarse_label=0):
    """parse_hsl (mapping.py:923-950): per fine label count / mean / stddev
    of its HSL (one value each in the FMM workflow), indexed (fine, coarse)."""
    import pandas as pd
    df = pd.DataFrame({"fine": res.selected, "length": res.hsl})
    df = df[df.fine != 0]
    stats_df = pd.DataFrame(res.selected, columns=["fine"])
    for stat, name in (("count", "count"), ("mean", "mean [m]"), ("std", "stddev [m]")):
        stats_df = stats_df.join(getattr(df.groupby("fine"), stat)(), on="fine")
        stats_df.rename(index=str, columns={"length": name}, inplace=True)
    stats_df["coarse"] = coarse_label
    stats_df.set_index(["fine", "coarse"], inplace=True)
    return stats_df


def trace_mean_hsl(grid, hsl_averaging_threshold=5.0):
    """remap_hsl's trace_mean_hsl (mapping.py:977-985): numpy mean of the grid
    values >= the threshold, or None."""
    g = grid.cpu().numpy() if hasattr(grid, "cpu") else np.asarray(grid)
    v = g[~np.isnan(g)]
    v = v[v >= hsl_averaging_threshold]
    return np.mean(v) if v.shape[0] else None
