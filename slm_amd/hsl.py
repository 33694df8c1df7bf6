"""FMM hillslope lengths on the GPU: the reference's live HSL path.

Mirrors python/streamlines/mapping.py:
  select_subsegments(do_restrict=False)  :819-839
  fmm_lengths / measure_hsl_fmm          :841-908
  parse_hsl                              :923-950
  remap_hsl(n_hsl_averaging_threshold=0) :953-985
with the per-subsegment skfmm.distance loop replaced by one colouring of the
subsegments and one solve per colour on the GPU (slm_fmm_pairs, slm_fmm_hsl,
slm_hsl_remap in hsl.hip).  The distance is scikit-fmm's second-order fast
march by default (order=2: scikit-fmm's published update, run in acceptance
order, one GPU lane per subsegment; a subsegment whose march meets a negative
discriminant, where scikit-fmm raises, stays unmeasured as the reference's
except clause leaves it), or the first-order solution (order=1: a parallel
relaxation to the fast march's fixed point).  scikit-fmm is absent here, so
parity with it is unpinned beyond the published algorithm.  Each
subsegment's mean distance is summed in fixed point (2^-24), so the result
does not depend on summation order.

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
    failed: np.ndarray | None = None   # uint8 per |label|: 1 where the second-order march failed


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


@dataclass
class FmmPending:
    """slm_fmm_hsl in flight (fmm_start): its device outputs, workspace and
    stream, held until fmm_finish copies the results back."""
    sel_np: np.ndarray
    n_col: int
    order: int
    stream: object
    sums: object
    cnt: object
    nch: object
    nphi: object
    failed: object
    ws: object


def fmm_start(label, mapping, mask, is_thinchannel, pixel_size, stream=None, max_group=None, order=2,
              lanes_only=False):
    """The first half of fmm_lengths: select and colour the subsegments (host
    round trips) and launch the solve on stream (default: the current one)
    without waiting for it.  Work on other streams runs beside the march;
    fmm_finish(pending) waits for it and returns the FmmResult."""
    import torch
    hip._require_cuda(label, mapping, mask)
    dev = label.device
    stream = stream if stream is not None else torch.cuda.current_stream(dev)
    nx, ny = label.shape
    lib = hip.load()
    with torch.cuda.stream(stream):
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
        failed = torch.empty(n_labels, dtype=torch.uint8, device=dev)
        # all colours solved together when they fit in 8 GiB
        one = lib.slm_fmm_workspace_size_k(nx, ny, 1) - 512
        k = max(1, min(int(n_col), (8 << 30) // max(one, 1), int(max_group) if max_group else 1 << 30))
        ws_bytes = lib.slm_fmm_workspace_size_k(nx, ny, k)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        hip.check(lib.slm_fmm_hsl(hip._ptr(label), hip._ptr(mapping), int(is_thinchannel), hip._ptr(colour_d),
                                  n_labels, n_col, nx, ny, ctypes.c_double(float(pixel_size)), int(order),
                                  hip._ptr(sums), hip._ptr(cnt), hip._ptr(nch), hip._ptr(nphi), hip._ptr(failed),
                                  hip._ptr(ws), ws_bytes, hip._stream(stream),
                                  hip.SLM_FLAG_FMM_LANES if lanes_only else 0))
    return FmmPending(sel_np, int(n_col), int(order), stream, sums, cnt, nch, nphi, failed, ws)


def fmm_finish(pend: FmmPending):
    """Wait for fmm_start's solve and return its FmmResult."""
    import torch
    with torch.cuda.stream(pend.stream):
        s_np, c_np = pend.sums.cpu().numpy(), pend.cnt.cpu().numpy().view(np.uint32)
        f_np = pend.failed.cpu().numpy() if pend.order == 2 else np.zeros(s_np.shape[0], np.uint8)
        ch_np, ph_np = pend.nch.cpu().numpy().view(np.uint32), pend.nphi.cpu().numpy().view(np.uint32)
    sel_np = pend.sel_np
    # per selected label, in the reference's order of operations
    # (2 * mean distance; kept when n_phi / n_channel > 4; int32 truncation)
    a = np.abs(sel_np.astype(np.int64))
    q, c = s_np[a].astype(np.float64), c_np[a].astype(np.int64)
    ch, ph = ch_np[a].astype(np.int64), ph_np[a].astype(np.int64)
    ok = (ch > 0) & (c > 0) & (f_np[a] == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        h = 2.0 * (q / FIX) / c
        keep = ok & (ph / np.where(ch > 0, ch, 1) > 4.0)
    hsl = np.where(keep, h, 0.0).astype(np.int32)
    pos = hsl[hsl > 0]
    mean = float(np.mean(pos)) if pos.size else None
    return FmmResult(sel_np, hsl, mean, pend.n_col, s_np, c_np, ch_np, ph_np, f_np)


def fmm_lengths(label, mapping, mask, is_thinchannel, pixel_size, stream=None, max_group=None, order=2,
                lanes_only=False):
    """fmm_lengths (mapping.py:841-900) on device tensors: label int32
    [nx, ny] (signed flank labels), mapping int32 [nx, ny] (uint32 bits),
    mask bool [nx, ny].  order: 2 (scikit-fmm's default) or 1.  max_group
    caps the colours solved together (default: all that fit in 8 GiB).
    lanes_only (test hook): second-order marches one lane per subsegment,
    without the LDS heaps of the larger ones (the same bits).
    Returns an FmmResult (host arrays)."""
    return fmm_finish(fmm_start(label, mapping, mask, is_thinchannel, pixel_size, stream, max_group, order,
                                lanes_only))


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
    values >= the threshold, or None.  A device grid is filtered on the device
    (NaN compares false, so the NaN pixels drop out as in the reference's
    isnan filter) and only the kept values, in the grid's order, come back:
    numpy averages the same float32 array."""
    if hasattr(grid, "cpu"):
        v = grid[grid >= hsl_averaging_threshold].cpu().numpy()
    else:
        g = np.asarray(grid)
        v = g[~np.isnan(g)]
        v = v[v >= hsl_averaging_threshold]
    return np.mean(v) if v.shape[0] else None


# ------------------------------------------------------------ map_hsl / map_aspect

def map_hsl(hsl_grid, mask, hsl_dilation_width=51, hsl_mean_radius=30, pixel_size=1.0, stream=None):
    """Mapping.map_hsl (mapping.py:987-1020) on the device: hsl_grid float32
    [nx, ny] (remap_hsl's grid), mask bool (True = masked, the merged active
    masks).  Returns hsl_smoothed float32 [nx, ny]."""
    import torch
    hip._require_cuda(hsl_grid, mask)
    nx, ny = hsl_grid.shape
    dfw = int(hsl_dilation_width / pixel_size)
    mdr = int(hsl_mean_radius / pixel_size)
    lib = hip.load()
    ws_bytes = lib.slm_map_hsl_workspace_size(nx, ny)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=hsl_grid.device)
    m8 = mask.to(torch.uint8).contiguous()
    out = torch.empty_like(hsl_grid)
    hip.check(lib.slm_map_hsl(hip._ptr(hsl_grid), hip._ptr(m8), nx, ny, dfw, mdr, hip._ptr(out), hip._ptr(ws),
                              ws_bytes, hip._stream(stream), hip.SLM_FLAG_SYNC))
    return out


def slope_deg(roi, pad, pixel_size, stream=None):
    """Preprocess.slope_array (preprocess.py:597) from the ROI DTM (float32
    device [nx, ny]): degrees, 'edge'-padded by pad."""
    import torch
    hip._require_cuda(roi)
    nx, ny = roi.shape
    out = torch.empty((nx + 2 * pad, ny + 2 * pad), dtype=torch.float32, device=roi.device)
    hip.check(hip.load().slm_slope_deg(hip._ptr(roi), nx, ny, int(pad), float(np.float32(pixel_size)),
                                       hip._ptr(out), hip._stream(stream), hip.SLM_FLAG_SYNC))
    return out


def _int_radius(r, name):
    if abs(r - round(r)) > 1e-9:
        raise ValueError("%s = %g pixels: only whole-pixel disk radii are supported" % (name, r))
    return int(round(r))


def map_aspect(slope, uv, mapping, is_channel, aspect_median_filter_radius=4.0, uv_median_radius=50.0,
               pixel_size=1.0, aspect_slope_threshold=15.0, do_aspect_median_filtering=True, stream=None):
    """Mapping.map_aspect (mapping.py:1022-1053) on the device: slope float32
    [nx, ny] (degrees), uv float32 [nx, ny, 2], mapping int32 [nx, ny]
    (uint32 bits).  Returns (aspect float32 [nx, ny], mask bool [nx, ny]).
    The disk radii (metres / pixel_size) must be whole pixels."""
    import torch
    hip._require_cuda(slope, uv, mapping)
    nx, ny = slope.shape
    r_s = _int_radius(aspect_median_filter_radius / pixel_size, "aspect_median_filter_radius")
    r_uv = _int_radius(uv_median_radius / pixel_size, "uv_median_radius")
    lib = hip.load()
    ws_bytes = lib.slm_map_aspect_workspace_size(nx, ny)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=slope.device)
    aspect = torch.empty((nx, ny), dtype=torch.float32, device=slope.device)
    m8 = torch.empty((nx, ny), dtype=torch.uint8, device=slope.device)
    hip.check(lib.slm_map_aspect(hip._ptr(slope), hip._ptr(uv), hip._ptr(mapping), int(is_channel), nx, ny, r_s,
                                 r_uv, 1 if do_aspect_median_filtering else 0, float(aspect_slope_threshold),
                                 hip._ptr(aspect), hip._ptr(m8), hip._ptr(ws), ws_bytes, hip._stream(stream),
                                 hip.SLM_FLAG_SYNC))
    return aspect, m8.bool()


# ------------------------------------------------- HSL by aspect: host statistics

def combine_hsl_aspect(aspect, aspect_mask, hsl_smoothed, mask, pad, n_bins=60, hsl_averaging_threshold=5.0):
    """Mapping.combine_hsl_aspect (mapping.py:1055-1102) on host arrays (the
    device grids of map_aspect and map_hsl, copied back): the unmasked
    (aspect in degrees, HSL >= threshold) pairs sorted by aspect, the mean HSL
    in n_bins + 1 aspect bins centred on -180, ..., 180 degrees (the two end
    bins averaged, since they are one direction), and the [hsl, bin centre in
    radians] array.  mask: the merged active masks (True = masked).  The
    grids are numpy arrays, or torch tensors (the selection then runs on
    their device).  Returns (DataFrame of pairs, per-bin means, averages
    array)."""
    import pandas as pd
    inner = np.index_exp[pad:-pad, pad:-pad]
    if hasattr(aspect, "cpu"):
        # torch tensors (any device): the pixel selection runs where the grids
        # are, and only the kept pairs, in the same row-major order, come back
        import torch
        a, h = aspect.float()[inner], hsl_smoothed.float()[inner]
        m = mask.bool()[inner] | aspect_mask.bool()[inner]
        sel = ~m
        if not bool(sel.any()):
            # nothing unmasked: the reference falls back to every pixel
            sel = torch.ones_like(m)
        sel = sel & (h >= hsl_averaging_threshold)
        hv = h[sel].cpu().numpy()
        asp = np.rad2deg(a[sel].cpu().numpy()).astype(np.float32)
        pairs = np.stack((hv, asp), axis=1)
    else:
        a = np.asarray(aspect, np.float32)
        h = np.asarray(hsl_smoothed, np.float32)
        m = np.asarray(mask, bool)[inner] | np.asarray(aspect_mask, bool)[inner]
        asp, hv = a[inner][~m], h[inner][~m]
        if hv.shape == (0,):
            # nothing unmasked: the reference falls back to every pixel
            asp, hv = a[inner].ravel(), h[inner].ravel()
        asp = np.rad2deg(asp).astype(np.float32)
        keep = hv >= hsl_averaging_threshold
        pairs = np.stack((hv[keep], asp[keep]), axis=1)
    df = pd.DataFrame(columns=["hsl", "aspect"], data=pairs[pairs[:, 1].argsort()])
    half = 180 / n_bins
    width = 2 * half
    edges = np.linspace(-180, 180 + width, n_bins + 2) - half
    df["groups"] = pd.cut(df["aspect"], edges)
    means = df.groupby("groups", observed=False)["hsl"].mean()
    hsl = np.array(means.values, dtype=np.float64)
    west = (hsl[0] + hsl[-1]) / 2.0
    hsl[0] = hsl[-1] = west
    centres = np.deg2rad(edges[:-1] + half)
    return df, means, np.stack((hsl, centres), axis=1)


def hsl_aspect_stats(haa, df, n_hsl_split=4):
    """Mapping.hsl_aspect_stats (mapping.py:1104-1251) on combine_hsl_aspect's
    outputs: mean HSL overall, north- and south-facing (aspect >= 0 / <= 0),
    their disparity and its ratio to the mean split standard deviation, the
    mean HSL vector (magnitude, azimuth in degrees), and for the N-facing
    (45 < aspect < 135) and S-facing (-135 < aspect < -45) HSL values the
    two-sample tests (Kolmogorov-Smirnov, Mann-Whitney, rank-sum, Student and
    Welch t), the Q-Q and P-P curves and the P-P disparity.  Returns a dict."""
    from scipy.interpolate import interp1d
    from scipy.stats import ks_2samp, mannwhitneyu, ranksums, ttest_ind
    haa = np.array(haa, dtype=np.float64)
    ok = ~np.isnan(haa[:, 0])
    hsl, asp = haa[ok, 0], haa[ok, 1]
    south, north = hsl[asp <= 0.0], hsl[asp >= 0.0]
    out = {}
    out["hsl_mean"] = float(np.mean(hsl[1:]) if -asp[0] == asp[-1] else np.mean(hsl))
    out["hsl_mean_south"] = float(np.mean(south)) if south.size else 0.0
    out["hsl_mean_north"] = float(np.mean(north)) if north.size else 0.0
    out["hsl_ns_disparity"] = abs(out["hsl_mean_north"] - out["hsl_mean_south"])
    out["hsl_ns_disparity_normed"] = out["hsl_ns_disparity"] / out["hsl_mean"]
    out["hsl_stddev"] = float(np.std(hsl))
    splits = []
    for part in np.split(haa[1:, 0], n_hsl_split):
        v = part[~np.isnan(part)]
        splits.append(np.std(v) if v.size else 0.0)
    out["hsl_split_stddev"] = float(np.mean(splits)) if splits else 0.0
    out["hsl_split_stddev_normed"] = out["hsl_split_stddev"] / out["hsl_mean"]
    haa[np.isnan(haa[:, 0])] = 0.0
    vec = haa[:, 0] * np.exp(1j * haa[:, 1])
    vec = vec[~np.isnan(vec)]
    mean_vec = np.mean(vec)
    out["hsl_mean_magnitude"] = float(abs(mean_vec))
    out["hsl_mean_azimuth"] = float(np.rad2deg(np.angle(mean_vec)))
    out["hsl_ns_disparity_confidence"] = out["hsl_ns_disparity"] / out["hsl_split_stddev"]
    hsl_n = df.hsl[(df.aspect > +45) & (df.aspect < +135)]
    hsl_s = df.hsl[(df.aspect < -45) & (df.aspect > -135)]
    un, us = hsl_n.unique(), hsl_s.unique()
    nn, ns = un.shape[0], us.shape[0]
    if nn > 0 and ns > 0:
        out["hsl_ns_ks_nm"] = nn * ns
        out["hsl_ns_ks_nmfactor"] = float(np.sqrt((nn + ns) / (nn * ns)))
        ks, mw, rsum = ks_2samp(un, us), mannwhitneyu(un, us), ranksums(un, us)
        tt, welch = ttest_ind(un, us, equal_var=True), ttest_ind(un, us, equal_var=False)
        out["hsl_ns_ks"] = (float(ks.statistic), float(ks.pvalue))
        out["hsl_ns_ttest"] = (float(tt.statistic), float(tt.pvalue))
        out["hsl_ns_mannwhitney"] = (float(mw.statistic) / (nn * ns), float(mw.pvalue))
        out["hsl_ns_ranksum"] = (float(rsum.statistic), float(rsum.pvalue))
        out["hsl_ns_welch"] = (float(welch.statistic), float(welch.pvalue))
        # the reference's builtin min/max over the Series (Python floats, so the
        # averages below are float64): the same values, vectorised
        n_min, s_min = float(hsl_n.min()), float(hsl_s.min())
        n_max, s_max = float(hsl_n.max()), float(hsl_s.max())
        ns_min, ns_max = (n_min + s_min) / 2, (n_max + s_max) / 2
        pc = np.linspace(0, 100, num=50)
        pn, ps = np.percentile(hsl_n, pc), np.percentile(hsl_s, pc)
        pns = (pn + ps) / 2
        qn, qs = interp1d(pns, pn), interp1d(pns, ps)
        fn, fs = interp1d(pn, pc), interp1d(ps, pc)
        s_from_n = fs(np.clip(pn, s_min, s_max))
        n_from_s = fn(np.clip(ps, n_min, n_max))
        grid = np.linspace(ns_min, ns_max, num=50)
        out["hsl_ns_qq_array"] = np.stack((qs(grid), qn(grid)), axis=1).T
        fx = np.concatenate((pc, n_from_s))
        fy = np.concatenate((s_from_n, pc))
        fm = np.concatenate((s_from_n + pc, pc + n_from_s)) / 2
        fxym = np.stack((fx, fy, fm), axis=1)
        fxym = fxym[fxym[:, 0].argsort()].T
        fxm = interp1d(fxym[2], fxym[0])(pc)
        fym = interp1d(fxym[2], fxym[1])(pc)
        fmm = interp1d(fxym[2], fxym[2])(pc)
        pp = np.stack((pc, fxm, fym, fmm), axis=1).T
        out["hsl_ns_pp_array"] = pp
        trap = getattr(np, "trapezoid", None) or getattr(np, "trapz")
        out["hsl_ns_pp_diff"] = float((50 - (trap(pp[1], pp[2]) / 100)) * 2)
        out["hsl_ns_min"], out["hsl_ns_max"] = float(ns_min), float(ns_max)
    return out
