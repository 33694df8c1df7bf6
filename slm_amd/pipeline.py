"""The full mapping pipeline of one DTM tile on the GPU, stage by stage.

The device-side sequence behind Trace.do() and the mapping passes
(python/streamlines/trace.py:74-166, mapping.py:627-776, lengths.py,
kde.py): uv from the DTM, the trajectory pass for the ds statistics, both
fields directions, the fields finishing, channel mapping and segmentation,
traced hillslope lengths from every hillslope pixel, and the Gaussian KDE of
those lengths.  Inputs stay in HBM between stages; only the small statistics
(1000 trajectories, the KDE histogram) come back to the host.  Each stage is
timed with HIP events on the current stream.
"""
from __future__ import annotations

import time

import numpy as np

from . import hip
from . import hsl as fmm_hsl
from .kde import Univariate_distribution
from .mapping import ChannelMapper, pick_seeds
from .params import kernel_params
from .useful import compute_stats, create_seeds


class _Stages:
    def __init__(self):
        import torch
        self.torch = torch
        self.marks = []

    def mark(self, name):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record()
        self.marks.append((name, ev, time.perf_counter()))

    def report(self):
        self.torch.cuda.synchronize()
        out = {}
        for (name, e0, w0), (_, e1, w1) in zip(self.marks[:-1], self.marks[1:]):
            out[name] = {"gpu_ms": e0.elapsed_time(e1), "wall_ms": (w1 - w0) * 1e3}
        return out


def map_tile(job, segmentation_threshold=20, channel_quantile=0.97, fp_contract=False, device=None, condition=False,
             midslopes=True, midslope_filter_sigma=5.0, midslope_threshold=0.03, ridge_filter_sigma=2.0,
             ridge_threshold=0.8, fmm=False, return_grids=False):
    """Run the pipeline on a workloads.Job; returns (results dict, stage timings).
    condition: the uv stage runs the loop and blockage conditioning
    (slm_condition_uv, the reference's do_condition) instead of the plain
    gradient field the job's grids were built with.  midslopes: the
    mid-slope and ridge flags of Mapping.pass1 (mapping.py:286-287, the JSON
    mapping defaults) are set before the channels are mapped.  fmm: also the
    FMM hillslope lengths of every subsegment (slm_amd.hsl, the reference's
    live HSL path) and their grid.  return_grids: results["grids"] holds the
    device grids (uv, mask, sla, slt, slc, mapping, label, fmm_hsl) and the
    trajectories' host lengths and step counts [n, 2] (traj_length,
    traj_nsteps)."""
    import torch
    dev = device or torch.device("cuda", torch.cuda.current_device())
    info = job.info
    pad = int(info.pad_width)
    st = _Stages()
    roi = torch.from_numpy(np.ascontiguousarray(job.roi, dtype=np.float32)).to(dev)
    mask_np = job.mask
    nxp, nyp = mask_np.shape
    seeds = torch.from_numpy(job.seeds).to(dev)
    # host seed lists (useful.create_seeds, legacy MT19937 shuffle) are inputs, made first
    tseeds, n_t, _ = create_seeds(mask_np, pad, int(info.n_work_items), n_seed_points=info.n_trajectory_seed_points,
                                  do_shuffle=info.do_shuffle, rng_seed=info.shuffle_rng_seed)
    ts = torch.from_numpy(tseeds).to(dev)
    torch.cuda.synchronize()

    st.mark("uv")
    if condition:
        uv_full, mask_full = hip.condition_uv(roi, pad, float(job.pixel_size))
    else:
        uv_full, mask_full = hip.gradient_uv(roi, pad, float(job.pixel_size))
    ox, oy = job.origin
    # the tracer's grids: the bbox crop of the full padded grids (Data, useful.py:37-115)
    uv = uv_full[ox:ox + nxp, oy:oy + nyp].contiguous()
    mask_b = mask_full[ox:ox + nxp, oy:oy + nyp].contiguous()
    mask = mask_b.to(torch.uint8).reshape(-1)

    st.mark("fields")
    # the fields pass on the main stream, and beside it the two trajectory
    # directions (1000 seeds each) on the side stream: the fields pass needs
    # their statistics only at the finishing stage.  Every output is
    # allocated before the first launch and stays referenced until the side
    # stream has joined (a tensor freed while its kernel runs could otherwise
    # be handed to another output by the caching allocator).
    main = torch.cuda.current_stream(dev)
    side = hip.side_stream(dev)
    info.n_seed_points = np.uint32(n_t)
    ptraj = [kernel_params(info, sign, "INTEGRATE_TRAJECTORY", fp_contract) for sign in (1, -1)]
    outs = [(torch.empty((n_t, int(p.max_n_steps), 2), dtype=torch.int8, device=dev),
             torch.empty(n_t, dtype=torch.int16, device=dev),
             torch.empty(n_t, dtype=torch.float32, device=dev)) for p in ptraj]
    nxy = nxp * nyp
    slc = torch.empty((2, nxy), dtype=torch.int32, device=dev)
    slt = torch.empty((2, nxy), dtype=torch.int32, device=dev)
    steps = torch.zeros(1, dtype=torch.int64, device=dev)
    side.wait_stream(main)
    for p, (traj, ns, ln) in zip(ptraj, outs):
        hip.trace_paths(p, ts, mask, uv.reshape(-1), traj, ns, ln, stream=side)
    info.n_seed_points = np.uint32(job.seeds.shape[0])
    # both directions in one persistent launch (no end-of-launch tail between them)
    jobs = hip.fields_jobs([(kernel_params(info, sign, "INTEGRATE_FIELDS", fp_contract), seeds, mask,
                             uv.reshape(-1), slc[k], slt[k]) for k, sign in enumerate((1, -1))])
    hip.trace_fields_batch(jobs, steps)
    with torch.cuda.stream(side):
        tl = torch.stack([o[2] for o in outs], 1).cpu().numpy()
        tn = torch.stack([o[1] for o in outs], 1).cpu().numpy().view(np.uint16)
    main.wait_stream(side)
    del outs
    stats = compute_stats(tl, tn, info.pixel_size)

    st.mark("finish")
    slt_f, sla_f = hip.fields_finish_raw(slc, slt, int(info.subpixel_seed_point_density))
    ps = np.float32(info.pixel_size)
    sd = np.float32(np.float32(stats["ds"]["downstream", "mean"]) / ps)
    su = np.float32(np.float32(stats["ds"]["upstream", "mean"]) / ps)
    slc_f = hip.fields_finish_scaled(slc, slt_f, int(info.subpixel_seed_point_density), sd, su)

    st.mark("mapping")
    slt_dn = slt_f[:, 0].reshape(nxp, nyp).contiguous()
    slc_dn = slc_f[:, 0].reshape(nxp, nyp).contiguous()
    ok = ~mask_b
    thr = float(torch.quantile(slt_dn[ok][:: max(1, int(ok.sum().item()) // 1000000)].double(),
                               channel_quantile).item())
    info.set_thresholds(channel_threshold=np.float32(thr), segmentation_threshold=np.uint32(segmentation_threshold))
    mapping0 = torch.zeros((nxp, nyp), dtype=torch.int32, device=dev)
    if midslopes:
        hip.map_midslopes_ridges(sla_f.reshape(nxp, nyp, 2), mask_b, mapping0, int(info.is_midslope),
                                 int(info.is_ridge), midslope_filter_sigma, midslope_threshold, ridge_filter_sigma,
                                 ridge_threshold)
    cm = ChannelMapper(info, mask_b, uv, mapping0, slt=slt_dn, slc=slc_dn, fp_contract=fp_contract)
    ok_chain = cm.map_channels_segments()

    def traced_lengths():
        # lengths.hsl (lengths.py): trace down from the mid-slope pixels, or
        # without a mid-slope map from every pixel that is not a thin channel;
        # then the Gaussian KDE of their logarithms
        if midslopes:
            hs = pick_seeds(mask=mask_b, map=cm.mapping, flag=info.is_midslope, pad=pad)
        else:
            hs = pick_seeds(mask=mask_b, map=~cm.mapping, flag=info.is_thinchannel, pad=pad)
        info.n_seed_points = np.uint32(hs.shape[0])
        p = kernel_params(info, 1, "HILLSLOPE_LENGTHS", fp_contract)
        length = torch.zeros(hs.shape[0], dtype=torch.float32, device=dev)
        if hs.shape[0]:
            hip.hillslope_lengths(p, hs, mask, uv.reshape(-1), cm.mapping.reshape(-1), int(info.is_thinchannel),
                                  int(info.is_midslope), length)
        return hs, length

    def kde(length):
        reached = length[length > 0]
        d = None
        if reached.numel() > 1:
            logx = torch.log(reached).cpu().numpy().astype(np.float32)
            d = Univariate_distribution(logx_array=logx, logy_array=logx, n_hist_bins=512, n_pdf_points=256)
            d.compute_kde_opencl(kernel="gaussian", bandwidth=0.15, fp_contract=fp_contract)
        return reached, d

    fmm_res = fmm_grid = fmm_trace_mean = ha_stats = None
    if not fmm:
        st.mark("hsl")
        hs_seeds, hsl = traced_lengths()
        st.mark("kde")
        pos, dist = kde(hsl)
    else:
        # the reference's live HSL path (mapping.py:490-494): FMM lengths per
        # subsegment, remapped onto the grid.  The march of the largest
        # subsegments holds a few CUs for most of the stage, so the traced
        # lengths, their KDE and the slope and aspect maps of the aspect
        # stage (none of which reads the FMM result) run beside it on the
        # side stream.
        st.mark("fmm")
        # the side stream joins the mapping outputs before the solve is queued
        # on the main stream, so it does not wait for the solve
        side.wait_stream(main)
        pend = fmm_hsl.fmm_start(cm.label, cm.mapping, mask_b, int(info.is_thinchannel), job.pixel_size,
                                 stream=main)
        pad_w = int(info.pad_width)
        with torch.cuda.stream(side):
            hs_seeds, hsl = traced_lengths()
            pos, dist = kde(hsl)
            slope_full = fmm_hsl.slope_deg(roi, pad_w, job.pixel_size)
            slope_c = slope_full[ox:ox + nxp, oy:oy + nyp].contiguous()
            aspect, aspect_mask = fmm_hsl.map_aspect(slope_c, uv, cm.mapping, int(info.is_channel), 4.0, 50.0,
                                                     job.pixel_size)
        fmm_res = fmm_hsl.fmm_finish(pend)
        del pend
        fmm_grid = fmm_hsl.remap_hsl(cm.label, fmm_res)
        fmm_trace_mean = fmm_hsl.trace_mean_hsl(fmm_grid)
        # map_hsl, map_aspect, combine_hsl_aspect, hsl_aspect_stats (mapping.py:987-1251)
        hsl_smoothed = fmm_hsl.map_hsl(fmm_grid, mask_b, 51, 30, job.pixel_size)
        main.wait_stream(side)
        st.mark("aspect")
        df_ha, _, haa = fmm_hsl.combine_hsl_aspect(aspect, aspect_mask, hsl_smoothed, mask_b, pad_w)
        ha_stats = fmm_hsl.hsl_aspect_stats(haa, df_ha)
    st.mark("end")
    timings = st.report()
    results = {
        "steps": int(steps.item()),
        "channel_threshold": thr,
        "chain_ok": bool(ok_chain),
        "n_segments": int(cm.n_segments),
        "channel_pixels": int(((cm.mapping & int(info.is_thinchannel)) != 0).sum().item()),
        "hsl_seeds": int(hs_seeds.shape[0]),
        "hsl_reached": int(pos.numel()),
        "hsl_median_m": float(pos.median().item()) if pos.numel() else None,
        "kde_mode_x": None,
    }
    if dist is not None:
        dist.find_modes()
        results["kde_mode_x"] = float(np.asarray(dist.mode_x).reshape(-1)[0])
    if fmm_res is not None:
        results["fmm_subsegments"] = int(fmm_res.selected.size)
        results["fmm_kept"] = int((fmm_res.hsl > 0).sum())
        results["fmm_colours"] = int(fmm_res.n_colours)
        results["fmm_mean_hsl_m"] = fmm_res.fmm_mean_hsl
        results["fmm_trace_mean_hsl_m"] = None if fmm_trace_mean is None else float(fmm_trace_mean)
        results["hsl_aspect"] = {k: ha_stats[k] for k in ("hsl_mean", "hsl_mean_north", "hsl_mean_south",
                                                           "hsl_ns_disparity", "hsl_mean_azimuth",
                                                           "hsl_ns_disparity_confidence") if k in ha_stats}
    results["midslope_pixels"] = int(((cm.mapping & int(info.is_midslope)) != 0).sum().item())
    results["ridge_pixels"] = int(((cm.mapping & int(info.is_ridge)) != 0).sum().item())
    if return_grids:
        results["grids"] = {"uv": uv, "mask": mask_b, "sla": sla_f, "slt": slt_f, "slc": slc_f,
                            "mapping": cm.mapping, "label": cm.label, "fmm_hsl": fmm_grid,
                            "traj_length": tl, "traj_nsteps": tn}
    return results, timings
