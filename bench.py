"""Benchmark: jittered slc/slt field mapping of a 2048x2048 DTM on one MI355X.

One step = the full fields pass of the tracing path: integrate_fields
downstream and upstream over every seed (2x2 sub-pixel seeds per pixel,
jitter 1.0), i.e. both slc/slt maps of the 2048^2 tile at 2 m.  Inputs (uv,
mask, seeds) are resident in HBM before timing starts; each step sorts the
seeds into tile order, clears and re-accumulates the fields.

metric: streamline-steps/s = accepted RK2 steps of all streamlines, both
directions, counted exactly by the kernel, divided by wall time of the timed
steps.  ms_per_step is the 2048^2 slc/slt map time.

The default path traces both directions in one persistent launch
(slm_trace_fields_batch, one job per direction); --path queue and --path
tiled time one launch per direction.

roofline: dominant kernel k_fields_batch, algorithmic 73 B per step (64 B of
uv gathers, 1 B mask, 8 B of slc/slt atomics) x steps per launch / mean
launch time from HIP events on the launch stream (the events bracket the
whole call: seed sort, uv quads and commit kernels included); peak 8 TB/s HBM.
cpu_baseline: the CPU oracle (C restatement of the reference kernels, OpenMP)
on a bounded random sample of the same seeds, steps/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_STEP = 73
HBM_PEAK_GBS = 8000.0
CALIB_ITERS = 8192          # slm_calibrate_valu rounds: 16 v_fma_f32 each per wave
CALIB_BLOCKS_PER_CU = 8     # 32 waves per CU, 8 per SIMD
VALU_PEAK_KINDS = (0, 2, 8)  # v_fma_f32, v_add_u32, v_mov_b32


def _profile_values(path):
    """'key: value ...' lines of a committed profile summary, as floats."""
    vals = {}
    if not os.path.exists(path):
        return vals
    for line in open(path):
        if line.startswith("#") or ":" not in line:
            continue
        k, v = line.split(":", 1)
        try:
            vals[k.strip()] = float(v.split()[0])
        except (ValueError, IndexError):
            pass
    return vals


def valu_peak(stream=None, reps=5, kind=0):
    """Sustained wave64 VALU issue rate of the whole chip (wave-instructions
    per second): slm_calibrate_valu, whose waves issue a known number of
    independent instructions of one kind (0 v_fma_f32, 1 v_add_f32,
    2 v_add_u32, 3 v_xor_b32, 4 v_mul_f32, 5 the two alternating), timed with HIP events on its
    stream."""
    import torch
    from slm_amd import hip
    dev = torch.device("cuda", torch.cuda.current_device())
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    n_blocks = cus * CALIB_BLOCKS_PER_CU
    out = torch.zeros(n_blocks, dtype=torch.float32, device=dev)
    stream = stream or torch.cuda.current_stream(dev)
    hip.calibrate_valu(n_blocks, CALIB_ITERS, out, kind=kind, stream=stream, sync=True)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(stream)
        hip.calibrate_valu(n_blocks, CALIB_ITERS, out, kind=kind, stream=stream)
        b.record(stream)
    torch.cuda.synchronize(dev)
    ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
    instr = n_blocks * 4 * 16 * CALIB_ITERS
    return instr / (ms / 1e3), {"kind": kind, "blocks": n_blocks, "iters": CALIB_ITERS, "wave_instructions": instr,
                                "ms": ms}


def _cpu_info():
    """The host the CPU baseline runs on: logical CPUs (nproc), the affinity
    mask, the thread share the environment grants (OMP_NUM_THREADS, the
    process's CPU share on a shared GPU host), the CPU model (lscpu) and
    whether an OpenCL CPU device (pocl) is installed, which would let the
    reference's own kernels run there."""
    import subprocess
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except Exception:
        info["affinity"] = info["nproc"]
    omp = os.environ.get("OMP_NUM_THREADS", "")
    info["omp_num_threads"] = int(omp) if omp.isdigit() and int(omp) > 0 else None
    info["model"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                info["model"] = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    icds = []
    for d in ("/etc/OpenCL/vendors",):
        try:
            icds += sorted(f for f in os.listdir(d) if f.endswith(".icd"))
        except OSError:
            pass
    info["opencl_icds"] = icds
    cl = None
    try:
        r = subprocess.run(["clinfo", "-l"], capture_output=True, text=True, timeout=30)
        cl = r.stdout
    except Exception:
        pass
    info["pocl_device"] = bool(cl and "portable computing language" in cl.lower()) or \
        any("pocl" in f.lower() for f in icds)
    info["clinfo"] = " | ".join(l.strip() for l in cl.splitlines() if l.strip())[:300] if cl else None
    return info


def _cores(info):
    """Threads for the CPU baseline: every CPU of the affinity mask, within
    the OMP_NUM_THREADS share when the environment sets one."""
    n = info["affinity"]
    if info["omp_num_threads"]:
        n = min(n, info["omp_num_threads"])
    return max(1, n)


def cpu_baseline(job, target_s=12.0):
    """Oracle steps/s on the first n seeds of the shuffled seed list (a uniform
    random sample of the map, RNG keys as in the full run), both directions;
    n grows until one measurement takes a few seconds."""
    from oracle import oracle
    host = _cpu_info()
    cores = _cores(host)
    m = job.mask.astype(np.uint8).reshape(-1)
    uv = job.uv.reshape(-1)
    n_total = job.seeds.shape[0]
    n = 4096
    while True:
        n = min(n, n_total)
        steps = 0
        t0 = time.perf_counter()
        for sign in (1, -1):
            p = job.params(sign, n_seed_points=n)
            steps += oracle.trace_fields(p, job.seeds, m, uv, nthreads=cores)[2]
        dt = time.perf_counter() - t0
        if dt > target_s / 4 or n >= n_total:
            break
        n *= 4
    return {"value": steps / dt, "unit": "streamline-steps/s", "cores": cores, "kind": "port",
            "sample": "first %d of %d shuffled seeds (x%d sub-seeds, both directions): %d steps in "
                      "%.1f s; C restatement of the reference kernels, OpenMP, %d threads" %
                      (n, n_total, int(job.info.subpixel_seed_point_density) ** 2, steps, dt, cores),
            "host": host,
            "reference_kernels": "not run: no OpenCL CPU device (pocl) on this host" if not host["pocl_device"]
            else "pocl present; the reference's .cl sources do not travel with the build"}


def tracer_roofline(args, kname, steps_per_launch, mean_launch_s, alg_gbs, stream, profile=None):
    """Roofline of the fields tracer.  Two resources are compared, each from
    the kernel's mean launch time measured here:
      - VALU issue: a SIMD issues VALU work in quad-cycle slots, one
        instruction or a dual-issued pair per slot.  The slots one launch
        takes (SQ_INSTS_VALU minus SQ_ACTIVE_INST_VALU2 per launch, from the
        committed PMC summary, scaled by what SQ_INSTS_VALU reads per issued
        instruction on the calibration kernel) per second, against the slot
        rate of the v_fma_f32 calibration kernel (never dual-issued, so one
        instruction per slot), measured now; the instruction rate against
        the fastest calibrated kind is kept beside it;
      - HBM: measured DRAM-side bytes per launch (FETCH_SIZE + WRITE_SIZE,
        copy-calibrated, from the same summary) per second, against 8 TB/s.
    The bound is the resource with the larger fraction.  The algorithmic
    bytes (73 B per step: 64 B of uv gathers, 1 B mask, 8 B of slc/slt
    atomics) are reported beside them: they mostly hit L2, so their rate is
    not an HBM rate.  profile: the PMC summary under profiles/ (default: the
    kernel's tile_2048 one for that workload); when it records the steps of
    its launches and they differ from these (other options), its per-launch
    counts are scaled per step."""
    prof = {}
    if profile is None and args.workload == "tile_2048":
        profile = kname + "_tile_2048_traffic.txt"
    if profile:
        prof = _profile_values(os.path.join(ROOT, "profiles", profile))
    scale = steps_per_launch / prof["steps_per_launch"] if prof.get("steps_per_launch") else 1.0
    prof = {k: (v * scale if k.endswith("_per_launch") else v) for k, v in prof.items()}
    cal = _profile_values(os.path.join(ROOT, "profiles", "valu_calibration.txt"))
    traffic = prof.get("bytes_per_launch")
    hbm = {"peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic": alg_gbs,
           "algorithmic_bytes_per_step": BYTES_PER_STEP, "algorithmic_frac": alg_gbs / HBM_PEAK_GBS,
           "measured": None, "frac": None}
    if traffic:
        hbm["measured"] = traffic / mean_launch_s / 1e9
        hbm["frac"] = hbm["measured"] / HBM_PEAK_GBS
    out = {"kernel": kname, "path": args.path, "steps_per_launch": steps_per_launch,
           "mean_launch_ms": mean_launch_s * 1e3, "traffic": traffic, "hbm": hbm,
           "l2_hit_rate": prof.get("TCC_hit_rate")}
    valu = prof.get("SQ_INSTS_VALU_per_launch")
    dual = prof.get("SQ_ACTIVE_INST_VALU2_per_launch")
    per_instr = cal.get("SQ_INSTS_VALU_per_issued_instruction")
    if valu and dual is not None and per_instr:
        instr = valu / per_instr
        slots = instr - dual / per_instr
        fma, meta = valu_peak(stream, kind=0)
        # the fastest sustained instruction rate of any calibrated kind
        # (v_mov_b32, which pairs), for the instruction view
        ipeak, _ = max((valu_peak(stream, kind=k) for k in VALU_PEAK_KINDS), key=lambda r: r[0])
        rate = slots / mean_launch_s
        irate = instr / mean_launch_s
        out["valu"] = {"achieved": rate / 1e9, "peak": fma / 1e9, "unit": "G VALU issue slots/s",
                       "frac": rate / fma, "slots_per_launch": slots, "instructions_per_launch": instr,
                       "dual_issued_pairs_per_launch": dual / per_instr,
                       "instructions_per_step": instr / steps_per_launch,
                       "instruction_rate": irate / 1e9, "fastest_kind_instruction_rate": ipeak / 1e9,
                       "instruction_frac": irate / ipeak, "calibration": meta}
    if "valu" in out and (hbm["frac"] is None or out["valu"]["frac"] >= hbm["frac"]):
        v = out["valu"]
        out.update({"bound": "valu", "achieved": v["achieved"], "peak": v["peak"], "unit": v["unit"],
                    "frac": v["frac"]})
    elif hbm["frac"] is not None:
        out.update({"bound": "hbm", "achieved": hbm["measured"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm["frac"]})
    else:
        # no PMC summary for this workload: only the algorithmic rate is known
        out.update({"bound": "hbm", "achieved": alg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": alg_gbs / HBM_PEAK_GBS, "note": "algorithmic bytes; no measured counters"})
    return out


def bench_batch(args):
    """One step = every tile of the batched workload traced in both directions by
    one slm_trace_fields_batch call (seed sort included), replayed from a hipGraph."""
    import torch
    from slm_amd import autotune, hip, workloads
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    tiles = workloads.batch_c3()
    gen_s = time.perf_counter() - t0
    jobs = []
    for t in tiles:
        seeds = torch.from_numpy(t.seeds).to(dev)
        mask = torch.from_numpy(t.mask.astype(np.uint8).reshape(-1)).to(dev)
        uv = torch.from_numpy(np.ascontiguousarray(t.uv).reshape(-1)).to(dev)
        for sign in (1, -1):
            p = t.params(sign)
            slc = torch.empty(p.nxy_padded, dtype=torch.int32, device=dev)
            jobs.append((p, seeds, mask, uv, slc, torch.empty_like(slc)))
    arr = hip.fields_jobs(jobs)
    ws = torch.empty(hip.batch_workspace_size(arr), dtype=torch.uint8, device=dev)
    steps = torch.zeros(1, dtype=torch.int64, device=dev)
    hip.trace_fields_batch(arr, steps=steps, workspace=ws, sync=True)
    default_ms = autotune.time_batch(arr, hip.batch_opts(), ws, repeats=2)
    table = []
    if args.no_tune:
        opts = hip.batch_opts()
    else:
        opts, table = autotune.tune_batch(arr, workspace=ws, repeats=2)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            hip.trace_fields_batch(arr, steps=steps, workspace=ws, opts=opts)
    torch.cuda.current_stream().wait_stream(side)
    for _ in range(args.warmup):
        g.replay()
    torch.cuda.synchronize()
    steps.zero_()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record()
        g.replay()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_steps = int(steps.item())
    mean_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    steps_per_call = total_steps / args.steps
    # the tracer's launch time: the call with the seed order and quads of the
    # graph's last replay kept (no sort or quad build; the queue set-up and
    # the commit into the jobs' outputs remain, under 1% of it)
    launch_ms = autotune.time_batch(arr, opts, ws, repeats=3)
    achieved = steps_per_call * BYTES_PER_STEP / (launch_ms / 1e3) / 1e9
    roofline = tracer_roofline(args, "k_fields_batch", steps_per_call, launch_ms / 1e3, achieved,
                               torch.cuda.current_stream(), profile="k_fields_batch_c3_traffic.txt")
    out = {
        "metric": "streamline-steps/s",
        "value": total_steps / wall,
        "unit": "streamline-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (64 seeded DTM tiles, 16 fans with unfilled pits)",
        "config": {"workload": "batch_c3", "tiles": len(tiles), "jobs": len(jobs),
                   "sides": [list(t.roi.shape) for t in tiles],
                   "pixels": int(sum(t.mask.size for t in tiles)), "seeds": int(sum(t.seeds.shape[0] for t in tiles)),
                   "max_n_steps": int(tiles[0].info.max_n_steps), "subpixel_density": 1, "jitter": 1.0,
                   "directions": 2, "host_generation_s": round(gen_s, 1)},
        "batch": {"graph": True, "opts": {"chunk": opts.chunk, "blocks_per_cu": opts.blocks_per_cu,
                                          "queue_slices": opts.queue_slices, "lds_slots": opts.lds_slots,
                                          "refill_batch": opts.refill_batch},
                  "default_opts_ms": round(default_ms, 3), "graph_replay_ms": round(mean_ms, 3),
                  "autotune_table": [list(r) for r in table]},
        "roofline": roofline,
    }
    print(json.dumps(out), flush=True)


def bench_surrogate(args):
    """Surrogate training on traced tiles, then the surrogate-gated map of a
    large DTM against plain tracing of the same DTM.  One step = one gated
    map (inference on overlapping 256^2 tiles, 2% of the tiles re-traced and
    checked, failures replaced); the plain map (every pixel traced) is timed
    the same way.  Both start from the DTM's uv, mask and h resident in HBM."""
    import torch
    from slm_amd import gated
    from slm_amd import surrogate as S
    dev = torch.device("cuda", 0)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    gated.TRACE["max_length"] = float(args.max_length)
    if args.weights and os.path.exists(args.weights):
        z = np.load(args.weights)
        params = {name: (z[name + ".w"], z[name + ".b"]) for name, *_ in S.LAYERS}
        theta = float(z["theta"])
        train_rep = {"loaded": args.weights}
    else:
        data = gated.training_set(n_dtms=args.train_dtms, n=2048, size=128, log=log, device=dev)
        const = gated.constant_loss(data, w_len=args.w_len, w_bce=args.w_bce)
        net, hist = gated.train(data, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log, device=dev,
                                w_len=args.w_len, w_bce=args.w_bce, cosine=not args.no_cosine)
        cal = gated.calibrate(net, data)
        log("channel logit threshold %.2f: F1 %.4f (at 0: %.4f)" % cal)
        train_scores = gated.tile_scores(net, data, n_tiles=1024)
        hold = gated.training_set(n_dtms=2, n=2048, size=128, seed0=90000, theta=data.theta, device=dev)
        hold_scores = gated.tile_scores(net, hold)
        del hold
        log("tiles: train %s, held-out DTMs %s" % (train_scores, hold_scores))
        params, theta = net.params_numpy(), data.theta
        train_rep = {"tiles": data.n_tiles, "tile": data.size, "epochs": args.epochs, "batch": args.batch,
                     "lr": args.lr, "lr_schedule": "constant" if args.no_cosine else "warmup 200 + cosine",
                     "loss_weights": {"length": args.w_len, "density": 1.0, "channel": args.w_bce},
                     "theta": theta, "constant_loss": const, "loss_first_epoch": hist["loss"][0],
                     "loss_last_epoch": hist["loss"][-1], "train_s": hist["train_ms"] / 1e3,
                     "tiles_per_s": hist["tiles_per_s"], "train_tflops": hist["tflops"],
                     "data_generate_s": data.timings["generate_s"], "data_trace_s": data.timings["trace_s"],
                     "channel_logit_threshold": cal[0], "train_channel_f1": cal[1],
                     "train_tiles": train_scores, "heldout_tiles": hold_scores,
                     "train_channel_f1_uncalibrated": cal[2]}
        del data, net
        if args.save_weights:
            np.savez(args.save_weights, theta=theta,
                     **{k + ".w": w for k, (w, b) in params.items()}, **{k + ".b": b for k, (w, b) in params.items()})
    torch.cuda.empty_cache()
    inf = S.UNet(params, device=dev, inference=True)
    t0 = time.perf_counter()
    roi = gated.family_dtm(args.map_n, args.map_seed)
    gen_s = time.perf_counter() - t0
    grid = gated.DeviceGrid(roi, device=dev)
    log("map grid %d x %d, generated in %.1f s" % (grid.nx, grid.ny, gen_s))

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        r = fn()
        b.record()
        torch.cuda.synchronize()
        return r, a.elapsed_time(b)

    steps = torch.zeros(1, dtype=torch.int64, device=dev)
    plain, _ = gated.plain_map(grid, theta, steps=steps)            # warm-up (and the reference fields)
    plain_ms = []
    for _ in range(args.steps):
        grid._seeds = None                                          # seed list built inside the timed map
        (plain, _), ms = timed(lambda: gated.plain_map(grid, theta))
        plain_ms.append(ms)
    plain = plain.clone()
    # the metric's noise floor: a second plain map whose seeds draw other jitter
    grid._seeds = grid.seeds().flip(0).contiguous()
    plain2, _ = gated.plain_map(grid, theta)
    floor = gated.compare(plain2, plain, grid.mask, theta, (grid.nx, grid.ny))
    grid._seeds = None
    size, halo = args.map_tile, args.map_halo
    batch = max(1, (64 * 256 * 256) // (size * size))
    gkw = dict(size=size, halo=halo, batch=batch, check_frac=0.02, rng_seed=2, max_err=0.05, min_f1=0.95)
    ckw = {k: v for k, v in gkw.items() if k != "batch"}
    # the gate's checks passed to a trace of other jitter: the share of checked
    # tiles that any estimate as close to the truth as the tracer itself passes
    floor_checks = gated.check_tiles(plain2, grid, theta, **ckw)
    del plain2
    for _ in range(args.warmup):
        gated.gated_map(inf, grid, theta, **gkw)
    gated_ms = []
    for _ in range(args.steps):
        (res, ms) = timed(lambda: gated.gated_map(inf, grid, theta, **gkw))
        gated_ms.append(ms)
    out, rep = res
    (pred, _), pred_ms = timed(lambda: gated.predict(inf, grid, size, halo, batch))
    fwd = []
    for _ in range(3):
        _, ms = timed(lambda: inf.forward(inf.A["x"], batch, size, size))
        fwd.append(ms)
    fwd_ms = float(np.median(fwd))
    fwd_tflops = inf.flops_per_pixel() * batch * size * size / (fwd_ms / 1e3) / 1e12
    net_checks = gated.check_tiles(pred, grid, theta, **ckw)       # every checked tile, no early stop
    pass_rate = lambda cs: sum(c["ok"] for c in cs) / max(1, len(cs))
    m_gated = gated.compare(out, plain, grid.mask, theta, (grid.nx, grid.ny))
    m_pred = gated.compare(pred, plain, grid.mask, theta, (grid.nx, grid.ny))
    tp, tg = float(np.median(plain_ms)), float(np.median(gated_ms))
    peak = 2500.0
    res = {
        "metric": "surrogate-gated vs plain mapping speedup",
        "value": tp / tg,
        "unit": "x",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tg,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded DTMs, surrogate trained on their traces; random-init weights trained here)",
        "config": {"workload": "surrogate_gated_%d" % args.map_n, "map_px": grid.nx * grid.ny,
                   "tile": size, "halo": halo, "check_frac": 0.02, "check_rng_seed": 2,
                   "tolerance": {"median_rel_err": 0.05, "channel_f1": 0.95},
                   "trace": gated.TRACE, "directions": 1, "train_dtms": args.train_dtms},
        "surrogate": {"train": train_rep,
                      "map": {"plain_ms": tp, "gated_ms": tg, "predict_only_ms": pred_ms,
                              "plain_steps": int(steps.item()), "host_generation_s": round(gen_s, 1),
                              "gated_vs_plain": m_gated, "prediction_vs_plain": m_pred,
                              "plain_vs_plain_other_jitter": floor,
                              "tiles": rep["tiles"], "checked": rep["checked"], "evaluated": rep["evaluated"],
                              "failed": rep["failed"],
                              "fallback": rep["fallback"], "check_steps": rep["check_steps"],
                              "gate_pass_rate": {"network": pass_rate(net_checks),
                                                 "trace_other_jitter": pass_rate(floor_checks),
                                                 "tiles_checked": len(net_checks)},
                              "checks_other_jitter": [{k: c[k] for k in ("median_rel_err", "f1", "channel_px", "ok")}
                                                      for c in floor_checks],
                              "checks": [{k: c[k] for k in ("median_rel_err", "f1", "channel_px", "ok")}
                                         for c in net_checks]}},
        "roofline": {"bound": "mfma", "achieved": fwd_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": fwd_tflops / peak, "traffic": None, "kernel": "network forward (k_conv)",
                     "flop_per_px": inf.flops_per_pixel(), "px_per_launch": batch * size * size,
                     "mean_forward_ms": fwd_ms},
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tile_2048", choices=["tile_2048", "small_256", "batch_c3", "surrogate"],
                    help="batch_c3: 64 uneven tiles (log-uniform 256..1536 px, 1 or 2 m, 16 fans with pits, "
                         "MAX_N_STEPS 5000), both directions, one persistent launch captured in a hipGraph "
                         "with autotuned queue parameters")
    ap.add_argument("--no-tune", action="store_true", help="batch_c3: skip the autotuner (library defaults)")
    ap.add_argument("--train-dtms", type=int, default=32, help="surrogate: traced 2048^2 DTMs for training")
    ap.add_argument("--epochs", type=int, default=20, help="surrogate: training epochs")
    ap.add_argument("--batch", type=int, default=32, help="surrogate: training batch (tiles)")
    ap.add_argument("--lr", type=float, default=1e-3, help="surrogate: Adam peak learning rate")
    ap.add_argument("--no-cosine", action="store_true", help="surrogate: constant learning rate after warm-up")
    ap.add_argument("--w-len", type=float, default=0.25, help="surrogate: loss weight of log1p(slt)")
    ap.add_argument("--w-bce", type=float, default=5.0, help="surrogate: loss weight of the channel term")
    ap.add_argument("--max-length", type=float, default=100.0, help="surrogate: streamline cap (m) of all tracing")
    ap.add_argument("--map-n", type=int, default=8192, help="surrogate: side of the gated map's DTM")
    ap.add_argument("--map-tile", type=int, default=256, help="surrogate: inference tile side")
    ap.add_argument("--map-halo", type=int, default=16, help="surrogate: inference tile overlap per side")
    ap.add_argument("--map-seed", type=int, default=2, help="surrogate: seed of the gated map's DTM")
    ap.add_argument("--weights", default="", help="surrogate: load trained weights (.npz) instead of training")
    ap.add_argument("--save-weights", default="", help="surrogate: save the trained weights (.npz)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--path", default="batch", choices=["batch", "queue", "tiled"],
                    help="batch: both directions in one persistent launch (slm_trace_fields_batch); "
                         "queue: one slm_trace_fields launch per direction; "
                         "tiled: LDS-owned tiles (slm_trace_fields_tiled), one launch per direction")
    ap.add_argument("--tile", type=int, default=0, help="tile side for --path tiled (0: library default)")
    ap.add_argument("--block", type=int, default=0, help="threads per block for --path tiled")
    ap.add_argument("--unit", type=int, default=0, help="states per unit for --path tiled")
    ap.add_argument("--cap", type=int, default=0, help="steps per tile visit for --path tiled")
    ap.add_argument("--halo", type=int, default=0, help="hysteresis band for --path tiled")
    ap.add_argument("--pipeline", action="store_true",
                    help="also time the full tile pipeline stage by stage (uv, trajectories, fields, "
                         "finishing, channel mapping and segmentation, hillslope lengths, KDE)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (scales PMC byte counters)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the VALU calibration kernel four times (a known count of v_fma_f32 per "
                         "wave; calibrates the VALU instruction counter and the issue rate)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the tracing path is measured on a single MI355X "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from slm_amd import hip, workloads

    dev = torch.device("cuda", 0)
    if args.valu_calibration:
        for kind in range(9):
            rate, meta = valu_peak(reps=3, kind=kind)
            print(json.dumps({"valu_calibration": meta, "wave_instructions_per_s": rate}), flush=True)
        return
    if args.copy_calibration:
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}), flush=True)
        return
    if args.workload == "batch_c3":
        return bench_batch(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)
    job = getattr(workloads, args.workload)()
    seeds = torch.from_numpy(job.seeds).to(dev)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(dev)
    uv = torch.from_numpy(np.ascontiguousarray(job.uv).reshape(-1)).to(dev)
    pd, pu = job.params(1), job.params(-1)
    nxy = pd.nxy_padded
    slc = torch.empty((2, nxy), dtype=torch.int32, device=dev)
    slt = torch.empty((2, nxy), dtype=torch.int32, device=dev)
    steps = torch.zeros(2, dtype=torch.int64, device=dev)
    topts = hip.tile_opts(args.tile, args.block, args.unit, args.cap, args.halo)
    jobs = None
    if args.path == "tiled":
        ws = torch.empty(hip.tiled_workspace_size(pd, topts), dtype=torch.uint8, device=dev)
    elif args.path == "batch":
        jobs = hip.fields_jobs([(pd, seeds, mask, uv, slc[0], slt[0]), (pu, seeds, mask, uv, slc[1], slt[1])])
        ws = torch.empty(hip.batch_workspace_size(jobs), dtype=torch.uint8, device=dev)
    else:
        ws = torch.empty(hip.fields_workspace_size(pd), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    launches = 1 if args.path == "batch" else 2

    def one_step(keep_order, events=None):
        if args.path == "batch":
            # both directions (each job its own slc/slt) in one launch; the
            # seeds are sorted into tile order for both jobs inside the call
            if events is not None:
                events[0][0].record(stream)
            hip.trace_fields_batch(jobs, steps[0:1], workspace=ws, stream=stream, keep_order=keep_order)
            if events is not None:
                events[0][1].record(stream)
            return
        for i, p in enumerate((pd, pu)):
            if events is not None:
                events[i][0].record(stream)
            if args.path == "tiled":
                hip.trace_fields_tiled(p, seeds, mask, uv, slc[i], slt[i], steps[i:i + 1], workspace=ws,
                                       opts=topts, stream=stream)
            else:
                hip.trace_fields(p, seeds, mask, uv, slc[i], slt[i], steps[i:i + 1], workspace=ws,
                                 stream=stream, keep_order=keep_order or i == 1)
            if events is not None:
                events[i][1].record(stream)

    one_step(False)
    for _ in range(args.warmup):
        one_step(False)
    torch.cuda.synchronize()
    steps.zero_()
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(launches)]
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        one_step(False, ev[k])  # each step re-sorts the seeds (the per-direction paths reuse them upstream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_steps = int(steps.sum().item())
    launch_ms = [e[i][0].elapsed_time(e[i][1]) for e in ev for i in range(launches)]
    mean_launch_s = float(np.mean(launch_ms)) / 1e3
    steps_per_launch = total_steps / (launches * args.steps)
    achieved = steps_per_launch * BYTES_PER_STEP / mean_launch_s / 1e9
    value = total_steps / wall
    kname = {"batch": "k_fields_batch", "queue": "k_fields", "tiled": "k_tl_process"}[args.path]
    roofline = tracer_roofline(args, kname, steps_per_launch, mean_launch_s, achieved, stream)
    out = {
        "metric": "streamline-steps/s",
        "value": value,
        "unit": "streamline-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spectral DTM, V-valleys, sinks filled)",
        "config": {"workload": job.name, "grid": list(job.roi.shape),
                   "pixel_size_m": job.pixel_size, "seeds": int(job.seeds.shape[0]),
                   "subpixel_density": int(job.info.subpixel_seed_point_density),
                   "streamlines_per_direction": job.n_streamlines(),
                   "jitter": float(job.info.jitter_magnitude), "max_n_steps": int(job.info.max_n_steps),
                   "directions": 2, "map_ms": wall / args.steps * 1e3},
        "roofline": roofline,
    }
    if args.path == "tiled":
        roofline["kernel"] = "k_tl_process (with its per-round binning)"
        roofline["tile_opts"] = {"tile": topts.tile, "block": topts.block, "unit": topts.unit,
                                 "cap": topts.cap_steps, "halo": topts.halo}
    if args.pipeline:
        from slm_amd.pipeline import map_tile
        map_tile(job, fmm=True)  # warm-up
        res, tim = map_tile(job, fmm=True)
        # the loop and blockage conditioning of the same DTM (the reference's
        # do_condition), timed alone: the uv stage with it instead of the plain gradient
        roi_d = torch.from_numpy(np.ascontiguousarray(job.roi, dtype=np.float32)).to(dev)
        cws = torch.empty(int(hip.load().slm_condition_workspace_size(*job.roi.shape)), dtype=torch.uint8,
                          device=dev)
        hip.condition_uv(roi_d, int(job.info.pad_width), float(job.pixel_size), workspace=cws, sync=True)
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(5):
            hip.condition_uv(roi_d, int(job.info.pad_width), float(job.pixel_size), workspace=cws)
        c1.record()
        torch.cuda.synchronize()
        res["condition_uv_ms"] = round(c0.elapsed_time(c1) / 5, 3)
        out["pipeline"] = {"stages": {k: round(v["gpu_ms"], 3) for k, v in tim.items()},
                           "total_ms": round(sum(v["gpu_ms"] for v in tim.values()), 3), "results": res}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(job)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
