"""Surrogate-gated mapping: training data from the tracer, training on the
GPU, and the gated map of a large DTM.

The surrogate has no counterpart in the reference.  The plain tracer
(slm_trace_fields, downstream direction, jittered RK2, one seed per pixel) is
its only source of truth: it makes the training targets, it re-traces a random
sample of the tiles of every gated map to check the prediction, and it
replaces every checked tile that fails (or the whole map, when too many fail).

Fields compared (planar float32 [3][nx*ny] over the tracer's grid):
  0  log1p(slt / d^2)    streamline length sum per pixel
  1  log1p(slc / d^2)    streamline density (the "log-density")
  2  channel logit       traced: +1 where slc/d^2 >= theta else -1
                         predicted: the network's logit (channel where > 0)

  family_dtm      the synthetic DTMs of training and of the large map
  DeviceGrid      a DTM's tracer grids on the device (h, uv, mask, seeds)
  training_set    traced DTMs cut into tiles: inputs, targets, valid masks
  train           Adam on the GPU through the surrogate kernels
  predict         inference on overlapping tiles, stitched into the grid
  gated_map       predict, re-trace a random share of the tiles, replace
                  failures, fall back to plain_map when too many fail
  check_tiles     the gate's checks applied to any whole-map fields
  plain_map       the same fields by tracing every pixel
  compare         median relative log-density error and channel-pixel F1
"""
from __future__ import annotations

import copy
import ctypes
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import hip, synth
from . import surrogate as S
from .parameters import load_parameters, section
from .params import kernel_params
from .useful import Info

# tracing of the surrogate workloads: 1 m pixels, one seed per pixel, jitter
# +-0.5 px, RK2 step 0.5 px, streamlines capped at 100 m
TRACE = {"subpixel_seed_point_density": 1, "jitter_magnitude": 1.0, "max_length": 100.0,
         "integrator_step_factor": 0.5}
H_SCALE = 10.0          # metres per unit of the height input channel


# stationary noise scale: about 60 m of relief on a 2048^2 grid
NOISE_SCALE = 60.0 / 21800.0


def family_dtm(n: int, seed: int) -> np.ndarray:
    """n x n DTM at 1 m with the same local statistics at every n: stationary
    spectral noise (amplitude ~ k^-2, about 60 m of relief per 2048^2
    window), 5% tilt, V-valley segments 30 m deep, 61 px half width and
    2400 px long, centred uniformly over the grid at 3 per 2048^2, sinks
    filled, 2-px NaN border."""
    scale = n / 2048.0
    return synth.make_dtm(n, pixel_size=1.0, tilt=0.05, seed=seed, nan_border=2,
                          n_valleys=max(1, int(round(3 * scale * scale))), valley_depth=30.0,
                          valley_half_width_px=61.44, noise_scale=NOISE_SCALE, valley_half_length_px=1200.0,
                          valley_centre_range=(0.0, 1.0))


def _dev(device=None):
    import torch
    return device or torch.device("cuda", torch.cuda.current_device())


class DeviceGrid:
    """One DTM on the device, on the tracer's grid (the bounding box of valid
    pixels plus pad, as Data in useful.py:37-115): h (NaN = no data), uv,
    mask, and the seed list (every valid pixel, row-major)."""

    def __init__(self, roi, pixel_size=1.0, trace=None, pad=2, device=None):
        import torch
        self.torch = torch
        self.dev = _dev(device)
        params = load_parameters(overrides={"trace": dict(TRACE if trace is None else trace)})
        tr, state = section(params, "trace"), section(params, "state")
        r = roi if hasattr(roi, "data_ptr") else torch.from_numpy(np.ascontiguousarray(roi, np.float32))
        r = r.to(self.dev)
        uv_full, mask_full = hip.gradient_uv(r, pad, float(pixel_size))
        ok = ~mask_full
        rows = torch.nonzero(ok.any(1)).flatten()
        cols = torch.nonzero(ok.any(0)).flatten()
        x0, x1, y0, y1 = int(rows[0]), int(rows[-1]), int(cols[0]), int(cols[-1])
        bnx, bny = x1 + 1 - x0, y1 + 1 - y0
        ox, oy = x0 - pad, y0 - pad
        self.nx, self.ny = bnx + 2 * pad, bny + 2 * pad
        self.uv = uv_full[ox:ox + self.nx, oy:oy + self.ny].contiguous()
        self.mask = mask_full[ox:ox + self.nx, oy:oy + self.ny].to(torch.uint8).contiguous()
        hp = torch.nn.functional.pad(r[None], (pad, pad, pad, pad), value=float("nan"))[0]
        self.h = hp[ox:ox + self.nx, oy:oy + self.ny].contiguous()
        self.origin = (ox, oy)
        self.pad = pad
        self.info = Info(state, tr, pixel_size)
        self.info.set_xy(bnx, bny, pad)
        self.density = int(self.info.subpixel_seed_point_density)
        self.max_length_px = float(self.info.max_length)
        self._seeds = None

    @property
    def nxy(self):
        return self.nx * self.ny

    def seeds(self):
        if self._seeds is None:
            t = self.torch
            self._seeds = (t.nonzero(self.mask == 0).to(t.float32) - float(self.pad)).contiguous()
        return self._seeds

    def params(self, n_seeds, info=None, sign=1):
        info = info or self.info
        info.n_seed_points = np.uint32(n_seeds)
        return kernel_params(info, sign, "INTEGRATE_FIELDS", True)

    def trace_down(self, steps=None, workspace=None):
        """Raw downstream fields (int32 holding uint32) over the grid."""
        t = self.torch
        seeds = self.seeds()
        slc = t.empty(self.nxy, dtype=t.int32, device=self.dev)
        slt = t.empty_like(slc)
        ws = hip.trace_fields(self.params(seeds.shape[0]), seeds, self.mask.reshape(-1), self.uv.reshape(-1),
                              slc, slt, steps, workspace=workspace)
        return slc, slt, ws

    def region(self, r0, r1, c0, c1, n_seeds=None):
        """Job tuple (params, seeds, mask, uv, slc, slt) tracing the grid's
        window [r0, r1) x [c0, c1) as a grid of its own (streamlines stop at
        its edge); its outputs are fresh int32 buffers.  n_seeds: the
        window's valid pixels, when the caller has counted them (the seed
        list is then built without waiting for the device)."""
        t = self.torch
        uv = self.uv[r0:r1, c0:c1].contiguous()
        mask = self.mask[r0:r1, c0:c1].contiguous()
        valid = mask == 0
        idx = t.nonzero(valid) if n_seeds is None else t.nonzero_static(valid, size=int(n_seeds))
        seeds = (idx.to(t.float32) - float(self.pad)).contiguous()
        info = copy.copy(self.info)
        info.set_xy(r1 - r0 - 2 * self.pad, c1 - c0 - 2 * self.pad, self.pad)
        p = self.params(seeds.shape[0], info)
        slc = t.empty((r1 - r0) * (c1 - c0), dtype=t.int32, device=self.dev)
        slt = t.empty_like(slc)
        return (p, seeds, mask.reshape(-1), uv.reshape(-1), slc, slt)


def fields_log(slc, slt, mask, density, theta, out=None):
    """slm_nn_fields_log: raw traced fields -> planar [3, nxy] float32."""
    import torch
    nxy = slc.numel()
    if out is None:
        out = torch.empty((3, nxy), dtype=torch.float32, device=slc.device)
    hip.check(hip.load().slm_nn_fields_log(hip._ptr(slc), hip._ptr(slt), hip._ptr(mask), nxy, int(density),
                                           ctypes.c_float(theta), hip._ptr(out), hip._stream(None), 0))
    return out


def tile_inputs(grid, tiles, n_tiles, size, mean_ws, x):
    hip.check(hip.load().slm_nn_tile_inputs(hip._ptr(grid.h), hip._ptr(grid.uv), hip._ptr(grid.mask), grid.nx,
                                            grid.ny, hip._ptr(tiles), n_tiles, size, ctypes.c_float(H_SCALE),
                                            hip._ptr(mean_ws), hip._ptr(x), hip._stream(None), 0))


def tile_targets(grid, slc, slt, tiles, n_tiles, size, theta, target, valid):
    hip.check(hip.load().slm_nn_tile_targets(hip._ptr(slc), hip._ptr(slt), hip._ptr(grid.mask), grid.nx, grid.ny,
                                             hip._ptr(tiles), n_tiles, size, grid.density, ctypes.c_float(theta),
                                             hip._ptr(target), hip._ptr(valid), hip._stream(None), 0))


def tile_stitch(head, tiles, n_tiles, size, halo, nx, ny, out):
    hip.check(hip.load().slm_nn_tile_stitch(hip._ptr(head), S.OUT_CH, hip._ptr(tiles), n_tiles, size, halo, nx, ny,
                                            hip._ptr(out), hip._stream(None), 0))


# ------------------------------------------------------------------ training data

class TrainingSet:
    """Tiles of traced DTMs in HBM: x bf16 [n*S*S, 16], target float32
    [n*S*S, 3], valid uint8 [n*S*S]; n_valid per tile (host); theta = the
    channel density threshold (the theta_quantile of the pooled density)."""

    def __init__(self, x, target, valid, n_valid, size, theta, timings):
        self.x, self.target, self.valid = x, target, valid
        self.n_valid = n_valid
        self.size = size
        self.n_tiles = len(n_valid)
        self.theta = theta
        self.timings = timings


def training_set(n_dtms=32, n=2048, size=128, seed0=1000, theta_quantile=0.97, threads=8, device=None, log=None,
                 theta=None):
    """Trace n_dtms seeded n x n family DTMs (downstream) and cut each grid
    into (n/size)^2 tiles of size^2, as the training data of the surrogate.
    theta: the channel threshold (default: the theta_quantile of the pooled
    density)."""
    import torch
    dev = _dev(device)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        futures = [pool.submit(family_dtm, n, seed0 + k) for k in range(n_dtms)]
        rois = [f.result() for f in futures]
    t_gen = time.perf_counter() - t0
    tiles_np = np.array([(i * size, j * size) for i in range(n // size) for j in range(n // size)], np.int32)
    tiles = torch.from_numpy(tiles_np).to(dev)
    per = len(tiles_np)
    M = n_dtms * per * size * size
    x = torch.empty((M, S.IN_CH), dtype=torch.bfloat16, device=dev)
    target = torch.empty((M, 3), dtype=torch.float32, device=dev)
    valid = torch.empty(M, dtype=torch.uint8, device=dev)
    mean_ws = torch.empty(per, dtype=torch.float32, device=dev)
    raw, steps = [], torch.zeros(1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for k, roi in enumerate(rois):
        g = DeviceGrid(roi, device=dev)
        slc, slt, _ = g.trace_down(steps=steps)
        sl = slice(k * per * size * size, (k + 1) * per * size * size)
        tile_inputs(g, tiles, per, size, mean_ws, x[sl])
        raw.append((g, slc, slt))
    torch.cuda.synchronize()
    t_trace = time.perf_counter() - t1
    # theta: quantile of the pooled density over valid pixels (a strided sample)
    if theta is None:
        dens = torch.cat([slc[g.mask.reshape(-1) == 0].to(torch.float32) / g.density ** 2 for g, slc, _ in raw])
        theta = float(torch.quantile(dens[:: max(1, dens.numel() // 4000000)], theta_quantile).item())
        del dens
    for k, (g, slc, slt) in enumerate(raw):
        sl = slice(k * per * size * size, (k + 1) * per * size * size)
        tile_targets(g, slc, slt, tiles, per, size, theta, target[sl], valid[sl])
    n_valid = valid.view(n_dtms * per, size * size).sum(1, dtype=torch.int64).cpu().numpy()
    timings = {"generate_s": t_gen, "trace_s": t_trace, "steps": int(steps.item())}
    if log:
        log("training set: %d DTMs %d^2 -> %d tiles of %d^2, theta %.1f, generate %.1f s, trace %.2f s"
            % (n_dtms, n, n_dtms * per, size, theta, t_gen, t_trace))
    return TrainingSet(x, target, valid, n_valid, size, theta, timings)


def constant_loss(data, w_len=0.25, w_bce=5.0):
    """Loss per valid pixel of the best constant prediction (channel means for
    the squared errors, the prevalence logit for the cross-entropy), with the
    weights of train()."""
    import torch
    v = data.valid.bool()
    t = data.target[v].double()
    var = t[:, :2].var(0, unbiased=False).cpu().numpy()
    p = min(max(t[:, 2].mean().item(), 1e-9), 1 - 1e-9)
    return w_len * var[0] + var[1] - w_bce * (p * math.log(p) + (1 - p) * math.log(1 - p))


def train(data, epochs=20, batch=32, lr=1e-3, seed=0, device=None, log=None, w_len=0.25, w_bce=5.0,
          warmup=200, cosine=True):
    """Adam over shuffled batches of tiles; returns (net, history) where
    history holds the mean loss per valid pixel of each epoch (weighted:
    w_len on the length term, w_bce on the channel term) and timings.  The
    learning rate rises linearly to lr over warmup steps, then (cosine)
    decays to 0 along half a cosine over the remaining steps."""
    import torch
    dev = _dev(device)
    size = data.size
    params = S.init_params(seed)
    # head biases at the constant predictor: target means and the prevalence logit
    v = data.valid.bool()
    mu = data.target[v].double().mean(0).cpu().numpy()
    prev = min(max(float(mu[2]), 1e-6), 1 - 1e-6)
    params["head"][1][:3] = [mu[0], mu[1], math.log(prev / (1 - prev))]
    net = S.UNet(params, device=dev)
    net._alloc(batch, size, size)
    xv = data.x.view(data.n_tiles, size * size * S.IN_CH)
    tv = data.target.view(data.n_tiles, size * size * 3)
    vv = data.valid.view(data.n_tiles, size * size)
    tb = torch.empty((batch, size * size * 3), dtype=torch.float32, device=dev)
    vb = torch.empty((batch, size * size), dtype=torch.uint8, device=dev)
    xb = net.A["x"].view(batch, size * size * S.IN_CH)
    sums = torch.zeros((epochs, 3), dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed)
    n_batches = data.n_tiles // batch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    totals = []
    total_steps = epochs * n_batches
    step = 0
    ev[0].record()
    for ep in range(epochs):
        perm = rng.permutation(data.n_tiles)
        permd = torch.from_numpy(perm.astype(np.int64)).to(dev)
        tot = 0
        for b in range(n_batches):
            idx = permd[b * batch:(b + 1) * batch]
            torch.index_select(xv, 0, idx, out=xb)
            torch.index_select(tv, 0, idx, out=tb)
            torch.index_select(vv, 0, idx, out=vb)
            nv = int(data.n_valid[perm[b * batch:(b + 1) * batch]].sum())
            tot += nv
            net.forward(net.A["x"], batch, size, size)
            dh, _, _ = net.loss(tb.view(-1, 3), vb.view(-1), n_valid=nv, out=sums[ep], w_len=w_len, w_bce=w_bce)
            net.backward(dh)
            f = min(1.0, (step + 1) / max(1, warmup))
            if cosine and step >= warmup:
                f *= 0.5 * (1.0 + math.cos(math.pi * (step - warmup) / max(1, total_steps - warmup)))
            net.adam(lr=lr * f)
            step += 1
        totals.append(tot)
        if log:
            s = sums[ep].cpu().numpy()
            log("epoch %d: loss %.4f (sse %.4f %.4f, bce %.4f)" % (ep, (w_len * s[0] + s[1] + w_bce * s[2]) / tot,
                                                                   s[0] / tot, s[1] / tot, s[2] / tot))
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1])
    s = sums.cpu().numpy().astype(np.float64) * np.array([w_len, 1.0, w_bce])
    hist = {"loss": [float(s[e].sum() / totals[e]) for e in range(epochs)],
            "train_ms": ms, "steps": epochs * n_batches, "batch": batch,
            "tiles_per_s": epochs * n_batches * batch / (ms / 1e3),
            # forward + backward-data + backward-weight: three passes of the forward FLOP
            "tflops": 3 * net.flops_per_pixel() * size * size * batch * epochs * n_batches / (ms / 1e3) / 1e12}
    return net, hist


def calibrate(net, data, n_tiles=1024, batch=64):
    """Shift the channel logit so that 0 is the threshold with the best
    channel-pixel F1 on (up to) n_tiles training tiles: the head's channel
    bias absorbs the threshold.  Returns (threshold, F1 at it, F1 at 0)."""
    import torch
    size = data.size
    n = min(n_tiles, data.n_tiles) // batch * batch
    logits, truth = [], []
    xv = data.x.view(data.n_tiles, size * size * S.IN_CH)
    for b0 in range(0, n, batch):
        net._alloc(batch, size, size)
        net.A["x"].view(batch, -1).copy_(xv[b0:b0 + batch])
        head = net.forward(net.A["x"], batch, size, size)
        sl = slice(b0 * size * size, (b0 + batch) * size * size)
        v = data.valid[sl].bool()
        logits.append(head[:, 2].float()[v])
        truth.append(data.target[sl, 2][v] > 0.5)
    z, t = torch.cat(logits), torch.cat(truth)
    thr = torch.arange(-4.0, 4.0001, 0.05, device=z.device)
    pc = z[None, :] > thr[:, None] if z.numel() * thr.numel() < 2 ** 31 else None
    f1s = []
    for k in range(thr.numel()):
        p = pc[k] if pc is not None else z > thr[k]
        tp = (p & t).sum()
        f1s.append((2 * tp / (2 * tp + (p & ~t).sum() + (~p & t).sum()).clamp_min(1)).item())
    k = int(np.argmax(f1s))
    best = float(thr[k].item())
    f0 = f1s[int(np.argmin(np.abs(thr.cpu().numpy())))]
    _, _, ob, _ = net.offsets["head"]
    net.p[ob + 2] -= best
    net.pb[ob + 2] = net.p[ob + 2].to(net.pb.dtype)
    return best, f1s[k], f0


def tile_scores(net, data, n_tiles=None, batch=64):
    """Channel F1 (logit > 0) and median relative log-density error of the
    network on the tiles of a TrainingSet (e.g. held-out DTMs)."""
    import torch
    size = data.size
    n = (data.n_tiles if n_tiles is None else min(n_tiles, data.n_tiles)) // batch * batch
    xv = data.x.view(data.n_tiles, size * size * S.IN_CH)
    tp = fp = fn = 0
    errs = []
    for b0 in range(0, n, batch):
        net._alloc(batch, size, size)
        net.A["x"].view(batch, -1).copy_(xv[b0:b0 + batch])
        head = net.forward(net.A["x"], batch, size, size)
        sl = slice(b0 * size * size, (b0 + batch) * size * size)
        v = data.valid[sl].bool()
        pc, tc = head[:, 2].float()[v] > 0, data.target[sl, 2][v] > 0.5
        tp += int((pc & tc).sum().item())
        fp += int((pc & ~tc).sum().item())
        fn += int((~pc & tc).sum().item())
        t1 = data.target[sl, 1][v]
        errs.append(((head[:, 1].float()[v] - t1).abs() / t1.clamp_min(1e-6))[::16])
    return {"channel_f1": 2 * tp / max(1, 2 * tp + fp + fn),
            "median_rel_err_log_density": float(torch.cat(errs).median().item())}


# ------------------------------------------------------------------ maps

def tile_origins(nx, ny, size, halo):
    """Origins of overlapping size^2 tiles whose cores (size - 2*halo)^2
    partition the grid."""
    c = size - 2 * halo
    return np.array([(-halo + i * c, -halo + j * c) for i in range(int(math.ceil(nx / c)))
                     for j in range(int(math.ceil(ny / c)))], np.int32)


def predict_tiles(net, grid, origins, size, halo, batch, out):
    """Inference on the tiles at origins (int32 [T, 2]), their cores stitched
    into out [3, nxy], in batches of at most batch tiles (the last one as
    large as what is left).  Each tile's result depends on that tile alone,
    whatever batch it is in."""
    import torch
    dev = grid.dev
    if not len(origins):
        return out
    tiles = torch.from_numpy(np.ascontiguousarray(origins, np.int32)).to(dev)
    net._alloc(batch, size, size)
    mean_ws = torch.empty(batch, dtype=torch.float32, device=dev)
    for b0 in range(0, len(origins), batch):
        tl = tiles[b0:b0 + batch]
        nb = int(tl.shape[0])
        tile_inputs(grid, tl, nb, size, mean_ws, net.A["x"])
        head = net.forward(net.A["x"], nb, size, size)
        tile_stitch(head, tl, nb, size, halo, grid.nx, grid.ny, out)
    return out


def predict(net, grid, size=256, halo=16, batch=64, out=None):
    """Surrogate fields [3, nxy] of a grid, by inference on overlapping tiles."""
    import torch
    origins = tile_origins(grid.nx, grid.ny, size, halo)
    if out is None:
        out = torch.empty((3, grid.nxy), dtype=torch.float32, device=grid.dev)
    predict_tiles(net, grid, origins, size, halo, batch, out)
    return out, origins


def plain_map(grid, theta, steps=None, workspace=None):
    """The fields by tracing every pixel (downstream)."""
    slc, slt, ws = grid.trace_down(steps=steps, workspace=workspace)
    return fields_log(slc, slt, grid.mask.reshape(-1), grid.density, theta), ws


def _tile_stats(pred, truth, valid):
    """Device float64 [4] of one tile (views [3, a, b]): the median relative
    log-density error over the valid pixels, and the channel confusion counts
    tp, fp, fn."""
    import torch
    v = valid
    t1 = truth[1][v]
    err = ((pred[1][v] - t1).abs() / t1.clamp_min(1e-6))
    pc, tc = (pred[2] > 0) & v, (truth[2] > 0) & v
    med_t = err.median().double() if err.numel() else torch.zeros((), dtype=torch.float64, device=err.device)
    return torch.stack([med_t, (pc & tc).sum().double(), (pc & ~tc).sum().double(), (~pc & tc).sum().double()])


def _tile_verdict(got, max_err, min_f1, min_channel_px):
    """(ok, median error, F1, traced channel pixels) from _tile_stats' values."""
    med = float(np.float32(got[0]))
    tp, fp, fn = int(got[1]), int(got[2]), int(got[3])
    f1 = 2 * tp / max(1, 2 * tp + fp + fn) if (tp + fp + fn) else 1.0
    ok = med <= max_err and (f1 >= min_f1 or (tp + fn) < min_channel_px)
    return ok, med, f1, tp + fn



def _check_windows(grid, origins, chosen, size, halo):
    """Trace the core of each chosen tile as a grid of its own with a margin
    of max_length + 2 px (every streamline that can reach the core starts
    inside that window), all windows in one batched launch.  Returns (jobs,
    cores, steps): the job tuples, (i0, i1, j0, j1, r0, r1, c0, c1) per tile
    (core and window in grid pixels), and the device step count."""
    import torch
    margin = int(math.ceil(grid.max_length_px)) + 2
    c = size - 2 * halo
    cores = []
    for t in chosen:
        i0, j0 = int(origins[t][0]) + halo, int(origins[t][1]) + halo
        i1, j1 = min(grid.nx, i0 + c), min(grid.ny, j0 + c)
        r0, r1 = max(0, i0 - margin), min(grid.nx, i1 + margin)
        c0, c1 = max(0, j0 - margin), min(grid.ny, j1 + margin)
        cores.append((i0, i1, j0, j1, r0, r1, c0, c1))
    # the windows' seed counts in one copy, then their seed lists without waiting
    counts = torch.stack([(grid.mask[r0:r1, c0:c1] == 0).sum() for _, _, _, _, r0, r1, c0, c1 in cores]).cpu()
    jobs = [grid.region(r0, r1, c0, c1, n_seeds=int(n))
            for (_, _, _, _, r0, r1, c0, c1), n in zip(cores, counts.tolist())]
    steps = torch.zeros(1, dtype=torch.int64, device=grid.dev)
    hip.trace_fields_batch(hip.fields_jobs(jobs), steps=steps)
    return jobs, cores, steps


def _checked(grid, jobs, cores, theta, fields, max_err, min_f1, min_channel_px):
    """Per checked tile, in order: (check dict, predicted core view of
    fields, traced core).  The statistics of all the tiles are computed
    before the first is returned, and come back in one copy."""
    import torch
    view = fields.view(3, grid.nx, grid.ny)
    gm = grid.mask.view(grid.nx, grid.ny)
    items, stats = [], []
    for (p, seeds, mask, uv, slc, slt), (i0, i1, j0, j1, r0, r1, c0, c1) in zip(jobs, cores):
        f = fields_log(slc, slt, mask, grid.density, theta).view(3, r1 - r0, c1 - c0)
        tr = f[:, i0 - r0:i1 - r0, j0 - c0:j1 - c0]
        pr = view[:, i0:i1, j0:j1]
        stats.append(_tile_stats(pr, tr, gm[i0:i1, j0:j1] == 0))
        items.append(([i0, i1, j0, j1], pr, tr))
    got = torch.stack(stats).cpu().numpy() if stats else []
    for (core, pr, tr), g in zip(items, got):
        ok, med, f1, nch = _tile_verdict(g, max_err, min_f1, min_channel_px)
        yield {"core": core, "median_rel_err": med, "f1": f1, "channel_px": nch, "ok": ok}, pr, tr


def _chosen(n_tiles, check_frac, rng_seed):
    k = max(1, int(round(check_frac * n_tiles)))
    return np.sort(np.random.default_rng(rng_seed).choice(n_tiles, k, replace=False))


def check_tiles(fields, grid, theta, size=256, halo=16, check_frac=0.02, rng_seed=2, max_err=0.05, min_f1=0.95,
                min_channel_px=200):
    """The gate's checks (the same tiles, windows and criteria as gated_map)
    applied to given fields [3, nxy] over the grid, e.g. a whole-map
    prediction or another trace: every chosen tile is checked and fields is
    not changed.  Returns the list of check dicts."""
    origins = tile_origins(grid.nx, grid.ny, size, halo)
    jobs, cores, _ = _check_windows(grid, origins, _chosen(len(origins), check_frac, rng_seed), size, halo)
    return [c for c, _, _ in _checked(grid, jobs, cores, theta, fields, max_err, min_f1, min_channel_px)]


def gated_map(net, grid, theta, size=256, halo=16, batch=64, check_frac=0.02, rng_seed=2, max_err=0.05,
              min_f1=0.95, min_channel_px=200, max_fail_frac=0.25, workspace=None):
    """Surrogate map with tracer checks.  A random check_frac of the tiles
    (numpy default_rng(rng_seed)) is predicted and re-traced first, each
    tile's core as a grid of its own with a margin of max_length + 2 px
    (every streamline that can reach the core starts inside it).  A checked
    tile whose median relative log-density error exceeds max_err, or whose
    channel F1 is below min_f1 (when it has at least min_channel_px traced
    channel pixels), fails.  When more than max_fail_frac of the checked
    tiles fail, the map is plain_map and no other tile is predicted;
    otherwise the other tiles are predicted and each failed checked tile
    keeps its trace.  The checks run in their order and stop at the failure
    that decides the fallback (report["evaluated"] of report["checked"]);
    the first floor(max_fail_frac * k) + 1 of them, the fewest that can
    decide it, are predicted, set up and traced first, and the others only
    when the map has not fallen back by then.  Returns (fields [3, nxy],
    report)."""
    import torch
    origins = tile_origins(grid.nx, grid.ny, size, halo)
    n_tiles = len(origins)
    chosen = _chosen(n_tiles, check_frac, rng_seed)
    k = len(chosen)
    out = torch.empty((3, grid.nxy), dtype=torch.float32, device=grid.dev)
    first = min(k, int(math.floor(max_fail_frac * k)) + 1)
    checks, n_fail, check_steps = [], 0, 0
    for part in (chosen[:first], chosen[first:]):
        if not len(part) or n_fail > max_fail_frac * k:
            continue
        predict_tiles(net, grid, origins[part], size, halo, batch, out)
        jobs, cores, steps = _check_windows(grid, origins, part, size, halo)
        check_steps += int(steps.item())
        for check, pr, tr in _checked(grid, jobs, cores, theta, out, max_err, min_f1, min_channel_px):
            checks.append(check)
            if not check["ok"]:
                n_fail += 1
                if n_fail > max_fail_frac * k:
                    break   # the map falls back whatever the remaining checks say
                pr.copy_(tr)
    report = {"tiles": n_tiles, "checked": k, "evaluated": len(checks), "failed": n_fail,
              "check_steps": check_steps, "fallback": False, "checks": checks}
    if n_fail > max_fail_frac * k:
        out, _ = plain_map(grid, theta, workspace=workspace)
        report["fallback"] = True
        return out, report
    predict_tiles(net, grid, np.delete(origins, chosen, axis=0), size, halo, batch, out)
    return out, report


def _f1(pc, tc):
    tp = int((pc & tc).sum().item())
    fp = int((pc & ~tc).sum().item())
    fn = int((~pc & tc).sum().item())
    return 2 * tp / max(1, 2 * tp + fp + fn), tp + fn, tp + fp


def compare(fields, ref, mask, theta=None, shape=None):
    """Median relative error of log-density and channel-pixel F1 of fields
    against ref over the valid pixels (mask == 0); given theta, also the F1
    of channels read from the log-density (density >= theta); given the grid
    shape, also the F1 with one pixel of tolerance."""
    v = mask.reshape(-1) == 0
    r1 = ref[1][v]
    rel = (fields[1][v] - r1).abs() / r1.clamp_min(1e-6)
    rel0 = (fields[0][v] - ref[0][v]).abs() / ref[0][v].clamp_min(1e-6)
    tc = ref[2][v] > 0
    f1, n_ref, n_pred = _f1(fields[2][v] > 0, tc)
    out = {"median_rel_err_log_density": float(rel.median().item()),
           "median_rel_err_log_length": float(rel0.median().item()),
           "channel_f1": f1, "channel_px_ref": n_ref, "channel_px_pred": n_pred}
    if shape is not None:
        # F1 with one pixel of tolerance: a predicted channel pixel counts when
        # a reference channel pixel is among its 8 neighbours, and vice versa
        import torch
        pm = (fields[2] > 0).view(1, 1, *shape).float()
        tm = (ref[2] > 0).view(1, 1, *shape).float()
        vm = v.view(*shape)
        pd = torch.nn.functional.max_pool2d(pm, 3, 1, 1)[0, 0] > 0
        td = torch.nn.functional.max_pool2d(tm, 3, 1, 1)[0, 0] > 0
        p_ok = int(((pm[0, 0] > 0) & td & vm).sum().item())
        t_ok = int(((tm[0, 0] > 0) & pd & vm).sum().item())
        prec, recall = p_ok / max(1, n_pred), t_ok / max(1, n_ref)
        out["channel_f1_tol1px"] = 2 * prec * recall / max(1e-12, prec + recall)
    if theta is not None:
        out["channel_f1_from_density"] = _f1(fields[1][v] >= math.log1p(theta), tc)[0]
    return out
