"""Autotuner for the batched fields tracer's LDS table and queue parameters.

slm_trace_fields_batch takes SlmBatchOpts: the number of work items a wave
grabs per queue atomic (chunk), the resident blocks per CU, the number of
queue heads the work list is cut into (one per XCD by default), and the size
of each wave's LDS accumulation table (256, 512 or 1024 slots: the pixel
window whose runs are summed in LDS before they reach global memory; larger
tables catch more runs but take LDS from occupancy), and the number of idle
lanes a wave gathers before it refills them (a refill costs the wave a fixed
stretch of serial work; idle lanes cost issue slots).  The best values depend
on the batch: many short trajectories favour small chunks (less tail
imbalance), long uniform ones large chunks (fewer atomics); converging flow
favours a larger table.  The tuner times candidates on the batch itself with
HIP events (coordinate descent: chunk, LDS table, refill batch, queue
slices, occupancy)
and keeps the winner per batch signature in memory and, if given, in a JSON
file.
"""
from __future__ import annotations

import hashlib
import json
import os

from . import hip

CHUNKS = (32, 64, 128, 256, 512, 1024)
SLICES = (1, 2, 4, 8)
BLOCKS = (0, 4, 3, 2)
LDS_SLOTS = (256, 512, 1024)
REFILL = (0, 12, 20, 28, 36)     # 0: the library's choice from the sub-pixel density

_cache: dict = {}


def signature(job_array) -> str:
    """Batch signature: grid sizes, seed counts, sub-pixel density and step caps."""
    h = hashlib.sha1()
    for j in job_array:
        p = j.params
        h.update(("%d,%d,%d,%d,%d;" % (p.nx_padded, p.ny_padded, p.n_seed_points, p.subpixel_seed_point_density,
                                       p.max_n_steps)).encode())
    return h.hexdigest()[:16]


def time_batch(job_array, opts, workspace, repeats=2, stream=None) -> float:
    """Mean ms of slm_trace_fields_batch with opts (seed order reused)."""
    import torch
    s = stream or torch.cuda.current_stream()
    hip.trace_fields_batch(job_array, workspace=workspace, opts=opts, stream=s, keep_order=True)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(repeats)]
    for a, b in ev:
        a.record(s)
        hip.trace_fields_batch(job_array, workspace=workspace, opts=opts, stream=s, keep_order=True)
        b.record(s)
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / repeats


def tune_batch(job_array, workspace=None, cache_path=None, repeats=2, verbose=False):
    """Returns (SlmBatchOpts, table of (chunk, blocks, slices, lds_slots, refill_batch, ms))."""
    import torch
    sig = signature(job_array)
    if cache_path and os.path.exists(cache_path) and not _cache:
        _cache.update(json.load(open(cache_path)))
    if sig in _cache:
        c = _cache[sig]
        return (hip.batch_opts(c["chunk"], c["blocks_per_cu"], c["queue_slices"], c.get("lds_slots", 512),
                               c.get("refill_batch", 0)), c.get("table", []))
    need = hip.batch_workspace_size(job_array)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=torch.device("cuda", torch.cuda.current_device()))
    hip.trace_fields_batch(job_array, workspace=workspace)   # builds the seed order and quads
    table = []
    best = {"chunk": 64, "blocks_per_cu": 0, "queue_slices": 8, "lds_slots": 512, "refill_batch": 0}

    def run(cfg):
        ms = time_batch(job_array, hip.batch_opts(cfg["chunk"], cfg["blocks_per_cu"], cfg["queue_slices"],
                                                  cfg["lds_slots"], cfg["refill_batch"]), workspace, repeats)
        table.append((cfg["chunk"], cfg["blocks_per_cu"], cfg["queue_slices"], cfg["lds_slots"], cfg["refill_batch"],
                      ms))
        if verbose:
            print("autotune", cfg, "%.2f ms" % ms, flush=True)
        return ms

    best_ms = run(best)
    for key, values in (("chunk", CHUNKS), ("lds_slots", LDS_SLOTS), ("refill_batch", REFILL),
                        ("queue_slices", SLICES), ("blocks_per_cu", BLOCKS)):
        for v in values:
            if v == best[key]:
                continue
            cfg = dict(best, **{key: v})
            ms = run(cfg)
            if ms < best_ms:
                best, best_ms = cfg, ms
    _cache[sig] = dict(best, ms=best_ms, table=table)
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(_cache, f, indent=1)
    return (hip.batch_opts(best["chunk"], best["blocks_per_cu"], best["queue_slices"], best["lds_slots"],
                           best["refill_batch"]), table)
