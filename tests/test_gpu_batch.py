"""GPU tests of the batched fields tracer (slm_trace_fields_batch).

Uneven tiles (sizes, pixel sizes, fans with unfilled pits, both directions,
d = 1 and 2) traced by one persistent launch must give, job for job, the
bits of a separate slm_trace_fields call; the packed-sum fallback and
accumulation behave as for one tile; the call captured in a hipGraph and
replayed (with eager work between replays) gives the eager result every
time; and the autotuner's choice gives the same bits."""
import numpy as np
import pytest

from slm_amd import workloads

pytestmark = pytest.mark.gpu


def _batch(gpu, n_tiles=6, d=1, seed=3):
    import torch
    from slm_amd import hip
    tiles = workloads.batch_c3(n_tiles=n_tiles, seed=seed, n_cones=2, d=d, side_range=(40, 140))
    jobs, host = [], []
    for t in tiles:
        seeds = torch.from_numpy(t.seeds).to(gpu)
        mask = torch.from_numpy(t.mask.astype(np.uint8).reshape(-1)).to(gpu)
        uv = torch.from_numpy(np.ascontiguousarray(t.uv).reshape(-1)).to(gpu)
        for sign in (1, -1):
            p = t.params(sign)
            slc = torch.full((p.nxy_padded,), 7, dtype=torch.int32, device=gpu)
            slt = torch.full_like(slc, 7)
            jobs.append((p, seeds, mask, uv, slc, slt))
            host.append(t.name)
    return jobs, hip.fields_jobs(jobs)


def _single(gpu, jobs):
    import torch
    from slm_amd import hip
    out, total = [], 0
    for p, seeds, mask, uv, _, _ in jobs:
        slc = torch.empty(p.nxy_padded, dtype=torch.int32, device=gpu)
        slt = torch.empty_like(slc)
        st = torch.zeros(1, dtype=torch.int64, device=gpu)
        hip.trace_fields(p, seeds, mask, uv, slc, slt, st, sync=True)
        out.append((hip.as_u32(slc), hip.as_u32(slt)))
        total += int(st.item())
    return out, total


@pytest.mark.parametrize("d", [1, 2])
def test_batch_matches_single_calls(gpu, d):
    import torch
    from slm_amd import hip
    jobs, arr = _batch(gpu, d=d)
    print("jobs", len(jobs), [(p.nx_padded, p.ny_padded) for p, *_ in jobs][:6], flush=True)
    ref, total = _single(gpu, jobs)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    hip.trace_fields_batch(arr, steps=steps, sync=True)
    assert int(steps.item()) == total
    for (p, _, _, _, slc, slt), (rc, rt) in zip(jobs, ref):
        assert np.array_equal(hip.as_u32(slc), rc) and np.array_equal(hip.as_u32(slt), rt)


def test_batch_fallback_and_accumulate(gpu):
    import torch
    from slm_amd import hip
    jobs, arr = _batch(gpu, n_tiles=4, seed=5)
    ref, _ = _single(gpu, jobs)
    ws = hip.trace_fields_batch(arr, force_fallback=True, sync=True)
    for (_, _, _, _, slc, slt), (rc, rt) in zip(jobs, ref):
        assert np.array_equal(hip.as_u32(slc), rc) and np.array_equal(hip.as_u32(slt), rt)
    hip.trace_fields_batch(arr, workspace=ws, accumulate=True, keep_order=True, sync=True)
    for (_, _, _, _, slc, slt), (rc, rt) in zip(jobs, ref):
        assert np.array_equal(hip.as_u32(slc), rc * 2) and np.array_equal(hip.as_u32(slt), rt * 2)


def test_batch_graph_replay(gpu):
    import torch
    from slm_amd import hip
    jobs, arr = _batch(gpu, n_tiles=5, seed=7)
    ws = torch.empty(hip.batch_workspace_size(arr), dtype=torch.uint8, device=gpu)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    hip.trace_fields_batch(arr, steps=steps, workspace=ws, sync=True)   # eager: order, device info
    eager = [(hip.as_u32(j[4]).copy(), hip.as_u32(j[5]).copy()) for j in jobs]
    eager_steps = int(steps.item())
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            hip.trace_fields_batch(arr, steps=steps, workspace=ws)
    torch.cuda.current_stream().wait_stream(side)
    for r in range(3):
        for j in jobs:                      # eager work between replays: scribble the outputs
            j[4].fill_(r + 11)
            j[5].fill_(r + 13)
        steps.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(steps.item()) == eager_steps, r
        for j, (ec, et) in zip(jobs, eager):
            assert np.array_equal(hip.as_u32(j[4]), ec) and np.array_equal(hip.as_u32(j[5]), et), r


def test_autotuned_opts_same_bits(gpu):
    from slm_amd import autotune, hip
    jobs, arr = _batch(gpu, n_tiles=4, seed=9)
    ref, _ = _single(gpu, jobs)
    opts, table = autotune.tune_batch(arr, repeats=1)
    assert len(table) >= 10 and opts.chunk in autotune.CHUNKS
    hip.trace_fields_batch(arr, opts=opts, sync=True)
    for (_, _, _, _, slc, slt), (rc, rt) in zip(jobs, ref):
        assert np.array_equal(hip.as_u32(slc), rc) and np.array_equal(hip.as_u32(slt), rt)
    assert opts.lds_slots in autotune.LDS_SLOTS


@pytest.mark.parametrize("slots", [256, 512, 1024])
def test_lds_table_sizes_same_bits(gpu, slots):
    """Every LDS accumulation-table size the tuner may pick gives the bits of
    separate single-tile calls (and the packed-sum fallback still agrees)."""
    import torch
    from slm_amd import hip
    jobs, arr = _batch(gpu, n_tiles=5, seed=11, d=2)
    ref, total = _single(gpu, jobs)
    for force in (False, True):
        steps = torch.zeros(1, dtype=torch.int64, device=gpu)
        hip.trace_fields_batch(arr, steps=steps, opts=hip.batch_opts(lds_slots=slots), force_fallback=force, sync=True)
        if not force:
            assert int(steps.item()) == total
        for (_, _, _, _, slc, slt), (rc, rt) in zip(jobs, ref):
            assert np.array_equal(hip.as_u32(slc), rc) and np.array_equal(hip.as_u32(slt), rt)


@pytest.mark.parametrize("refill,d", [(1, 1), (7, 2), (20, 2), (28, 1), (64, 1)])
def test_refill_batches_same_bits(gpu, refill, d):
    """Any refill batch (the idle lanes a wave gathers before it refills them)
    gives the bits and step count of separate single-tile calls."""
    import torch
    from slm_amd import hip
    jobs, arr = _batch(gpu, n_tiles=4, seed=13, d=d)
    ref, total = _single(gpu, jobs)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    hip.trace_fields_batch(arr, steps=steps, opts=hip.batch_opts(refill_batch=refill), sync=True)
    assert int(steps.item()) == total
    for (_, _, _, _, slc, slt), (rc, rt) in zip(jobs, ref):
        assert np.array_equal(hip.as_u32(slc), rc) and np.array_equal(hip.as_u32(slt), rt)


def test_refill_batch_out_of_range_rejected(gpu):
    from slm_amd import hip
    jobs, arr = _batch(gpu, n_tiles=2, seed=3)
    with pytest.raises(hip.SlmError):
        hip.trace_fields_batch(arr, opts=hip.batch_opts(refill_batch=65), sync=True)
