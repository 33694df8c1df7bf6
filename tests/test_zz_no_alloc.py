"""slm_trace_fields_noalloc (the fields tracer with its workspace in the
library's fixed device scratch array) against slm_trace_fields with a
caller workspace, through the C ABI: bit-identical slc, slt and step counts
on two calls in a row, the second with a smaller grid and seed list, both
directions; and a grid too large for the scratch array is refused before
anything is enqueued."""
import ctypes

import numpy as np
import pytest

from slm_amd import workloads

pytestmark = pytest.mark.gpu


def _dev(gpu, job):
    import torch
    return (torch.from_numpy(job.seeds).to(gpu),
            torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(gpu),
            torch.from_numpy(np.ascontiguousarray(job.uv).reshape(-1)).to(gpu))


def test_noalloc_matches_workspace_entry_point(gpu):
    import torch
    from slm_amd import hip
    big = workloads.small_256(jitter=1.0, d=2)
    small = workloads.small_256(jitter=1.0, n=96, d=1, seed=4)
    assert small.mask.size < big.mask.size and small.seeds.shape[0] < big.seeds.shape[0]
    for job in (big, small):
        seeds, mask, uv = _dev(gpu, job)
        for sign in (1, -1):
            p = job.params(sign)
            assert hip.fields_workspace_size(p) <= hip.noalloc_max_bytes()
            print("noalloc", job.name, job.mask.shape, "seeds", job.seeds.shape[0], "sign", sign, flush=True)
            out = []
            for noalloc in (True, False):
                slc = torch.full((p.nxy_padded,), 5, dtype=torch.int32, device=gpu)
                slt = torch.full_like(slc, 5)
                steps = torch.zeros(1, dtype=torch.int64, device=gpu)
                if noalloc:
                    hip.trace_fields_noalloc(p, seeds, mask, uv, slc, slt, steps, sync=True)
                else:
                    hip.trace_fields(p, seeds, mask, uv, slc, slt, steps, sync=True)
                out.append((hip.as_u32(slc), hip.as_u32(slt), int(steps.item())))
            assert out[0][2] == out[1][2] > 0
            assert np.array_equal(out[0][0], out[1][0])
            assert np.array_equal(out[0][1], out[1][1])


def test_noalloc_refuses_oversize_grid(gpu):
    import torch
    from slm_amd import hip
    job = workloads.small_256(jitter=1.0, n=96, d=1)
    seeds, mask, uv = _dev(gpu, job)
    p = job.params(1)
    p.nx_padded = p.ny_padded = 2052
    p.nxy_padded = 2052 * 2052
    assert hip.fields_workspace_size(p) > hip.noalloc_max_bytes()
    slc = torch.zeros(16, dtype=torch.int32, device=gpu)
    rc = hip.load().slm_trace_fields_noalloc(ctypes.byref(p), hip._ptr(seeds), hip._ptr(mask), hip._ptr(uv),
                                             hip._ptr(slc), hip._ptr(slc), None, hip._stream(None), 0)
    assert rc == -5
    assert b"maximum" in hip.load().slm_last_error()
