"""The oracle's fields pass against the closed form of a planar DTM
(tests/planar.py): with jitter off every step is the same lattice vector, so
slc, slt and the step count follow by arithmetic.  Three cases bind each
stop rule: the grid edge (max_length 1000 m), the step cap (MAX_N_STEPS =
max_length / integrator_step_factor) and the length cap (MAX_LENGTH =
max_length / pixel_size, below the step cap at 2 m pixels); both directions,
flow along x and along y, one and four sub-pixel starts."""
import numpy as np
import pytest

from planar import planar_fields, planar_job, step_vector

CASES = [
    # (nx, ny, axis, pixel_size, d, max_length)
    (48, 40, 0, 1.0, 1, 1000.0),    # grid edge
    (40, 56, 1, 1.0, 2, 12.0),      # step cap: 24 steps of < 0.5 px
    (56, 48, 0, 2.0, 2, 12.0),      # length cap: 6 px
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("sign", (1, -1))
def test_oracle_fields_planar_closed_form(case, sign):
    from oracle import oracle
    nx, ny, axis, ps, d, ml = case
    job = planar_job(nx, ny, axis, ps, d, ml)
    p = job.params(sign)
    assert p.jitter_magnitude == 0.0
    s, dd = step_vector(p)
    assert 0.0 < abs(float(dd)) <= 0.5 and (dd > 0) == (sign > 0)
    slc, slt, steps = planar_fields(job, p, axis)
    assert steps > 0 and slc.sum() == steps + job.seeds.shape[0] * d * d
    oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, job.mask.astype(np.uint8).reshape(-1),
                                             job.uv.reshape(-1))
    assert osteps == steps
    assert np.array_equal(oslc, slc)
    assert np.array_equal(oslt, slt)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_fields_planar_closed_form(gpu, case):
    """The HIP tracer (slm_trace_fields and the two-job slm_trace_fields_batch)
    against the same closed form."""
    import torch
    from slm_amd import hip
    nx, ny, axis, ps, d, ml = case
    job = planar_job(nx, ny, axis, ps, d, ml)
    seeds = torch.from_numpy(job.seeds).to(gpu)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(gpu)
    uv = torch.from_numpy(np.ascontiguousarray(job.uv).reshape(-1)).to(gpu)
    want = {sign: planar_fields(job, job.params(sign), axis) for sign in (1, -1)}
    nxy = job.params(1).nxy_padded
    for sign in (1, -1):
        p = job.params(sign)
        slc = torch.empty(nxy, dtype=torch.int32, device=gpu)
        slt = torch.empty_like(slc)
        steps = torch.zeros(1, dtype=torch.int64, device=gpu)
        hip.trace_fields(p, seeds, mask, uv, slc, slt, steps, sync=True)
        assert int(steps.item()) == want[sign][2]
        assert np.array_equal(hip.as_u32(slc), want[sign][0])
        assert np.array_equal(hip.as_u32(slt), want[sign][1])
    slc = torch.empty((2, nxy), dtype=torch.int32, device=gpu)
    slt = torch.empty_like(slc)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    jobs = hip.fields_jobs([(job.params(s), seeds, mask, uv, slc[k], slt[k]) for k, s in enumerate((1, -1))])
    hip.trace_fields_batch(jobs, steps, sync=True)
    assert int(steps.item()) == want[1][2] + want[-1][2]
    for k, sign in enumerate((1, -1)):
        assert np.array_equal(hip.as_u32(slc[k]), want[sign][0])
        assert np.array_equal(hip.as_u32(slt[k]), want[sign][1])
