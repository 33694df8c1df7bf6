"""The whole device pipeline of one tile (slm_amd.pipeline.map_tile) on the
256^2 workload: the stages run end to end and their outputs agree with the
stage-level parity tests' subjects (uv on the device equals the host uv the
job was built with, so the fields equal a direct slm_trace_fields call)."""
import numpy as np
import pytest

from slm_amd import workloads

pytestmark = pytest.mark.gpu


def test_map_tile_small_256(gpu):
    import torch
    from slm_amd import hip
    from slm_amd.pipeline import map_tile
    job = workloads.small_256(jitter=1.0, d=2)
    res, tim = map_tile(job)
    print(res, {k: round(v["gpu_ms"], 2) for k, v in tim.items()}, flush=True)
    # the trajectories run beside the fields pass, inside its stage
    assert set(tim) == {"uv", "fields", "finish", "mapping", "hsl", "kde"}
    assert res["chain_ok"] and res["n_segments"] > 0 and res["channel_pixels"] > 0
    # HSL traces start at the mid-slope pixels (lengths.py) and stop, unmeasured,
    # when they meet another mid-slope line: a share of them reaches a channel
    assert 0 < res["hsl_seeds"] <= res["midslope_pixels"]
    assert res["hsl_reached"] > 0.2 * res["hsl_seeds"]
    assert res["kde_mode_x"] is not None and np.isfinite(res["kde_mode_x"])
    # the batched both-directions launch takes as many steps as two direct
    # one-direction calls on the host-prepared inputs
    seeds = torch.from_numpy(job.seeds).to(gpu)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(gpu)
    uv = torch.from_numpy(job.uv.reshape(-1)).to(gpu)
    nxy = job.mask.size
    slc = torch.empty(nxy, dtype=torch.int32, device=gpu)
    slt = torch.empty_like(slc)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    for sign in (1, -1):
        hip.trace_fields(job.params(sign), seeds, mask, uv, slc, slt, steps, sync=True)
    assert int(steps.item()) == res["steps"]


def test_map_tile_with_conditioning(gpu):
    """The pipeline runs on the conditioned uv (the reference's default)."""
    from slm_amd import pipeline, workloads
    job = workloads.small_256()
    res, tim = pipeline.map_tile(job, condition=True)
    assert res["chain_ok"] and res["n_segments"] > 0 and res["hsl_seeds"] > 0
    assert "uv" in tim


def test_two_stream_trajectories_match_serial_calls(gpu):
    """The pipeline traces the two trajectory directions on the side stream
    while the fields pass runs on the main one. With a short max_length every
    output (traj [1000, 200, 2] int8 = 400 KB, steps, lengths) sits in the
    caching allocator's small pool, where a tensor freed while its kernel still
    ran could be handed to another output; two passes in a row must give what
    serial slm_trace_paths calls give."""
    import torch
    from slm_amd import hip
    from slm_amd.params import kernel_params
    from slm_amd.pipeline import map_tile
    from slm_amd.useful import create_seeds
    job = workloads.small_256(jitter=1.0, d=1, max_length=100.0)
    info = job.info
    outs = [map_tile(job, midslopes=False, return_grids=True)[0]["grids"] for _ in range(2)]
    tseeds, n_t, _ = create_seeds(job.mask, int(info.pad_width), int(info.n_work_items),
                                  n_seed_points=info.n_trajectory_seed_points, do_shuffle=info.do_shuffle,
                                  rng_seed=info.shuffle_rng_seed)
    info.n_seed_points = np.uint32(n_t)
    ts = torch.from_numpy(tseeds).to(gpu)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(gpu)
    uv = torch.from_numpy(np.ascontiguousarray(job.uv).reshape(-1)).to(gpu)
    want_l, want_n = [], []
    for sign in (1, -1):
        p = kernel_params(info, sign, "INTEGRATE_TRAJECTORY", False)
        assert n_t * int(p.max_n_steps) * 2 <= 1 << 20
        traj = torch.empty((n_t, int(p.max_n_steps), 2), dtype=torch.int8, device=gpu)
        ns = torch.empty(n_t, dtype=torch.int16, device=gpu)
        ln = torch.empty(n_t, dtype=torch.float32, device=gpu)
        hip.trace_paths(p, ts, mask, uv, traj, ns, ln, sync=True)
        want_l.append(ln.cpu().numpy())
        want_n.append(ns.cpu().numpy().view(np.uint16))
    for g in outs:
        assert np.array_equal(g["traj_length"], np.stack(want_l, 1))
        assert np.array_equal(g["traj_nsteps"], np.stack(want_n, 1))
