"""The benchmark configurations at their own sizes on the GPU.

- The 2048^2 tile (2 m, relief 60 m, three V-valleys, d = 2, jitter 1.0,
  MAX_N_STEPS 2000): both directions in one slm_trace_fields_batch launch,
  bit-exact against the CPU oracle on a 65536-seed prefix of the shuffled
  seed list (a uniform random sample, RNG keyed as in the full map); the full
  map through size-independent properties; the whole tile pipeline (fields,
  mapping and segmentation, hillslope lengths, KDE, FMM lengths, aspect) run
  end to end with its invariants checked.
- Batched uneven tiles: 8 tiles with sides 256-512 px, 1 m and 2 m, two of
  them fans with unfilled pits, MAX_N_STEPS 5000, both directions: 16 jobs in
  one persistent launch, every job bit-exact against the CPU oracle.
- The surrogate's training set (8192 tiles of 128^2) and gated 8192^2 map:
  one epoch beats the constant predictor; the gate's fallback is the plain
  map bit for bit.
"""
import numpy as np
import pytest

from slm_amd import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tile_2048():
    return workloads.tile_2048()


def _to_dev(gpu, job):
    import torch
    seeds = torch.from_numpy(job.seeds).to(gpu)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(gpu)
    uv = torch.from_numpy(np.ascontiguousarray(job.uv).reshape(-1)).to(gpu)
    return seeds, mask, uv


def test_tile_2048_batch_prefix_bit_exact(gpu, tile_2048):
    import torch
    from oracle import oracle
    from slm_amd import hip
    job = tile_2048
    n = 65536
    print("tile_2048 prefix", n, "seeds x 4 sub-seeds, both directions", flush=True)
    seeds, mask, uv = _to_dev(gpu, job)
    nxy = job.params(1).nxy_padded
    slc = torch.empty((2, nxy), dtype=torch.int32, device=gpu)
    slt = torch.empty_like(slc)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    jobs = hip.fields_jobs([(job.params(s, n_seed_points=n), seeds, mask, uv, slc[k], slt[k])
                            for k, s in enumerate((1, -1))])
    hip.trace_fields_batch(jobs, steps, sync=True)
    m8 = job.mask.astype(np.uint8).reshape(-1)
    total = 0
    for k, s in enumerate((1, -1)):
        oc, ot, ost = oracle.trace_fields(job.params(s, n_seed_points=n), job.seeds, m8, job.uv.reshape(-1),
                                          nthreads=16)
        total += ost
        assert np.array_equal(hip.as_u32(slc[k]), oc), s
        assert np.array_equal(hip.as_u32(slt[k]), ot), s
    assert int(steps.item()) == total


def test_tile_2048_full_map_properties(gpu, tile_2048):
    """The full map: every streamline writes its seed pixel once, so sum(slc)
    = streamlines + accepted steps per direction; no masked pixel is written;
    the two-direction launch equals two one-direction launches bit for bit."""
    import torch
    from slm_amd import hip
    job = tile_2048
    seeds, mask, uv = _to_dev(gpu, job)
    nxy = job.params(1).nxy_padded
    slc = torch.empty((2, nxy), dtype=torch.int32, device=gpu)
    slt = torch.empty_like(slc)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    jobs = hip.fields_jobs([(job.params(s), seeds, mask, uv, slc[k], slt[k]) for k, s in enumerate((1, -1))])
    hip.trace_fields_batch(jobs, steps, sync=True)
    m = job.mask.reshape(-1)
    n_lines = job.n_streamlines()
    total = 0
    for k, s in enumerate((1, -1)):
        c = hip.as_u32(slc[k]).astype(np.int64)
        one = torch.empty(nxy, dtype=torch.int32, device=gpu)
        one_t = torch.empty_like(one)
        st = torch.zeros(1, dtype=torch.int64, device=gpu)
        hip.trace_fields(job.params(s), seeds, mask, uv, one, one_t, st, sync=True)
        assert np.array_equal(hip.as_u32(one), hip.as_u32(slc[k]))
        assert np.array_equal(hip.as_u32(one_t), hip.as_u32(slt[k]))
        assert c.sum() == n_lines + int(st.item())
        assert not c[m].any() and not hip.as_u32(slt[k])[m].any()
        total += int(st.item())
    assert int(steps.item()) == total
    # mean accepted steps per streamline sits where the CPU oracle's sample puts it
    assert 200 < total / (2 * n_lines) < 400


def test_tile_2048_pipeline_end_to_end(gpu, tile_2048):
    from slm_amd.pipeline import map_tile
    res, tim = map_tile(tile_2048, fmm=True)
    print({k: round(v["gpu_ms"], 2) for k, v in tim.items()}, flush=True)
    assert res["chain_ok"]
    assert res["steps"] > 9e9
    assert res["n_segments"] > 10 and res["channel_pixels"] > 1000
    assert res["midslope_pixels"] > 0 and res["ridge_pixels"] > 0
    assert 0 < res["hsl_reached"] <= res["hsl_seeds"]
    assert 0 < res["hsl_median_m"] < 1000.0
    assert res["kde_mode_x"] is not None and np.isfinite(res["kde_mode_x"])
    assert res["fmm_kept"] > 0 and res["fmm_subsegments"] >= res["fmm_kept"]
    assert 0 < res["fmm_mean_hsl_m"] < 1000.0


def test_batched_uneven_tiles_bit_exact_vs_oracle(gpu):
    import torch
    from oracle import oracle
    from slm_amd import hip
    tiles = workloads.batch_c3(n_tiles=8, seed=5, n_cones=2, side_range=(256, 512))
    assert sum(t.name.endswith("_cone") for t in tiles) == 2
    jobs = []
    for t in tiles:
        seeds, mask, uv = _to_dev(gpu, t)
        for s in (1, -1):
            p = t.params(s)
            assert p.max_n_steps == 5000
            c = torch.empty(p.nxy_padded, dtype=torch.int32, device=gpu)
            jobs.append((p, seeds, mask, uv, c, torch.empty_like(c)))
    print("batch", len(jobs), "jobs", [t.name for t in tiles], flush=True)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    hip.trace_fields_batch(hip.fields_jobs(jobs), steps, sync=True)
    total = 0
    for t in tiles:
        m8 = t.mask.astype(np.uint8).reshape(-1)
        for s in (1, -1):
            p, _, _, _, c, tt = jobs.pop(0)
            oc, ot, ost = oracle.trace_fields(p, t.seeds, m8, t.uv.reshape(-1), nthreads=16)
            total += ost
            assert np.array_equal(hip.as_u32(c), oc), (t.name, s)
            assert np.array_equal(hip.as_u32(tt), ot), (t.name, s)
    assert int(steps.item()) == total


def test_surrogate_training_set_and_gated_map_at_size(gpu):
    """The surrogate's configurations at their own sizes.  Training: 32
    traced family DTMs of 2048^2 cut into 8192 tiles of 128^2, one epoch of
    Adam (batch 32, 256 steps); the epoch's mean loss is below the best
    constant predictor's.  Gated map: the 8192^2 family DTM (seed 2) in 1369
    tiles of 256^2 with 27 checked; a network one epoch old misses the
    tolerance, so the map falls back and equals the plain map bit for bit."""
    import torch
    from slm_amd import gated
    from slm_amd import surrogate as S
    print("training set 32 x 2048^2", flush=True)
    data = gated.training_set(n_dtms=32, n=2048, size=128, seed0=1000, device=gpu)
    assert data.n_tiles == 8192 and int(data.n_valid.sum()) > 0.99 * 8192 * 128 * 128
    const = gated.constant_loss(data)
    net, hist = gated.train(data, epochs=1, batch=32, lr=1e-3, device=gpu)
    print("constant %.4f, first epoch %.4f" % (const, hist["loss"][0]), flush=True)
    assert hist["steps"] == 256 and hist["loss"][0] < const
    inf = S.UNet(net.params_numpy(), device=gpu, inference=True)
    theta = data.theta
    del data, net
    torch.cuda.empty_cache()
    print("gated map 8192^2", flush=True)
    grid = gated.DeviceGrid(gated.family_dtm(8192, 2), device=gpu)
    out, rep = gated.gated_map(inf, grid, theta, size=256, halo=16, batch=64)
    print({k: rep[k] for k in ("tiles", "checked", "evaluated", "failed", "fallback")}, flush=True)
    assert rep["tiles"] == 1369 and rep["checked"] == 27
    assert rep["fallback"]
    plain, _ = gated.plain_map(grid, theta)
    assert torch.equal(out, plain)
