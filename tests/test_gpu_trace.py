"""GPU parity of the tracing path: libslm_hip.so against the CPU oracle.

Integer fields (slc, slt) and the accepted-step count must be bit-identical
to the oracle, jitter on or off, in both FP-contraction modes; recorded
positions and trajectory records must be bit-identical too.  The oracle is
pinned separately by tests/test_oracle.py.
"""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def _dev(job, dev):
    import torch
    seeds = torch.from_numpy(np.ascontiguousarray(job.seeds)).to(dev)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(dev)
    uv = torch.from_numpy(np.ascontiguousarray(job.uv).reshape(-1)).to(dev)
    return seeds, mask, uv


def _gpu_fields(p, seeds, mask, uv, dev, force_fallback=False):
    import torch
    from slm_amd import hip
    slc = torch.empty(p.nxy_padded, dtype=torch.int32, device=dev)
    slt = torch.empty(p.nxy_padded, dtype=torch.int32, device=dev)
    steps = torch.zeros(1, dtype=torch.int64, device=dev)
    hip.trace_fields(p, seeds, mask, uv, slc, slt, steps, sync=True, force_fallback=force_fallback)
    return hip.as_u32(slc), hip.as_u32(slt), int(steps.item())


def test_short_rounding_forms(gpu):
    """The short correctly rounded sqrt/divide forms the tracing kernels use
    return the operators' bits: every float in the sqrt and reciprocal ranges,
    6.6e9 hashed quotients plus every dividend significand for 2048 divisor
    significands, and the guarded wrappers over every bit pattern (NaN, inf,
    zero, denormal)."""
    from slm_amd import hip
    c = hip.check_rounding(8)
    print("rounding check counts", c, flush=True)
    assert c[1] > 1_600_000_000 and c[3] > 800_000_000 and c[5] == 8 * c[3] + 2048 * (1 << 23)
    assert c[0] == 0 and c[2] == 0 and c[4] == 0 and c[6] == 0 and c[7] == 0


CASES = [
    # (grid side, d, jitter, fp_contract[, max_integration_step_error])
    (96, 2, 1.0, True, 1e-6),
    (128, 1, 0.0, False, 1e-5),
    (64, 1, 0.0, True),
    (64, 2, 1.0, True),
    (96, 3, 1.0, False),
    (96, 3, 0.0, True),
    (256, 1, 1.0, True),
    (256, 1, 0.0, True),
    (256, 1, 1.0, False),
    (256, 2, 1.0, True),
]


@pytest.mark.parametrize("case", CASES)
def test_fields_bit_exact(gpu, case):
    from slm_amd import workloads
    n, d, jitter, contract = case[:4]
    job = workloads.small_256(jitter=jitter, n=n, d=d, max_error=case[4] if len(case) > 4 else None)
    seeds, mask, uv = _dev(job, gpu)
    m = job.mask.astype(np.uint8).reshape(-1)
    for sign in (1, -1):
        p = job.params(sign, fp_contract=contract)
        print("case n=%d d=%d jitter=%g contract=%d sign=%+d" % (n, d, jitter, contract, sign), flush=True)
        gslc, gslt, gsteps = _gpu_fields(p, seeds, mask, uv, gpu)
        oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1))
        assert gsteps == osteps
        assert np.array_equal(gslc, oslc), "slc differs at %d cells" % int((gslc != oslc).sum())
        assert np.array_equal(gslt, oslt), "slt differs at %d cells" % int((gslt != oslt).sum())


@pytest.mark.parametrize("res,d,contract", [(100, 2, True), (64, 1, False), (96, 2, False)])
def test_fields_other_resolutions(gpu, res, d, contract):
    """TRAJECTORY_RESOLUTION other than 128: a power of two (64) takes the
    exact-scaling form of approximate, others (100, 96) divide."""
    from slm_amd import workloads
    job = workloads.small_256(jitter=1.0, n=128, d=d, resolution=res)
    seeds, mask, uv = _dev(job, gpu)
    m = job.mask.astype(np.uint8).reshape(-1)
    for sign in (1, -1):
        p = job.params(sign, fp_contract=contract)
        assert p.trajectory_resolution == res
        print("case res=%d d=%d contract=%d sign=%+d" % (res, d, contract, sign), flush=True)
        gslc, gslt, gsteps = _gpu_fields(p, seeds, mask, uv, gpu)
        oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1))
        assert gsteps == osteps
        assert np.array_equal(gslc, oslc) and np.array_equal(gslt, oslt)


@pytest.mark.parametrize("n,d,jitter", [(96, 2, 1.0), (256, 1, 1.0)])
def test_fields_fallback_path(gpu, n, d, jitter):
    """The 32-bit fallback accumulation (taken when the packed 64-bit check
    fails) gives the same fields, also when accumulating into non-zero output."""
    import torch
    from slm_amd import hip, workloads
    job = workloads.small_256(jitter=jitter, n=n, d=d)
    seeds, mask, uv = _dev(job, gpu)
    m = job.mask.astype(np.uint8).reshape(-1)
    p = job.params(-1)
    gslc, gslt, gsteps = _gpu_fields(p, seeds, mask, uv, gpu, force_fallback=True)
    oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1))
    assert gsteps == osteps and np.array_equal(gslc, oslc) and np.array_equal(gslt, oslt)
    base = np.arange(p.nxy_padded, dtype=np.uint32) * np.uint32(2654435761)
    for force in (False, True):
        slc = torch.from_numpy(base.view(np.int32).copy()).to(gpu)
        slt = torch.from_numpy((base ^ np.uint32(0xdeadbeef)).view(np.int32).copy()).to(gpu)
        hip.trace_fields(p, seeds, mask, uv, slc, slt, accumulate=True, sync=True, force_fallback=force)
        assert np.array_equal(hip.as_u32(slc), base + oslc)
        assert np.array_equal(hip.as_u32(slt), (base ^ np.uint32(0xdeadbeef)) + oslt)


def test_fields_planar_and_empty(gpu):
    import torch
    from slm_amd import hip, workloads
    job = workloads.planar(n=48, d=2, jitter=0.0)
    seeds, mask, uv = _dev(job, gpu)
    p = job.params(1)
    gslc, gslt, gsteps = _gpu_fields(p, seeds, mask, uv, gpu)
    oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, job.mask.astype(np.uint8).reshape(-1), job.uv.reshape(-1))
    assert gsteps == osteps and np.array_equal(gslc, oslc) and np.array_equal(gslt, oslt)
    # empty seed list: fields cleared, no work
    p0 = job.params(1, n_seed_points=0)
    gslc, gslt, gsteps = _gpu_fields(p0, seeds, mask, uv, gpu)
    assert gsteps == 0 and not gslc.any() and not gslt.any()


def test_fields_chunk_invariance(gpu):
    """Seeds split into chunks (seeds_chunk_offset) accumulate to the same
    integer fields as one launch (the reference's adaptive chunking)."""
    import torch
    from slm_amd import hip, workloads
    job = workloads.small_256(jitter=1.0, n=128, d=2)
    seeds, mask, uv = _dev(job, gpu)
    p = job.params(1)
    ref = _gpu_fields(p, seeds, mask, uv, gpu)
    slc = torch.zeros(p.nxy_padded, dtype=torch.int32, device=gpu)
    slt = torch.zeros(p.nxy_padded, dtype=torch.int32, device=gpu)
    steps = torch.zeros(1, dtype=torch.int64, device=gpu)
    n = int(p.n_seed_points)
    bounds = [0, 1, 37, n // 2, n - 5, n]
    for a, b in zip(bounds[:-1], bounds[1:]):
        q = job.params(1, n_seed_points=b)
        q.seeds_chunk_offset = a
        hip.trace_fields(q, seeds, mask, uv, slc, slt, steps, accumulate=True, sync=True)
    assert int(steps.item()) == ref[2]
    assert np.array_equal(hip.as_u32(slc), ref[0]) and np.array_equal(hip.as_u32(slt), ref[1])


def test_rng_stream_bit_exact(gpu):
    import torch
    from slm_amd import hip
    n, nd = 1 << 16, 64
    rng0 = torch.empty(n, dtype=torch.int32, device=gpu)
    draws = torch.empty(n * nd, dtype=torch.int32, device=gpu)
    for idx0 in (0, 0xfffffff0 - n):
        hip.rng_stream(idx0, n, nd, rng0, draws, sync=True)
        o0, od = oracle.rng_stream(idx0, n, nd)
        assert np.array_equal(hip.as_u32(rng0), o0)
        assert np.array_equal(hip.as_u32(draws).reshape(n, nd), od)


@pytest.mark.parametrize("n,d,jitter,contract", [(256, 1, 1.0, True), (96, 3, 1.0, False), (256, 2, 0.0, True)])
def test_fields_record_positions(gpu, n, d, jitter, contract):
    """Recorded jittered positions: bit-identical to the FP32 oracle (so RMS 0
    <= 1e-4 px); the RMS against the FP64 oracle is reported."""
    import torch
    from slm_amd import hip, workloads
    job = workloads.small_256(jitter=jitter, n=n, d=d)
    seeds, mask, uv = _dev(job, gpu)
    p = job.params(1, fp_contract=contract)
    rs = np.random.default_rng(5)
    ksel = rs.integers(0, job.seeds.shape[0], 64).astype(np.uint32)
    ssel = rs.integers(0, d * d, 64).astype(np.uint32)
    max_rec = 2048
    xy = torch.zeros((64, max_rec, 2), dtype=torch.float32, device=gpu)
    nrec = torch.zeros(64, dtype=torch.int32, device=gpu)
    hip.trace_fields_record(p, seeds, mask, uv, torch.from_numpy(ksel.view(np.int32)).to(gpu),
                            torch.from_numpy(ssel.view(np.int32)).to(gpu), max_rec, xy, nrec, sync=True)
    gxy = xy.cpu().numpy()
    gn = nrec.cpu().numpy()
    m = job.mask.astype(np.uint8).reshape(-1)
    sq, cnt = 0.0, 0
    for j in range(64):
        oxy, _ = oracle.trace_fields_record(p, job.seeds, m, job.uv.reshape(-1), int(ksel[j]), int(ssel[j]), max_rec)
        assert gn[j] == len(oxy) or (gn[j] > max_rec and len(oxy) == max_rec)
        k = min(gn[j], max_rec)
        assert np.array_equal(gxy[j, :k], oxy[:k])
        dxy, _ = oracle.trace_fields_record(p, job.seeds, m, job.uv.reshape(-1), int(ksel[j]), int(ssel[j]),
                                            max_rec, f64=True)
        kk = min(k, len(dxy))
        sq += float(((gxy[j, :kk].astype(np.float64) - dxy[:kk]) ** 2).sum())
        cnt += kk
    print("RMS vs FP64 oracle over %d positions: %.3g px" % (cnt, (sq / max(cnt, 1)) ** 0.5))


@pytest.mark.parametrize("n,contract", [(256, True), (256, False), (96, True)])
def test_paths_bit_exact(gpu, n, contract):
    import torch
    from slm_amd import hip, workloads
    job = workloads.small_256(jitter=0.0, n=n, d=1)
    seeds, mask, uv = _dev(job, gpu)
    m = job.mask.astype(np.uint8).reshape(-1)
    ns = min(1000, job.seeds.shape[0])
    for sign in (1, -1):
        p = job.params(sign, kernel="INTEGRATE_TRAJECTORY", fp_contract=contract, n_seed_points=ns)
        traj = torch.zeros((ns, p.max_n_steps, 2), dtype=torch.int8, device=gpu)
        nst = torch.zeros(ns, dtype=torch.int16, device=gpu)
        ln = torch.zeros(ns, dtype=torch.float32, device=gpu)
        hip.trace_paths(p, seeds, mask, uv, traj, nst, ln, sync=True)
        otraj, onst, oln = oracle.trace_paths(p, job.seeds, m, job.uv.reshape(-1))
        assert np.array_equal(nst.cpu().numpy().view(np.uint16), onst)
        assert np.array_equal(ln.cpu().numpy().view(np.uint32), oln.view(np.uint32))
        assert np.array_equal(traj.cpu().numpy(), otraj)


@pytest.mark.parametrize("contract,offset", [(True, 12), (False, 8), (True, 20)])
def test_hsl_bit_exact(gpu, contract, offset):
    import torch
    from slm_amd import hip, workloads
    from slm_amd.params import kernel_params
    from slm_amd.useful import pick_seeds
    job, mapping = workloads.v_valley(midslope_offset=offset)
    seeds = pick_seeds(mask=job.mask, map=mapping, flag=int(job.info.is_midslope), pad=2)
    job.info.n_seed_points = np.uint32(seeds.shape[0])
    p = kernel_params(job.info, 1, "HILLSLOPE_LENGTHS", contract)
    thin, mid = int(job.info.is_thinchannel), int(job.info.is_midslope)
    init = np.full(seeds.shape[0], -1.0, dtype=np.float32)
    o = oracle.hillslope_lengths(p, seeds, job.mask.astype(np.uint8).reshape(-1), job.uv.reshape(-1),
                                 mapping.reshape(-1), thin, mid, length=init.copy())
    sd, m, uv = torch.from_numpy(seeds).to(gpu), *_dev(job, gpu)[1:]
    mp = torch.from_numpy(mapping.reshape(-1).view(np.int32)).to(gpu)
    L = torch.from_numpy(init.copy()).to(gpu)
    hip.hillslope_lengths(p, sd, m, uv, mp, thin, mid, L, sync=True)
    assert np.array_equal(L.cpu().numpy().view(np.uint32), o.view(np.uint32))
    assert (o == -1.0).sum() < 0.1 * o.size   # most reach the channel; the rest keep their value


@pytest.mark.parametrize("contract", [True, False])
def test_hsl_pits_bit_exact(gpu, contract):
    """Hillslope lengths from every pixel of a cone with unfilled pits, a few
    scattered thin-channel pixels: many traces circle in a pit until
    MAX_N_STEPS and must leave their length untouched, as in the oracle."""
    import torch
    from slm_amd import hip, synth, workloads
    from slm_amd.params import kernel_params
    from slm_amd.useful import pick_seeds
    roi = synth.make_dtm(128, pixel_size=1.0, relief=20.0, tilt=0.05, seed=7, nan_border=2, cone=True,
                         n_pits=10, fill=False)
    job = workloads.job_from_dtm("pits_128", roi, 1.0, {"subpixel_seed_point_density": 1, "jitter_magnitude": 0.0,
                                                        "max_length": 1000.0, "integrator_step_factor": 0.5})
    thin, mid = int(job.info.is_thinchannel), int(job.info.is_midslope)
    mapping = np.where(job.mask, 0, mid).astype(np.uint32)
    rs = np.random.default_rng(3)
    pick = rs.random(job.mask.shape) < 0.004
    mapping[pick & ~job.mask] = thin
    seeds = pick_seeds(mask=job.mask, map=mapping, flag=mid, pad=2)
    job.info.n_seed_points = np.uint32(seeds.shape[0])
    p = kernel_params(job.info, 1, "HILLSLOPE_LENGTHS", contract)
    init = np.full(seeds.shape[0], -1.0, dtype=np.float32)
    o = oracle.hillslope_lengths(p, seeds, job.mask.astype(np.uint8).reshape(-1), job.uv.reshape(-1),
                                 mapping.reshape(-1), thin, mid, length=init.copy())
    sd, m, uv = torch.from_numpy(seeds).to(gpu), *_dev(job, gpu)[1:]
    mp = torch.from_numpy(mapping.reshape(-1).view(np.int32)).to(gpu)
    L = torch.from_numpy(init.copy()).to(gpu)
    print("pits hsl seeds", seeds.shape[0], "reached", int((o != -1.0).sum()), flush=True)
    hip.hillslope_lengths(p, sd, m, uv, mp, thin, mid, L, sync=True)
    assert np.array_equal(L.cpu().numpy().view(np.uint32), o.view(np.uint32))
    assert 0 < (o != -1.0).sum() < o.size


TILED_CASES = [
    # (grid side, d, jitter, fp_contract, (tile, block, unit, cap, halo), max_error)
    (256, 2, 1.0, True, (32, 256, 0, 0, 0), None),
    (256, 1, 0.0, False, (16, 512, 0, 0, 2), None),
    (200, 3, 1.0, True, (48, 1024, 4096, 64, 4), None),
    (96, 2, 1.0, True, (8, 256, 100, 3, 1), 1e-6),
    (128, 1, 1.0, False, (24, 1024, 0, 0, 8), 1e-5),
    (256, 2, 1.0, True, (64, 1024, 0, 0, 4), None),
]


@pytest.mark.parametrize("n,d,jitter,contract,topts,max_error", TILED_CASES)
def test_fields_tiled_bit_exact(gpu, n, d, jitter, contract, topts, max_error):
    """The LDS-tiled tracer gives the oracle's fields, both directions."""
    import torch
    from slm_amd import hip, workloads
    job = workloads.small_256(jitter=jitter, n=n, d=d, max_error=max_error)
    seeds, mask, uv = _dev(job, gpu)
    m = job.mask.astype(np.uint8).reshape(-1)
    for sign in (1, -1):
        p = job.params(sign, fp_contract=contract)
        print("tiled n=%d d=%d jitter=%g opts=%s sign=%+d" % (n, d, jitter, topts, sign), flush=True)
        slc = torch.empty(p.nxy_padded, dtype=torch.int32, device=gpu)
        slt = torch.empty(p.nxy_padded, dtype=torch.int32, device=gpu)
        steps = torch.zeros(1, dtype=torch.int64, device=gpu)
        hip.trace_fields_tiled(p, seeds, mask, uv, slc, slt, steps, opts=hip.tile_opts(*topts), sync=True)
        oslc, oslt, osteps = oracle.trace_fields(p, job.seeds, m, job.uv.reshape(-1))
        assert int(steps.item()) == osteps
        assert np.array_equal(hip.as_u32(slc), oslc)
        assert np.array_equal(hip.as_u32(slt), oslt)


def test_fields_tiled_fallback_and_accumulate(gpu):
    import torch
    from slm_amd import hip, workloads
    job = workloads.small_256(jitter=1.0, n=128, d=2)
    seeds, mask, uv = _dev(job, gpu)
    p = job.params(1)
    oslc, oslt, _ = oracle.trace_fields(p, job.seeds, job.mask.astype(np.uint8).reshape(-1), job.uv.reshape(-1))
    base = np.arange(p.nxy_padded, dtype=np.uint32) * np.uint32(40503)
    for force in (False, True):
        slc = torch.from_numpy(base.view(np.int32).copy()).to(gpu)
        slt = torch.from_numpy(base.view(np.int32).copy()).to(gpu)
        hip.trace_fields_tiled(p, seeds, mask, uv, slc, slt, accumulate=True, sync=True, force_fallback=force)
        assert np.array_equal(hip.as_u32(slc), base + oslc)
        assert np.array_equal(hip.as_u32(slt), base + oslt)
