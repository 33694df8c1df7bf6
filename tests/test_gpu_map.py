"""GPU parity of the channel mapping and segmentation chain (libslm_hip.so,
slm_amd.mapping.ChannelMapper) against the CPU oracle, bit for bit.

Each step of Mapping.map_channels_segments runs on the GPU; before the step
the oracle is loaded with the GPU's current grids, runs the same step, and
every grid (mapping, count, link, label) must match exactly.  The chain is
also run end to end on the oracle alone and compared at the end."""
import numpy as np
import pytest

from slm_amd import workloads

from map_chain import OracleMapper, STEPS, channels_from_mapping

pytestmark = pytest.mark.gpu


def _gpu_step(gm, name):
    return {"connect": gm.connect_channel_pixels, "thin": gm.thin_channels, "heads": gm.map_channel_heads,
            "count": gm.count_downchannels, "flag": gm.flag_downchannels, "confluences": gm.label_confluences,
            "segment": gm.segment_downchannels, "link_hs": gm.link_hillslopes, "segment_hs": gm.segment_hillslopes,
            "flanks": gm.subsegment_flanks}[name]()


def _u32(t):
    return None if t is None else t.detach().cpu().numpy().view(np.uint32).copy()


def _case(kind):
    if kind == "v":
        job, m = workloads.v_valley(n=96)
        m = channels_from_mapping(m, job.info)
        slt = np.zeros(m.shape, np.float32)
    elif kind in ("y", "y100", "y_adapt"):
        job, m = workloads.y_valley(n=128, channel_width=3)
        slt = np.random.default_rng(1).random(m.shape).astype(np.float32)
        if kind == "y100":
            # a TRAJECTORY_RESOLUTION that is not a power of two: the traces
            # divide in approximate instead of scaling
            job.info.trajectory_resolution = np.uint32(100)
        if kind == "y_adapt":
            # a step-error tolerance small enough that dt adapts (the kernels
            # skip the step error only when it provably cannot)
            job.info.max_integration_step_error = np.float32(1e-7)
            job.info.adjusted_max_error = 0.85 * np.sqrt(np.float64(job.info.max_integration_step_error))
    elif kind == "pits":
        job, m, slt = _traced_case(_pitted_job())
    else:
        job, m, slt = _traced_case()
    job.info.set_thresholds(segmentation_threshold=np.uint32(20))
    return job, m, slt


def _pitted_job():
    """A cone with unfilled pits: deterministic traces caught in a pit circle
    on the position lattice until MAX_N_STEPS (the kernels skip whole laps)."""
    from slm_amd import synth
    roi = synth.make_dtm(160, pixel_size=1.0, relief=20.0, tilt=0.05, seed=5, nan_border=2, cone=True,
                         n_pits=12, fill=False)
    return workloads.job_from_dtm("pits_160", roi, 1.0, {"subpixel_seed_point_density": 1, "jitter_magnitude": 1.0,
                                                         "max_length": 1000.0, "integrator_step_factor": 0.5})


def _traced_case(job=None):
    """small_256 DTM (or job): downstream fields traced on the GPU, channels
    where Mapping.map_channels' test holds for the 97th-percentile slt."""
    import torch
    from slm_amd import hip
    from oracle.postprocess_ref import postprocess_raw
    job = job if job is not None else workloads.small_256(jitter=1.0)
    dev = torch.device("cuda", 0)
    p = job.params(1)
    seeds = torch.from_numpy(job.seeds).to(dev)
    mask = torch.from_numpy(job.mask.astype(np.uint8).reshape(-1)).to(dev)
    uv = torch.from_numpy(job.uv.reshape(-1)).to(dev)
    slc = torch.empty(p.nxy_padded, dtype=torch.int32, device=dev)
    slt = torch.empty_like(slc)
    hip.trace_fields(p, seeds, mask, uv, slc, slt, sync=True)
    slc_raw = hip.as_u32(slc).reshape(job.mask.shape)[..., None].repeat(2, axis=2)
    slt_raw = hip.as_u32(slt).astype(np.float32).reshape(job.mask.shape)[..., None].repeat(2, axis=2)
    c, t, _ = postprocess_raw(slc_raw, slt_raw, 1)
    t = t[..., 0]
    thr = np.float32(np.quantile(t[~job.mask], 0.97))
    job.info.set_thresholds(channel_threshold=thr)
    m = np.zeros(job.mask.shape, np.uint32)
    m[(t >= thr) & (t.astype(np.float64) * 2 >= c[..., 0])] = np.uint32(job.info.is_channel)
    return job, m, t.astype(np.float32)


def _load_oracle(om, gm):
    om.mapping = _u32(gm.mapping).reshape(gm.mapping.shape)
    for k in ("count", "link", "label", "channel_label"):
        v = getattr(gm, k)
        setattr(om, k, None if v is None else _u32(v).reshape(v.shape))


@pytest.mark.parametrize("kind,contract", [("v", True), ("y", True), ("y", False), ("y100", True), ("y_adapt", True),
                                           ("traced", True),
                                           ("pits", True)])
def test_chain_stepwise_bit_exact(gpu, kind, contract):
    import torch
    from slm_amd.mapping import ChannelMapper
    job, m0, slt = _case(kind)
    print("case", kind, "contract", contract, "grid", job.mask.shape, flush=True)
    dev = gpu
    gm = ChannelMapper(job.info, torch.from_numpy(job.mask).to(dev), torch.from_numpy(job.uv).to(dev),
                       torch.from_numpy(m0.view(np.int32)).to(dev), slt=torch.from_numpy(slt).to(dev),
                       fp_contract=contract)
    om = OracleMapper(job.info, job.mask, job.uv, m0, slt=slt, fp_contract=contract)
    solo = OracleMapper(job.info, job.mask, job.uv, m0, slt=slt, fp_contract=contract)
    for name in STEPS:
        print("step", name, flush=True)
        _load_oracle(om, gm)
        r_o = om.step(name)
        r_g = _gpu_step(gm, name)
        torch.cuda.synchronize()
        assert bool(r_o) == bool(r_g), name
        assert np.array_equal(_u32(gm.mapping).reshape(-1), om.mapping.reshape(-1)), (name, "mapping")
        for k in ("count", "link", "label"):
            if getattr(om, k) is not None:
                assert np.array_equal(_u32(getattr(gm, k)).reshape(-1), getattr(om, k).reshape(-1)), (name, k)
        if name == "segment":
            assert gm.n_segments == om.n_segments
        solo.step(name)
    # the oracle alone, end to end, lands on the same grids
    assert np.array_equal(_u32(gm.mapping).reshape(-1), solo.mapping.reshape(-1))
    assert np.array_equal(_u32(gm.label).reshape(-1), solo.label.reshape(-1))
    assert np.array_equal(_u32(gm.link).reshape(-1), solo.link.reshape(-1))
    assert gm.n_segments == solo.n_segments and gm.n_segments > 0


@pytest.mark.parametrize("kind", ["v", "y", "y100", "y_adapt", "traced", "pits"])
def test_chain_fused_walks_bit_exact(gpu, kind):
    """map_channels_segments as the pipeline runs it (the hillslope and flank
    walks in one launch, slm_segment_hillslopes_flanks) against the oracle's
    chain of the reference's separate steps: the same grids and result."""
    import torch
    from slm_amd.mapping import ChannelMapper
    job, m0, slt = _case(kind)
    print("case", kind, "grid", job.mask.shape, flush=True)
    gm = ChannelMapper(job.info, torch.from_numpy(job.mask).to(gpu), torch.from_numpy(job.uv).to(gpu),
                       torch.from_numpy(m0.view(np.int32)).to(gpu), slt=torch.from_numpy(slt).to(gpu))
    ok = gm.map_channels_segments(do_map_channels_from_scratch=False)
    torch.cuda.synchronize()
    solo = OracleMapper(job.info, job.mask, job.uv, m0, slt=slt)
    ok_o = True
    for name in STEPS:
        if not solo.step(name):
            ok_o = False
            break
    assert bool(ok) == ok_o
    assert np.array_equal(_u32(gm.mapping).reshape(-1), solo.mapping.reshape(-1))
    assert np.array_equal(_u32(gm.label).reshape(-1), solo.label.reshape(-1))
    assert np.array_equal(_u32(gm.link).reshape(-1), solo.link.reshape(-1))
    assert gm.n_segments == solo.n_segments and gm.n_segments > 0


@pytest.mark.parametrize("shape,density,grow", [((300, 420), 0.985, 4), ((257, 130), 0.996, 14), ((70, 65), 1.0, 0)])
def test_thinning_bit_exact(gpu, shape, density, grow):
    """Bit-exact skeleton and iteration count; thick blobs need more than one
    launch of the tiled kernel (8 iterations each); an empty image needs one
    iteration."""
    import torch
    from scipy import ndimage
    from oracle import oracle
    from slm_amd import hip
    rs = np.random.default_rng(7)
    img = rs.random(shape) > density
    if grow:
        img = ndimage.binary_dilation(img, iterations=grow)
    p = workloads.planar(n=8).params()
    p.nx_padded, p.ny_padded, p.nxy_padded = img.shape[0], img.shape[1], img.size
    m = np.where(img, np.uint32(1), np.uint32(0)).astype(np.uint32).reshape(-1)
    it_o = oracle.map_call("thin_channels", p, mapping=m)
    mg = torch.from_numpy(np.where(img, 1, 0).astype(np.int32).reshape(-1)).to(gpu)
    it_g = hip.thin_channels(p, mg, medial_axis=False)
    torch.cuda.synchronize()
    assert it_g == it_o
    assert np.array_equal(hip.as_u32(mg), m)


@pytest.mark.parametrize("shape,density,grow,seed", [((300, 420), 0.985, 4, 7), ((257, 130), 0.996, 14, 8),
                                                     ((130, 97), 0.97, 2, 9), ((70, 65), 1.0, 0, 1)])
def test_medial_axis_thinning_bit_exact(gpu, shape, density, grow, seed):
    """skeletonize(medial_axis(.)) as Mapping.thin_channels calls it: the
    GPU's rounds against the oracle's one-by-one visit of the medial-axis
    restatement, then the same thinning; wide blobs, many objects, and an
    empty image."""
    import torch
    from scipy import ndimage
    from oracle import medial_axis_ref, oracle
    from slm_amd import hip
    rs = np.random.default_rng(seed)
    img = rs.random(shape) > density
    if grow:
        img = ndimage.binary_dilation(img, iterations=grow)
    print("medial axis", shape, "set pixels", int(img.sum()), flush=True)
    p = workloads.planar(n=8).params()
    p.nx_padded, p.ny_padded, p.nxy_padded = img.shape[0], img.shape[1], img.size
    ma = medial_axis_ref.medial_axis(img)
    m = np.where(ma, np.uint32(1), np.uint32(0)).astype(np.uint32).reshape(-1)
    it_o = oracle.map_call("thin_channels", p, mapping=m)
    want = np.where(img, np.uint32(1), np.uint32(0)).reshape(-1) | (m & np.uint32(2))
    mg = torch.from_numpy(np.where(img, 1, 0).astype(np.int32).reshape(-1)).to(gpu)
    it_g = hip.thin_channels(p, mg)
    torch.cuda.synchronize()
    assert it_g == it_o
    assert np.array_equal(hip.as_u32(mg), want)


def test_mapping_entry_points_reject_bad_input(gpu):
    import ctypes
    import torch
    from slm_amd import hip
    job, m0, _ = _case("v")
    p = job.params(1)
    p.n_seed_points = 4
    seeds = torch.zeros((4, 2), dtype=torch.float32, device=gpu)
    mask = torch.zeros(p.nxy_padded, dtype=torch.uint8, device=gpu)
    mp = torch.zeros(p.nxy_padded, dtype=torch.int32, device=gpu)
    small = torch.empty(16, dtype=torch.uint8, device=gpu)
    lib = hip.load()
    rc = lib.slm_count_downchannels(ctypes.byref(p), hip._ptr(seeds), hip._ptr(mask), hip._ptr(mp), hip._ptr(mp),
                                    hip._ptr(mp), hip._ptr(mp), hip._ptr(small), 16, hip._stream(None), 0)
    assert rc == -5                                       # SLM_E_WORKSPACE
    p.n_seed_points = 0                                  # empty seed list: nothing to do, no error
    assert lib.slm_connect_channels(ctypes.byref(p), None, None, None, None, hip._stream(None), 1) == 0


def test_segment_downchannels_on_link_cycles(gpu):
    """segment_downchannels over a link graph with cycles (heads whose paths
    run into them): the GPU's cycle-skipping walks give the oracle's grids."""
    import torch
    from oracle import oracle
    from slm_amd import hip
    from test_oracle_map import link_cycle_case
    p, seeds, mask, count, link = link_cycle_case()
    mapping = np.zeros(p.nxy_padded, np.uint32)
    label = np.zeros(p.nxy_padded, np.uint32)
    oracle.map_call("segment_downchannels", p, seeds=seeds, mask=mask, mapping=mapping, count=count, link=link,
                    label=label)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32) if a.dtype == np.uint32 else a).to(gpu)
    mg, lg = t(np.zeros(p.nxy_padded, np.uint32)), t(np.zeros(p.nxy_padded, np.uint32))
    hip.map_call("segment_downchannels", p, seeds=t(seeds), mask=t(mask), mapping=mg, count=t(count), link=t(link),
                 label=lg, sync=True)
    assert np.array_equal(hip.as_u32(mg), mapping)
    assert np.array_equal(hip.as_u32(lg), label)
