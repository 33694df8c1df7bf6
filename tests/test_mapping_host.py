"""Host glue of slm_amd.mapping.ChannelMapper (seed picking, flag resets,
relabelling, the flank sign fix) checked on the CPU: the HIP entry points
are replaced, for this test only, by the CPU oracle acting on the tensors'
memory, and the whole chain must equal the numpy OracleMapper's."""
import numpy as np
import pytest

from oracle import medial_axis_ref, oracle
from slm_amd import hip, mapping, workloads

from map_chain import OracleMapper, STEPS


@pytest.fixture
def cpu_kernels(monkeypatch):
    def fake_map_call(name, p, workspace=None, stream=None, sync=False, **arrays):
        np_arrays = {}
        for k, t in arrays.items():
            a = t.numpy()
            np_arrays[k] = a.view(np.uint32) if a.dtype == np.int32 else a
        oracle.map_call(name, p, **np_arrays)
        return workspace

    def fake_thin(p, m, workspace=None, stream=None, medial_axis=True):
        mm = m.numpy().view(np.uint32)
        src = mm
        if medial_axis:
            img = ((mm & np.uint32(hip.CHANNEL_BITS)) != 0).reshape(p.nx_padded, p.ny_padded)
            src = np.where(medial_axis_ref.medial_axis(img), np.uint32(1), np.uint32(0)).astype(np.uint32).reshape(-1)
        tmp = np.ascontiguousarray(src, dtype=np.uint32).copy()
        it = int(oracle.map_call("thin_channels", p, mapping=tmp))
        mm |= tmp & np.uint32(2)
        return it

    monkeypatch.setattr(hip, "map_call", fake_map_call)
    monkeypatch.setattr(hip, "thin_channels", fake_thin)
    monkeypatch.setattr(hip, "map_workspace_size", lambda p: 16)


def test_channel_mapper_glue_matches_oracle_mapper(cpu_kernels):
    import torch
    job, m0 = workloads.y_valley(n=48, channel_width=3)
    job.info.set_thresholds(segmentation_threshold=np.uint32(10))
    slt = np.random.default_rng(2).random(m0.shape).astype(np.float32)
    gm = mapping.ChannelMapper(job.info, torch.from_numpy(job.mask), torch.from_numpy(job.uv),
                               torch.from_numpy(m0.view(np.int32).copy()), slt=torch.from_numpy(slt))
    gm.workspace = torch.empty(16, dtype=torch.uint8)
    om = OracleMapper(job.info, job.mask, job.uv, m0, slt=slt)
    for s in STEPS:
        r_g = {"connect": gm.connect_channel_pixels, "thin": gm.thin_channels, "heads": gm.map_channel_heads,
               "count": gm.count_downchannels, "flag": gm.flag_downchannels,
               "confluences": gm.label_confluences, "segment": gm.segment_downchannels,
               "link_hs": gm.link_hillslopes, "segment_hs": gm.segment_hillslopes,
               "flanks": gm.subsegment_flanks}[s]()
        r_o = om.step(s)
        assert bool(r_g) == bool(r_o), s
        assert np.array_equal(gm.mapping.numpy().view(np.uint32), om.mapping), s
    assert gm.n_segments == om.n_segments > 0
    assert np.array_equal(gm.label.numpy().view(np.uint32), om.label)
    assert (gm.label.numpy() < 0).any() and (gm.label.numpy() > 0).any()


def test_pick_seeds_matches_reference_order():
    import torch
    from slm_amd.useful import pick_seeds as np_pick
    rs = np.random.default_rng(3)
    m = rs.integers(0, 64, size=(30, 40)).astype(np.uint32)
    mask = rs.random((30, 40)) < 0.2
    for kw in (dict(map=m, flag=6), dict(mask=mask, flag=1), dict(mask=mask, map=m, flag=8),
               dict(mask=mask, map=~m, flag=2)):
        want = np_pick(pad=2, **kw)
        tkw = {k: (torch.from_numpy(v.view(np.int32)) if k == "map" else torch.from_numpy(v) if k == "mask" else v)
               for k, v in kw.items()}
        got = mapping.pick_seeds(pad=2, **tkw).numpy()
        assert np.array_equal(got, want)


def test_channel_mapper_fused_walks_match_oracle_mapper(cpu_kernels, monkeypatch):
    """map_channels_segments runs the channel edges before one hillslope and
    flank walk (slm_segment_hillslopes_flanks, faked here as the oracle's
    segment_hillslopes then subsegment_flanks on the same seeds): the labels
    equal the reference's order of segment_hillslopes, channel edges and
    flanks."""
    import torch
    fake = hip.map_call

    def map_call(name, p, **kw):
        if name == "segment_hillslopes_flanks":
            fake("segment_hillslopes", p, **kw)
            return fake("subsegment_flanks", p, **kw)
        return fake(name, p, **kw)

    monkeypatch.setattr(hip, "map_call", map_call)
    job, m0 = workloads.y_valley(n=48, channel_width=3)
    job.info.set_thresholds(segmentation_threshold=np.uint32(10))
    slt = np.random.default_rng(2).random(m0.shape).astype(np.float32)
    gm = mapping.ChannelMapper(job.info, torch.from_numpy(job.mask), torch.from_numpy(job.uv),
                               torch.from_numpy(m0.view(np.int32).copy()), slt=torch.from_numpy(slt))
    gm.workspace = torch.empty(16, dtype=torch.uint8)
    assert gm.map_channels_segments(do_map_channels_from_scratch=False)
    om = OracleMapper(job.info, job.mask, job.uv, m0, slt=slt)
    for s in STEPS:
        assert om.step(s), s
    assert gm.n_segments == om.n_segments > 0
    assert np.array_equal(gm.mapping.numpy().view(np.uint32), om.mapping)
    assert np.array_equal(gm.label.numpy().view(np.uint32), om.label)
    assert (gm.label.numpy() < 0).any() and (gm.label.numpy() > 0).any()
