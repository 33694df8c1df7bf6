"""CPU pins of the mapping / segmentation oracle (oracle/slm_oracle_map.c).

1. Against oracle/pyref_map.py, a line-by-line pure-Python restatement of the
   reference kernels run work item after work item: for every kernel whose
   reference result does not depend on the schedule (same output for seed
   order and reversed seed order), the C oracle must match it bit for bit.
   The kernels that are schedule-free by construction must come out so.
2. Zhang-Suen thinning (the stand-in for scikit-image): closed-form cases
   and invariants (subset, idempotence, 8-connectivity kept).
3. Invariants of the whole chain on synthetic valleys with a known channel
   line and known flank sides.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle, pyref_map
from slm_amd import workloads
from slm_amd.useful import pick_seeds

from map_chain import OracleMapper, channels_from_mapping, STEPS

ORDER_FREE = {"connect_channels", "map_channel_heads", "label_confluences", "link_hillslopes",
              "segment_hillslopes", "subsegment_channel_edges", "subsegment_flanks"}


def _job(kind, n):
    if kind == "v":
        job, m = workloads.v_valley(n=n)
        m = channels_from_mapping(m, job.info)
    else:
        job, m = workloads.y_valley(n=n)
    job.info.set_thresholds(segmentation_threshold=np.uint32(10))
    return job, m


# (kernel, seed rule, arrays, chain step that precedes it)
def _kernel_calls(om):
    """Yield (name, seeds, array names) for each kernel launch of the chain,
    advancing the oracle mapper om between launches."""
    info, pad, mask = om.info, om.pad, om.mask
    yield "connect_channels", pick_seeds(map=om.mapping, flag=info.is_channel, pad=pad), ("uv", "mapping")
    om.connect_channel_pixels()
    om.thin_channels()
    thin = np.uint32(info.is_thinchannel)
    om.mapping[(om.mapping & thin) == thin] |= np.uint32(info.is_channelhead)
    yield "map_channel_heads", pick_seeds(mask=mask, flag=info.is_channel, pad=pad), ("uv", "mapping")
    om.run("map_channel_heads", pick_seeds(mask=mask, flag=info.is_channel, pad=pad), uv=om.uv, mapping=om.mapping)
    om.count_downchannels()
    heads = pick_seeds(mask=mask, map=om.mapping, flag=info.is_channelhead, pad=pad)
    om.clear(info.is_thinchannel)
    om.count[:] = 0
    yield "flag_downchannels", heads, ("mapping", "count", "link")
    om.run("flag_downchannels", heads, mapping=om.mapping, count=om.count, link=om.link)
    s = pick_seeds(mask=mask, map=om.mapping, flag=info.is_thinchannel, pad=pad)
    yield "label_confluences", s, ("slt", "mapping", "count", "link")
    om.run("label_confluences", s, slt=om.slt, mapping=om.mapping, count=om.count, link=om.link)
    om.flag_downchannels()
    om.flag_downchannels(False)
    om.flag_downchannels(False)
    om.label = np.zeros(om.mapping.shape, np.uint32)
    heads = pick_seeds(mask=mask, map=om.mapping, flag=info.is_channelhead, pad=pad)
    yield "segment_downchannels", heads, ("mapping", "count", "link", "label")
    om.run("segment_downchannels", heads, mapping=om.mapping, count=om.count, link=om.link, label=om.label)
    lab = om.label.view(np.int32)
    for i, l in enumerate(np.unique(lab[lab > 0])):
        lab[lab == l] = i + 1
    hs = pick_seeds(mask=mask, map=~om.mapping, flag=info.is_thinchannel, pad=pad)
    yield "link_hillslopes", hs, ("uv", "link")
    om.run("link_hillslopes", hs, uv=om.uv, link=om.link)
    yield "segment_hillslopes", hs, ("uv", "mapping", "label")
    om.run("segment_hillslopes", hs, uv=om.uv, mapping=om.mapping, label=om.label)
    s = pick_seeds(mask=mask, map=om.mapping, flag=info.is_subsegmenthead, pad=pad)
    yield "subsegment_channel_edges", s, ("mapping", "link", "label")
    om.run("subsegment_channel_edges", s, mapping=om.mapping, link=om.link, label=om.label)
    ft = int(info.is_leftflank) | int(info.is_thinchannel)
    s = pick_seeds(mask=mask, map=~om.mapping, flag=ft, pad=pad)
    yield "subsegment_flanks", s, ("uv", "mapping", "label")
    om.run("subsegment_flanks", s, uv=om.uv, mapping=om.mapping, label=om.label)
    yield "fix_right_flanks", s, ("uv", "mapping", "label")
    om.run("fix_right_flanks", s, uv=om.uv, mapping=om.mapping, label=om.label)
    s = pick_seeds(mask=mask, map=om.mapping, flag=info.is_leftflank, pad=pad)
    yield "fix_left_flanks", s, ("uv", "mapping", "label")


@pytest.mark.parametrize("kind,n", [("v", 32), ("y", 40)])
def test_oracle_matches_literal_restatement_where_schedule_free(kind, n):
    job, m0 = _job(kind, n)
    om = OracleMapper(job.info, job.mask, job.uv, m0, fp_contract=False)
    om.slt = np.random.default_rng(0).random(om.mapping.shape).astype(np.float32)
    compared, free = [], []
    for name, seeds, names in _kernel_calls(om):
        arrays = {a: (om.uv if a == "uv" else om.slt if a == "slt" else getattr(om, a)) for a in names}
        inout = [a for a in names if a not in ("uv", "slt")]
        p = om.params(name, seeds.shape[0])
        c_out = {a: arrays[a].reshape(-1).copy() for a in inout}
        oracle.map_call(name, p, seeds=seeds, mask=om.mask_u8, **{a: arrays[a] for a in names if a not in inout},
                        **c_out)
        runs = []
        for order in (None, list(range(seeds.shape[0]))[::-1]):
            py_out = {a: arrays[a].reshape(-1).copy() for a in inout}
            kw = {a: (arrays[a] if a in ("uv",) else arrays[a].reshape(-1)) for a in names if a not in inout}
            getattr(pyref_map, name)(p, seeds, om.mask_u8.astype(bool), **kw, **py_out, order=order)
            runs.append(py_out)
        sched_free = all(np.array_equal(runs[0][a], runs[1][a]) for a in inout)
        if name in ORDER_FREE:
            assert sched_free, name
        if sched_free:
            free.append(name)
            for a in inout:
                assert np.array_equal(c_out[a], runs[0][a]), (name, a)
            compared.append(name)
    assert ORDER_FREE <= set(compared)
    assert "fix_left_flanks" in compared


def test_racy_kernels_match_when_schedule_free():
    """count / segment / fix_right on the V valley (one channel head), where
    every schedule of the reference gives one answer."""
    job, m0 = _job("v", 32)
    om = OracleMapper(job.info, job.mask, job.uv, m0, fp_contract=False)
    om.connect_channel_pixels()
    om.thin_channels()
    om.map_channel_heads()
    info = om.info
    om.count = np.zeros(om.mapping.shape, np.uint32)
    om.link = np.zeros(om.mapping.shape, np.uint32)
    om.clear(info.is_thinchannel)
    heads = pick_seeds(mask=om.mask, map=om.mapping, flag=info.is_channelhead, pad=om.pad)
    assert heads.shape[0] >= 1
    p = om.params("count_downchannels", heads.shape[0])
    ref = [(om.mapping.reshape(-1).copy(), om.count.reshape(-1).copy(), om.link.reshape(-1).copy())
           for _ in range(2)]
    for r, order in zip(ref, (None, list(range(heads.shape[0]))[::-1])):
        pyref_map.count_downchannels(p, heads, om.mask_u8.astype(bool), om.uv, *r, order=order)
    assert all(np.array_equal(a, b) for a, b in zip(*ref))
    om.run("count_downchannels", heads, uv=om.uv, mapping=om.mapping, count=om.count, link=om.link)
    assert np.array_equal(om.mapping.reshape(-1), ref[0][0])
    assert np.array_equal(om.count.reshape(-1), ref[0][1])
    assert np.array_equal(om.link.reshape(-1), ref[0][2])
    # segment_downchannels after the flag passes
    om.flag_downchannels()
    om.slt = np.zeros(om.mapping.shape, np.float32)
    om.label_confluences()
    om.label = np.zeros(om.mapping.shape, np.uint32)
    heads = pick_seeds(mask=om.mask, map=om.mapping, flag=info.is_channelhead, pad=om.pad)
    p = om.params("segment_downchannels", heads.shape[0])
    pm, pl = om.mapping.reshape(-1).copy(), om.label.reshape(-1).copy()
    pyref_map.segment_downchannels(p, heads, om.mask_u8.astype(bool), pm, om.count.reshape(-1),
                                   om.link.reshape(-1), pl)
    om.run("segment_downchannels", heads, mapping=om.mapping, count=om.count, link=om.link, label=om.label)
    assert np.array_equal(om.mapping.reshape(-1), pm)
    assert np.array_equal(om.label.reshape(-1), pl)


# ------------------------------------------------------------------ thinning

def _thin(img):
    p = workloads.planar(n=8).params()
    nx, ny = img.shape
    p.nx_padded, p.ny_padded, p.nxy_padded = nx, ny, nx * ny
    m = np.where(img, np.uint32(1), np.uint32(0)).astype(np.uint32).reshape(-1)
    it = oracle.map_call("thin_channels", p, mapping=m)
    return ((m.reshape(nx, ny) & 2) != 0), it


def test_thinning_closed_forms():
    z = np.zeros((12, 20), bool)
    line = z.copy()
    line[5, 2:18] = True
    out, it = _thin(line)
    assert np.array_equal(out, line) and it == 1            # a 1-px line is already thin
    dot = z.copy()
    dot[4, 4] = True
    assert np.array_equal(_thin(dot)[0], dot)                # isolated pixel stays (B = 0)
    bar = z.copy()
    bar[4:7, 2:18] = True                                    # 3-px bar -> its middle row,
    out, _ = _thin(bar)                                      # ends eroded 1 (W) and 2 (E):
    want = z.copy()                                          # the first subiteration removes
    want[5, 3:16] = True                                     # the E end (N*E*S = E*S*W = 0)
    assert np.array_equal(out, want)
    sq = z.copy()
    sq[3:6, 3:6] = True                                      # 3x3 block -> a single pixel
    out, _ = _thin(sq)
    assert out.sum() == 1 and out[4, 4]


def test_thinning_invariants():
    rs = np.random.default_rng(5)
    img = ndimage.binary_dilation(rs.random((48, 64)) > 0.93, iterations=2)
    out, it = _thin(img)
    assert np.all(out <= img)
    assert ndimage.label(out, structure=np.ones((3, 3)))[1] == ndimage.label(img, structure=np.ones((3, 3)))[1]
    again, it2 = _thin(out)
    assert np.array_equal(again, out) and it2 == 1
    assert out.sum() < 0.5 * img.sum()


# ------------------------------------------------------------- whole chain

def _chain(kind, n, contract=True):
    job, m0 = _job(kind, n)
    om = OracleMapper(job.info, job.mask, job.uv, m0, fp_contract=contract)
    om.slt = np.zeros(om.mapping.shape, np.float32)
    for s in STEPS:
        r = om.step(s)
        assert r is not False, s
    return job, om


def test_v_valley_chain_invariants():
    job, om = _chain("v", 48)
    info = om.info
    c = 48 // 2 + 2 - job.origin[0]                       # axis column, cropped padded grid
    thin = (om.mapping & np.uint32(info.is_thinchannel)) != 0
    xs = np.unique(np.nonzero(thin)[0])
    assert list(xs) == [c]
    heads = np.argwhere((om.mapping & np.uint32(info.is_channelhead)) != 0)
    assert len(heads) == 1 and heads[0][0] == c
    lab = om.label.view(np.int32)
    ok = ~om.mask
    # every pixel is labelled except the wedge at the outlet whose traces leave
    # through the downstream border before they reach the channel
    unl = np.nonzero(ok & (lab == 0))
    assert unl[1].size < 0.1 * ok.sum() and unl[1].min() >= ok.shape[1] - 8
    # flanks: x < c left (negative), x > c right (positive), away from the ends
    inner = np.zeros_like(ok)
    inner[:, 8:-8] = True
    assert np.all(lab[:c - 1][(ok & inner)[:c - 1]] < 0)
    assert np.all(lab[c + 1:][(ok & inner)[c + 1:]] > 0)
    # channel segments are contiguous runs along y, labels 1..n in order
    chan = lab[c][ok[c]]
    assert np.all(np.diff(chan) >= 0) and chan.min() == 1 and chan.max() == om.n_segments
    # every hillslope label is a channel segment label
    assert set(np.abs(lab[ok]).tolist()) <= set(chan.tolist()) | {0}


def test_y_valley_chain_invariants():
    job, om = _chain("y", 64)
    info = om.info
    heads = np.argwhere((om.mapping & np.uint32(info.is_channelhead)) != 0)
    assert len(heads) == 2                                    # main and tributary tops
    conf = np.argwhere((om.mapping & np.uint32(info.is_majorconfluence)) != 0)
    assert len(conf) == 1
    j = conf[0]
    ssh = (om.mapping & np.uint32(info.is_subsegmenthead)) != 0
    assert ssh[j[0], j[1]]                                    # the confluence starts a segment
    lab = om.label.view(np.int32)
    assert om.n_segments >= 4
    # each channel segment's pixels form one 8-connected run
    chan = (om.mapping & np.uint32(info.is_thinchannel)) != 0
    for s in range(1, om.n_segments + 1):
        region = chan & (np.abs(lab) == s)
        if region.any():
            assert ndimage.label(region, structure=np.ones((3, 3)))[1] == 1, s


def test_cycle_fast_forward_is_exact():
    """Traces that circle in pits and link-graph cycles run to the reference's
    1000*MAX_N_STEPS cap; the oracle (and the GPU) skip whole laps.  On the
    traced 256^2 workload, which has such cycles, every grid of the chain
    must be the same with the skip and without it."""
    import time
    from oracle.postprocess_ref import postprocess_raw
    job = workloads.small_256(jitter=1.0)
    p = job.params(1)
    c, t, _ = oracle.trace_fields(p, job.seeds, job.mask.reshape(-1).astype(np.uint8), job.uv.reshape(-1))
    sh = job.mask.shape
    cc, tt, _ = postprocess_raw(c.reshape(sh)[..., None].repeat(2, 2), t.astype(np.float32).reshape(sh)[..., None]
                                .repeat(2, 2), 1)
    tt = tt[..., 0]
    thr = np.float32(np.quantile(tt[~job.mask], 0.97))
    m0 = np.zeros(sh, np.uint32)
    m0[(tt >= thr) & (tt.astype(np.float64) * 2 >= cc[..., 0])] = np.uint32(job.info.is_channel)
    job.info.set_thresholds(segmentation_threshold=np.uint32(20))
    states, times = [], []
    try:
        for ff in (True, False):
            oracle.map_set_fast_forward(ff)
            om = OracleMapper(job.info, job.mask, job.uv, m0, slt=tt.astype(np.float32))
            t0 = time.perf_counter()
            for s in ("connect", "thin", "heads", "count", "flag", "confluences", "segment", "link_hs"):
                om.step(s)
            times.append(time.perf_counter() - t0)
            states.append(om.state())
    finally:
        oracle.map_set_fast_forward(True)
    print("chain seconds with / without the skip:", times)
    for k in states[1]:
        assert np.array_equal(states[0][k], states[1][k]), k


def link_cycle_case():
    """A link graph with cycles for segment_downchannels: three heads whose
    paths enter one 9-pixel cycle at different pixels, a second 5-pixel cycle
    reached by one head, counts that rise along the paths and vary around the
    cycles so that segment starts are marked on them (threshold 4)."""
    from slm_amd import workloads
    n = 24
    p = workloads.planar(n=8).params()
    p.nx_padded = p.ny_padded = n
    p.nxy_padded = n * n
    p.segmentation_threshold = 4
    pad = int(p.pad_width)
    link = np.arange(n * n, dtype=np.uint32)            # self links: walks stop
    count = np.zeros(n * n, np.uint32)
    rs = np.random.default_rng(5)

    def pix(x, y):
        return x * n + y

    cyc = [pix(10, 3 + k) for k in range(9)]
    for a, b in zip(cyc, cyc[1:] + cyc[:1]):
        link[a] = b
    cyc2 = [pix(18, 4 + k) for k in range(5)]
    for a, b in zip(cyc2, cyc2[1:] + cyc2[:1]):
        link[a] = b
    heads = []
    for hx, entry, length in ((3, cyc[0], 6), (5, cyc[4], 3), (14, cyc2[2], 8)):
        path = [pix(hx, 3 + k) for k in range(length)]
        for a, b in zip(path, path[1:]):
            link[a] = b
        link[path[-1]] = entry
        count[path] = np.arange(1, length + 1) * 3
        heads.append(path[0])
    count[cyc] = rs.integers(5, 40, len(cyc))
    count[cyc2] = rs.integers(5, 40, len(cyc2))
    seeds = np.array([[h // n - pad, h % n - pad] for h in heads], np.float32)
    p.n_seed_points = len(heads)
    mask = np.zeros(n * n, np.uint8)
    return p, seeds, mask, count, link


def test_segment_walks_on_link_cycles_fast_forward_exact():
    """The full ownership walks of segment_downchannels circle on link cycles
    (the walk never meets a segment head there); skipping them on detection
    gives the same grids as walking to the hop cap."""
    p, seeds, mask, count, link = link_cycle_case()
    out = []
    try:
        for ff in (True, False):
            oracle.map_set_fast_forward(ff)
            mapping = np.zeros(p.nxy_padded, np.uint32)
            label = np.zeros(p.nxy_padded, np.uint32)
            oracle.map_call("segment_downchannels", p, seeds=seeds, mask=mask, mapping=mapping, count=count,
                            link=link, label=label)
            out.append((mapping.copy(), label.copy()))
    finally:
        oracle.map_set_fast_forward(True)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert (out[0][0] & np.uint32(0x4000)).sum() > 3 * 0x4000       # segment starts beyond the heads
