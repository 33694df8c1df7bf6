"""GPU parity of Mapping.map_hsl / map_aspect and Preprocess.slope_array
(slm_map_hsl, slm_map_aspect, slm_slope_deg) against the CPU oracle
(oracle/hsl_map_ref.py: scipy's grey_dilation, the skimage rank mean and
median restated from their definitions).  map_hsl is compared bit for bit;
aspect and slope, whose arctangents are float64 library calls rounded to
float32 on both sides, to within one float32 ulp, and the masks exactly."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import hsl_map_ref as H

pytestmark = pytest.mark.gpu


def _ulps(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


@pytest.mark.parametrize("shape,width,radius,ps", [((97, 130), 51, 30, 2.0), ((64, 70), 26, 9, 1.0),
                                                   ((120, 90), 0, 12, 2.0)])
def test_map_hsl_bit_exact(gpu, shape, width, radius, ps):
    import torch
    from slm_amd import hsl
    rs = np.random.default_rng(shape[0])
    lab = rs.integers(0, 12, size=(shape[0] // 8 + 1, shape[1] // 8 + 1))
    vals = np.where(rs.random(12) < 0.2, 0.0, rs.random(12) * 300).astype(np.float32)
    grid = vals[np.kron(lab, np.ones((8, 8), int))[:shape[0], :shape[1]]].astype(np.float32)
    mask = np.zeros(shape, bool)
    mask[:3], mask[:, -4:] = True, True
    mask[40:50, 20:35] = True
    want = H.map_hsl(grid, mask, width, radius, ps)
    got = hsl.map_hsl(torch.from_numpy(grid).to(gpu), torch.from_numpy(mask).to(gpu), width, radius, ps)
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want.astype(np.float32).view(np.uint32))


def test_map_aspect(gpu):
    import torch
    from slm_amd import hsl
    rs = np.random.default_rng(3)
    n, m = 90, 77
    f = ndimage.gaussian_filter(rs.random((n, m)), 3)
    slope = ((f - f.min()) / (f.max() - f.min()) * 40.0).astype(np.float32)   # 0..40 degrees
    ang = ndimage.gaussian_filter(rs.random((n, m)), 4) * 12
    uv = np.stack([np.cos(ang), np.sin(ang)], -1).astype(np.float32)
    mapping = np.zeros((n, m), np.uint32)
    mapping[rs.random((n, m)) < 0.03] = 1
    for do_med in (True, False):
        a_o, m_o = H.map_aspect(slope, uv, mapping, 1, 4.0, 12.0, 2.0, 15.0, do_med)
        a_g, m_g = hsl.map_aspect(torch.from_numpy(slope).to(gpu), torch.from_numpy(uv).to(gpu),
                                  torch.from_numpy(mapping.view(np.int32)).to(gpu), 1, 4.0, 12.0, 2.0, 15.0, do_med)
        a_g, m_g = a_g.cpu().numpy(), m_g.cpu().numpy()
        assert np.array_equal(m_g, m_o)
        d = _ulps(a_g, a_o)
        print("aspect ulps: max", int(d.max()), "exact share", float((d == 0).mean()), flush=True)
        assert d.max() <= 1 and (d == 0).mean() > 0.999


def test_slope_deg(gpu):
    import torch
    from slm_amd import hsl, synth
    from oracle import condition_ref
    roi = synth.make_dtm(70, pixel_size=2.0, relief=30.0, seed=4, nan_border=0)
    got = hsl.slope_deg(torch.from_numpy(roi).to(gpu), 2, 2.0).cpu().numpy()
    _, _, raw = condition_ref.velocity_field(roi, 2.0)
    s = (np.arctan(raw.astype(np.float64)).astype(np.float32) * np.float32(180.0 / np.pi)).astype(np.float32)
    want = np.pad(s, 2, mode="edge")
    d = _ulps(got, want)
    assert d.max() <= 1 and (d == 0).mean() > 0.999


def test_map_tile_fmm_products(gpu):
    """The FMM HSL, map_hsl, map_aspect and HSL-aspect statistics chain of
    map_tile(fmm=True) on the traced 256^2 workload: finite, consistent
    products (each step is checked against the oracle above)."""
    from slm_amd import pipeline, workloads
    job = workloads.small_256()
    res, tim = pipeline.map_tile(job, fmm=True)
    print({k: res[k] for k in res if k.startswith("fmm")}, res.get("hsl_aspect"), flush=True)
    assert res["fmm_subsegments"] > 10 and res["fmm_kept"] > 0 and res["fmm_mean_hsl_m"] > 0
    ha = res["hsl_aspect"]
    assert ha["hsl_mean"] > 0 and np.isfinite(ha["hsl_mean_azimuth"])
    assert "fmm" in tim and "aspect" in tim


def test_map_tile_fmm_overlap_matches_serial(gpu):
    """map_tile(fmm=True) runs the traced hillslope lengths, their KDE and the
    slope and aspect maps on the side stream while the FMM march runs on the
    main one.  Their results equal those of the serial stages of
    map_tile(fmm=False), and the FMM results equal a serial fmm_lengths on the
    same grids."""
    from slm_amd import hsl as fmm_hsl
    from slm_amd import pipeline, workloads
    job = workloads.small_256()
    ov, _ = pipeline.map_tile(job, fmm=True, return_grids=True)
    se, _ = pipeline.map_tile(job, fmm=False)
    for k in ("hsl_seeds", "hsl_reached", "hsl_median_m", "kde_mode_x", "n_segments", "channel_pixels"):
        assert ov[k] == se[k], k
    g = ov["grids"]
    info = job.info
    ref = fmm_hsl.fmm_lengths(g["label"], g["mapping"], g["mask"], int(info.is_thinchannel), job.pixel_size)
    assert ov["fmm_subsegments"] == ref.selected.size and ov["fmm_kept"] == int((ref.hsl > 0).sum())
    assert ov["fmm_mean_hsl_m"] == ref.fmm_mean_hsl
    assert ov["fmm_trace_mean_hsl_m"] == float(fmm_hsl.trace_mean_hsl(fmm_hsl.remap_hsl(g["label"], ref).cpu().numpy()))
