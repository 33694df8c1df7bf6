"""GPU parity of the FMM hillslope lengths (slm_amd.hsl: slm_fmm_pairs,
slm_fmm_hsl, slm_hsl_remap) against the CPU oracle (oracle/fmm_ref.py).

Second order (the default, scikit-fmm's): the GPU marches every subsegment
in the oracle's acceptance order, so counts, fixed-point sums, failure flags
and HSL are the oracle's exactly.  First order: counts exactly, the
fixed-point distance sums to within one unit per point (the GPU relaxation
and the oracle's fast march reach the same discrete solution, but a value can
come from a different update by an ulp), and the int32 HSL exactly unless
twice the mean distance lies within 1e-6 of an integer."""
import numpy as np
import pytest

from oracle import fmm_ref

pytestmark = pytest.mark.gpu


def _voronoi_labels(n, m, k, seed):
    rs = np.random.default_rng(seed)
    cx, cy = rs.integers(2, n - 2, k), rs.integers(2, m - 2, k)
    x, y = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    d = (x[..., None] - cx) ** 2 + (y[..., None] - cy) ** 2
    owner = d.argmin(axis=2)
    labels = (rs.permutation(k) + 3) * np.where(rs.random(k) < 0.5, -1, 1)
    label = labels[owner].astype(np.int32)
    mapping = np.zeros((n, m), np.uint32)
    # a channel through each cell: its pixels on the centre's row
    for j in range(k):
        mapping[(owner == j) & (x == cx[j])] = 2
    mask = np.zeros((n, m), bool)
    mask[:2], mask[-2:], mask[:, :2], mask[:, -2:] = True, True, True, True
    label[mask] = 0
    return label, mapping, mask


def _compare(gpu, label, mapping, mask, dx, order=2):
    import torch
    from slm_amd import hsl
    lt = torch.from_numpy(label).to(gpu)
    mt = torch.from_numpy(mapping.view(np.int32)).to(gpu)
    kt = torch.from_numpy(mask).to(gpu)
    res = hsl.fmm_lengths(lt, mt, kt, 2, dx, order=order)
    sel, h_o, stats = fmm_ref.fmm_lengths(label, mapping, mask, 2, dx, order=order)
    print("order", order, "subsegments", sel.size, "colours", res.n_colours, flush=True)
    assert np.array_equal(res.selected, sel)
    for s, hg, ho in zip(sel, res.hsl, h_o):
        a = abs(int(s))
        q, cnt, n_phi, n_ch = stats[a]
        if order == 2:
            assert bool(res.failed[a]) == (q is None), a
            if q is None:
                assert hg == 0 and ho == 0
                continue
            assert (int(res.cnt[a]), int(res.nphi[a]), int(res.nch[a])) == (cnt, n_phi, n_ch), a
            assert int(res.sums[a]) == q, a
            assert hg == ho, a
            continue
        assert (int(res.cnt[a]), int(res.nphi[a]), int(res.nch[a])) == (cnt, n_phi, n_ch), a
        assert abs(int(res.sums[a]) - q) <= cnt, a
        if n_ch and n_phi / n_ch > 4.0:
            h = 2.0 * (q / fmm_ref.FIX) / cnt
            if abs(h - round(h)) > 1e-6:
                assert hg == ho, (a, hg, ho)
        else:
            assert hg == 0 and ho == 0
    grid = hsl.remap_hsl(lt, res).cpu().numpy()
    assert np.array_equal(grid, fmm_ref.remap(label, sel, res.hsl))
    assert res.fmm_mean_hsl == fmm_ref.fmm_mean_hsl(h_o)
    return res


# cells of ~330 px march one lane each; ~1600 px (the 32 KiB LDS heaps) and
# ~5600 px (the large LDS heaps) one 64-lane block each
@pytest.mark.parametrize("shape,k,seed,dx", [((64, 50), 9, 0, 1.0), ((130, 97), 40, 1, 2.0), ((200, 200), 120, 2, 1.0),
                                             ((256, 256), 40, 4, 2.0), ((300, 262), 14, 3, 1.0)])
@pytest.mark.parametrize("order", [2, 1])
def test_fmm_voronoi_subsegments(gpu, shape, k, seed, dx, order):
    label, mapping, mask = _voronoi_labels(shape[0], shape[1], k, seed)
    res = _compare(gpu, label, mapping, mask, dx, order)
    assert (res.hsl > 0).sum() > 0


def test_fmm_lds_heaps_match_lanes(gpu):
    """Second order on the traced 2048^2 tile's segmentation (subsegments up
    to ~10^4 px): the LDS-heap marches against one lane per subsegment, bit
    for bit (sums, counts, failures)."""
    import torch
    from slm_amd import hsl, pipeline, workloads
    job = workloads.tile_2048()
    res, _ = pipeline.map_tile(job, return_grids=True)
    g = res["grids"]
    lab = g["label"].cpu().numpy()
    sizes = np.bincount(np.abs(lab.astype(np.int64)).ravel())[1:]
    print("subsegments", int((sizes > 0).sum()), "largest", int(sizes.max()), "over 512 px", int((sizes > 512).sum()),
          flush=True)
    assert sizes.max() > 2048
    a = hsl.fmm_lengths(g["label"], g["mapping"], g["mask"], int(job.info.is_thinchannel), job.pixel_size)
    b = hsl.fmm_lengths(g["label"], g["mapping"], g["mask"], int(job.info.is_thinchannel), job.pixel_size,
                        lanes_only=True)
    for f in ("sums", "cnt", "nch", "nphi", "failed", "hsl", "selected"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_fmm_colour_groups(gpu):
    """One colour per relaxation group (the minimum workspace) and every
    colour in one group solve the same problems: identical counts, and sums
    within the oracle tolerance of each other."""
    import torch
    from slm_amd import hsl
    label, mapping, mask = _voronoi_labels(130, 97, 40, 1)
    lt = torch.from_numpy(label).to(gpu)
    mt = torch.from_numpy(mapping.view(np.int32)).to(gpu)
    kt = torch.from_numpy(mask).to(gpu)
    one = hsl.fmm_lengths(lt, mt, kt, 2, 2.0, max_group=1, order=1)
    two = hsl.fmm_lengths(lt, mt, kt, 2, 2.0, max_group=2, order=1)
    allc = hsl.fmm_lengths(lt, mt, kt, 2, 2.0, order=1)
    print("colours", allc.n_colours, flush=True)
    assert allc.n_colours >= 3
    for r in (two, allc):
        assert np.array_equal(r.cnt, one.cnt) and np.array_equal(r.nch, one.nch) and np.array_equal(r.nphi, one.nphi)
        assert np.all(np.abs(r.sums.astype(np.int64) - one.sums.astype(np.int64)) <= one.cnt.astype(np.int64))
        assert np.array_equal(r.selected, one.selected)
    # second order: the march is exact, so any grouping gives the same bits
    one2 = hsl.fmm_lengths(lt, mt, kt, 2, 2.0, max_group=1)
    all2 = hsl.fmm_lengths(lt, mt, kt, 2, 2.0)
    assert np.array_equal(one2.sums, all2.sums) and np.array_equal(one2.failed, all2.failed)


def test_fmm_traced_segmentation(gpu):
    """The labels and thin channels of the segmentation chain on the traced
    256^2 workload (the grids map_tile produces)."""
    from slm_amd import pipeline, workloads
    job = workloads.small_256()
    res, _ = pipeline.map_tile(job, return_grids=True)
    g = res["grids"]
    label = g["label"].cpu().numpy().reshape(job.mask.shape).astype(np.int32)
    mapping = g["mapping"].cpu().numpy().reshape(job.mask.shape).view(np.uint32)
    for order in (2, 1):
        r = _compare(gpu, label, mapping, job.mask, job.pixel_size, order)
        assert r.selected.size > 10 and (r.hsl > 0).sum() > 0


def test_fmm_second_order_failures_match(gpu):
    """Subsegments whose channel is a few scattered pixels: some second-order
    marches meet a negative discriminant (scikit-fmm raises; the reference
    leaves them unmeasured).  The GPU flags exactly the oracle's, and the
    others match it bit for bit."""
    import torch
    from slm_amd import hsl
    label, mapping, mask = _voronoi_labels(120, 110, 30, 7)
    rs = np.random.default_rng(7)
    mapping[:] = 0
    pts = rs.integers(3, 117, size=(150, 2))
    mapping[pts[:, 0], np.clip(pts[:, 1], 3, 106)] = 2
    res = _compare(gpu, label, mapping, mask, 1.0, order=2)
    sel, _, stats = fmm_ref.fmm_lengths(label, mapping, mask, 2, 1.0, order=2)
    n_failed = sum(1 for v in stats.values() if v[0] is None)
    print("failed second-order marches", n_failed, "of", len(stats), flush=True)
    assert n_failed > 0 and int(res.failed.sum()) == n_failed
