"""Pins of the eikonal (FMM) oracle behind the FMM hillslope lengths
(oracle/slm_oracle_fmm.c, oracle/fmm_ref.py), on the CPU.

scikit-fmm is absent, so the oracle restates the published fast marching
method, first order and scikit-fmm's second-order update; these tests pin
it by closed forms and invariants: the update formula's cases, a straight
channel (exact distances k - 1/2, both orders), a diagonal channel (a plane
front at 45 degrees: exact in both orders), a single channel pixel (the
hand-computed values, including the second-order update across the zero
contour), masked barriers, 1-Lipschitz continuity and the symmetry of the
grid."""
import numpy as np
import pytest

from oracle import fmm_ref


def test_update_cases():
    assert fmm_ref.fmm_update(0.5, np.inf, 1.0) == 1.5               # one axis
    assert fmm_ref.fmm_update(3.0, 1.0, 1.0) == 2.0                  # b - a >= dx: one-sided
    assert fmm_ref.fmm_update(0.5, 0.5, 1.0) == (1.0 + np.sqrt(2.0)) * 0.5
    a, b = 0.2, 0.9
    u = fmm_ref.fmm_update(a, b, 1.0)
    assert abs((u - a) ** 2 + (u - b) ** 2 - 1.0) < 1e-14              # the discrete eikonal equation


@pytest.mark.parametrize("dx", [1.0, 2.0])
def test_straight_channel_exact(dx):
    n, m, c = 21, 17, 9
    phi = np.ones((n, m), np.int8)
    phi[c, :] = -1
    dom = np.ones((n, m), bool)
    d = fmm_ref.fmm_distance(phi, dom, dx)
    x = np.arange(n)[:, None].repeat(m, 1)
    want = (np.abs(x - c) - 0.5) * dx
    want[c, :] = -0.5 * dx
    assert np.array_equal(d, want)


def test_single_channel_pixel():
    phi = np.ones((9, 9), np.int8)
    phi[4, 4] = -1
    d = fmm_ref.fmm_distance(phi, np.ones((9, 9), bool), 1.0)
    assert d[4, 4] == -1.0 / np.sqrt(8.0)                          # crossings on both axes: 1/sqrt(1/.25+1/.25)
    assert d[3, 4] == d[5, 4] == d[4, 3] == d[4, 5] == 0.5
    assert d[3, 3] == (1.0 + np.sqrt(2.0)) * 0.5                    # from 0.5 and 0.5
    assert d[2, 4] == 1.5
    assert np.array_equal(d, d.T) and np.array_equal(d, d[::-1, :]) and np.array_equal(d, d[:, ::-1])


def test_masked_barrier_and_lipschitz():
    rs = np.random.default_rng(0)
    n = 40
    phi = np.ones((n, n), np.int8)
    phi[rs.integers(0, n, 12), rs.integers(0, n, 12)] = -1
    dom = np.ones((n, n), bool)
    dom[5:35, 20] = False                                           # a wall
    d = fmm_ref.fmm_distance(phi, dom, 1.0)
    assert np.isnan(d[~dom]).all() and not np.isnan(d[dom]).any()
    pos = np.where(dom, np.abs(d), np.nan)
    for ax in (0, 1):
        g = np.abs(np.diff(pos, axis=ax))
        assert np.nanmax(g) <= 1.0 + 1e-12                          # neighbours differ by at most dx
    # a point behind the wall is farther than straight-line distance allows
    one = np.ones((n, n), np.int8)
    one[20, 5] = -1
    dw = fmm_ref.fmm_distance(one, dom, 1.0)
    dn = fmm_ref.fmm_distance(one, np.ones((n, n), bool), 1.0)
    assert dw[20, 30] > dn[20, 30] + 5.0


def test_fmm_lengths_two_subsegments():
    """Two rectangular subsegments side by side, each with its channel along
    one edge: hsl = 2 * mean distance, truncated; the ratio rule holds."""
    n, m = 30, 20
    label = np.zeros((n, m), np.int32)
    label[1:15, 1:19] = 7
    label[15:29, 1:19] = -9
    mapping = np.zeros((n, m), np.uint32)
    mapping[1, 1:19] = 2                                            # channel of 7 along its first row
    mapping[28, 1:19] = 2                                           # channel of -9 along its last row
    mask = label == 0
    sel, hsl, stats = fmm_ref.fmm_lengths(label, mapping, mask, 2, 1.0)
    assert list(sel) == [-9, 7]
    for s, h in zip(sel, hsl):
        q, cnt, n_phi, n_ch = stats[abs(int(s))]
        assert n_phi == 14 * 18 and n_ch == 18
        assert h == np.int32(2.0 * (q / fmm_ref.FIX) / cnt) and h > 0
    grid = fmm_ref.remap(label, sel, hsl)
    assert grid[5, 5] == hsl[1] and grid[20, 5] == hsl[0] and grid[0, 0] == 0


def test_colouring_separates_pairs():
    from slm_amd import hsl
    rs = np.random.default_rng(1)
    n_labels = 60
    problems = np.arange(1, n_labels, dtype=np.int64)
    a = rs.integers(1, n_labels, 300)
    b = rs.integers(1, n_labels, 300)
    keep = a != b
    lo, hi = np.minimum(a, b)[keep], np.maximum(a, b)[keep]
    pairs = (lo.astype(np.int64) << 32) | hi.astype(np.int64)
    colour, n_col = hsl.colour_labels(n_labels, problems, pairs)
    assert colour[0] == 0xFF and (colour[1:] < n_col).all()
    assert (colour[lo] != colour[hi]).all()


def test_rank_filters_closed_forms():
    """The oracle's skimage rank mean / median restatements on images whose
    results are known: a constant image; a lone spike (the median removes it,
    the mean spreads it as spike / footprint size); edges count only in-image
    pixels; the footprint disk(r) = {x^2 + y^2 <= r^2}."""
    from oracle import hsl_map_ref as H
    assert len(H.disk_offsets(1)) == 5 and len(H.disk_offsets(2)) == 13 and len(H.disk_offsets(3)) == 29
    c = np.full((9, 11), 77, np.uint16)
    assert np.array_equal(H.rank_mean(c, 3), c)
    assert np.array_equal(H.rank_median(c.astype(np.uint8), 2), c.astype(np.uint8))
    sp = np.zeros((15, 15), np.uint16)
    sp[7, 7] = 1300
    m = H.rank_mean(sp, 2)
    assert m[7, 7] == 1300 // 13 and m[7, 9] == 1300 // 13 and m[7, 10] == 0
    assert not H.rank_median(sp.astype(np.uint8) * 0 + (sp > 0).astype(np.uint8) * 200, 1).any()
    # corner: disk(1) has 3 in-image pixels at (0, 0)
    e = np.zeros((5, 5), np.uint16)
    e[0, 1] = 9
    assert H.rank_mean(e, 1)[0, 0] == 3
    # median: smallest v with 2 * #(<= v) >= pop; pop = 5 inside: values 1,2,3,4,5 -> 3
    g = np.zeros((3, 3), np.uint8)
    g[1, 1], g[0, 1], g[2, 1], g[1, 0], g[1, 2] = 3, 1, 5, 2, 4
    assert H.rank_median(g, 1)[1, 1] == 3


def _aspect_case(n_bins=60, north_extra=0.0, seed=0):
    rs = np.random.default_rng(seed)
    nx = ny = 64
    aspect = rs.uniform(-np.pi, np.pi, (nx, ny)).astype(np.float32)
    hsl = (50.0 + rs.normal(0, 5, (nx, ny))).astype(np.float32)
    hsl = np.where(np.sin(aspect) > 0, hsl + north_extra, hsl).astype(np.float32)
    return aspect, np.zeros((nx, ny), bool), hsl, np.zeros((nx, ny), bool)


def test_combine_hsl_aspect_bins():
    from slm_amd import hsl as H
    aspect, amask, h, mask = _aspect_case()
    df, means, haa = H.combine_hsl_aspect(aspect, amask, h, mask, pad=2, n_bins=60)
    assert haa.shape == (61, 2)
    assert np.isclose(np.rad2deg(haa[0, 1]), -180.0) and np.isclose(np.rad2deg(haa[-1, 1]), 180.0)
    assert haa[0, 0] == haa[-1, 0]                                   # +-180 degrees is one direction
    # a middle bin's mean by direct binning
    deg = np.rad2deg(aspect[2:-2, 2:-2]).astype(np.float32).ravel()
    hv = h[2:-2, 2:-2].ravel()
    sel = (deg > -3.0) & (deg <= 3.0) & (hv >= 5.0)
    assert np.isclose(haa[30, 0], hv[sel].astype(np.float64).mean())
    assert np.all(np.diff(df.aspect.values) >= 0)


def test_combine_hsl_aspect_tensor_inputs():
    """Grids given as torch tensors are selected where they live; the pairs,
    bin means and statistics equal those of the numpy path, with masked
    pixels, NaN lengths and the all-masked fallback."""
    import torch
    from slm_amd import hsl as H
    aspect, amask, h, mask = _aspect_case(north_extra=5.0, seed=3)
    rs = np.random.default_rng(4)
    amask = rs.random(amask.shape) < 0.2
    mask = rs.random(mask.shape) < 0.1
    h = np.where(rs.random(h.shape) < 0.05, np.nan, h).astype(np.float32)
    for m in (mask, np.ones_like(mask)):
        df0, me0, haa0 = H.combine_hsl_aspect(aspect, amask, h, m, pad=2)
        t = [torch.from_numpy(np.ascontiguousarray(x)) for x in (aspect, amask, h, m)]
        df1, me1, haa1 = H.combine_hsl_aspect(*t, pad=2)
        assert df0.equals(df1) and me0.equals(me1) and np.array_equal(haa0, haa1, equal_nan=True)
    s0, s1 = H.hsl_aspect_stats(haa0, df0), H.hsl_aspect_stats(haa1, df1)
    assert all(np.array_equal(np.asarray(s0[k]), np.asarray(s1[k])) for k in s0)


def test_hsl_aspect_stats_disparity():
    from slm_amd import hsl as H

    def stats(extra):
        df, _, haa = H.combine_hsl_aspect(*_aspect_case(north_extra=extra), pad=2)
        return H.hsl_aspect_stats(haa, df)

    even, skew = stats(0.0), stats(20.0)
    assert even["hsl_ns_disparity"] < 2.0
    assert skew["hsl_mean_north"] > skew["hsl_mean_south"] + 15.0
    assert skew["hsl_ns_disparity_confidence"] > even["hsl_ns_disparity_confidence"]
    assert 60.0 < skew["hsl_mean_azimuth"] < 120.0                  # the mean vector points north
    assert skew["hsl_ns_ks"][1] < 1e-6 and skew["hsl_ns_pp_diff"] > 50


@pytest.mark.parametrize("dx", [1.0, 2.0])
def test_second_order_straight_channel_exact(dx):
    n, m, c = 21, 17, 9
    phi = np.ones((n, m), np.int8)
    phi[c, :] = -1
    d = fmm_ref.fmm_distance2(phi, np.ones((n, m), bool), dx, order=2)
    x = np.arange(n)[:, None].repeat(m, 1)
    want = (np.abs(x - c) - 0.5) * dx
    want[c, :] = -0.5 * dx
    assert np.abs(d - want).max() < 1e-12 * dx


def test_diagonal_front_exact_both_orders():
    """Channel pixels on the diagonal: the zero contour is the line y = x + 1/2
    (and its mirror), so the distance of (x, y) is (|y - x| - 1/2) / sqrt(2)."""
    n = 24
    phi = np.ones((n, n), np.int8)
    phi[np.arange(n), np.arange(n)] = -1
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    k = np.abs(Y - X)
    want = (k - 0.5) / np.sqrt(2.0)
    inner = (k >= 1) & (X >= 3) & (X < n - 3) & (Y >= 3) & (Y < n - 3)
    for order in (1, 2):
        d = fmm_ref.fmm_distance2(phi, np.ones((n, n), bool), 1.0, order=order)
        assert np.abs(d[inner] - want[inner]).max() < 1e-12, order


def test_second_order_single_pixel_by_hand():
    """One channel pixel at (4, 4): the interface points carry 1/2 and the
    pixel -1/sqrt(8).  (2, 4) is updated when (3, 4) is accepted: along x, v1
    = 1/2 and the point beyond, (4, 4), lies across the contour, so v2 =
    -1/sqrt(8) and t = (4 v1 - v2)/3; along y nothing is frozen: u = t + 2/3.
    (3, 3) takes first-order terms on both axes (the points beyond are not
    frozen yet): (1 + sqrt(2))/2."""
    phi = np.ones((9, 9), np.int8)
    phi[4, 4] = -1
    d = fmm_ref.fmm_distance2(phi, np.ones((9, 9), bool), 1.0, order=2)
    assert d[4, 4] == -1.0 / np.sqrt(8.0)
    assert d[3, 4] == 0.5
    t = (1.0 / 3.0) * (4.0 * 0.5 + 1.0 / np.sqrt(8.0))
    assert abs(d[2, 4] - (t + 2.0 / 3.0)) < 1e-15
    assert abs(d[3, 3] - (1.0 + np.sqrt(2.0)) * 0.5) < 1e-15
    assert np.array_equal(d, d.T) and np.array_equal(d, d[::-1, :]) and np.array_equal(d, d[:, ::-1])


def _scattered(seed, n=48):
    rs = np.random.default_rng(seed)
    phi = np.ones((n, n), np.int8)
    phi[rs.integers(4, n - 4, 10), rs.integers(4, n - 4, 10)] = -1
    return phi, np.ones((n, n), bool)


def test_second_order_close_to_first_and_lipschitz():
    phi, dom = _scattered(2)
    d1 = fmm_ref.fmm_distance2(phi, dom, 1.0, order=1)
    d2 = fmm_ref.fmm_distance2(phi, dom, 1.0, order=2)
    assert np.array_equal(np.sign(d1), np.sign(d2))
    assert np.abs(d2 - d1).max() < 1.0
    for ax in (0, 1):
        assert np.abs(np.diff(np.abs(d2), axis=ax)).max() <= 1.0 + 1e-9


def test_second_order_negative_discriminant_raises():
    """Scattered single channel pixels: where two fronts meet, the
    second-order terms of the two axes can leave the quadratic without a
    real root; scikit-fmm raises there (the reference then leaves the
    subsegment unmeasured), and so does the oracle."""
    phi, dom = _scattered(0)
    with pytest.raises(fmm_ref.NegativeDiscriminant):
        fmm_ref.fmm_distance2(phi, dom, 1.0, order=2)
    fmm_ref.fmm_distance2(phi, dom, 1.0, order=1)      # first order always has a root
