"""Pins of the eikonal (FMM) oracle behind the FMM hillslope lengths
(oracle/slm_oracle_fmm.c, oracle/fmm_ref.py), on the CPU.

scikit-fmm is absent, so the oracle is a first-order restatement of the
fast marching method; these tests pin it by closed forms and invariants:
the update formula's cases, a straight channel (exact distances k - 1/2),
a single channel pixel (the hand-computed diagonal value), masked barriers,
1-Lipschitz continuity and the symmetry of the grid."""
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
