"""Pins of the surrogate's baselines (tests/baselines.py) on the CPU: D8
accumulation on planes where every path is known, the threshold sweep
against brute force, block means by hand."""
import numpy as np

from baselines import best_f1, block_mean, d8_bounded_accumulation, d8_receivers


def test_d8_on_a_plane_counts_the_upslope_column():
    """h = row index: every path runs straight down its column to row 0 (a
    drop of 1 per pixel beats 1/sqrt(2) on the diagonals), so pixel (i, j)
    counts the pixels i..i+L of its column, L = min(max_len, rows above)."""
    nx, ny, max_len = 30, 7, 10
    h = np.repeat(np.arange(nx, dtype=np.float32)[:, None], ny, axis=1)
    recv, step = d8_receivers(h)
    ii, jj = np.mgrid[:nx, :ny]
    want = np.where(ii > 0, (ii - 1) * ny + jj, ii * ny + jj).reshape(-1)
    assert np.array_equal(recv, want)
    assert np.array_equal(step, (ii > 0).reshape(-1).astype(float))
    c = d8_bounded_accumulation(h, max_len)
    assert np.array_equal(c, np.minimum(nx - 1 - ii, max_len) + 1)


def test_d8_diagonal_plane_and_no_data():
    """h = i + j: the steepest neighbour is the diagonal (drop 2/sqrt(2)), each
    step sqrt(2) long, so max_len 5 allows 3 steps along a diagonal line.  A
    pixel without data starts no path and takes none: its upslope diagonal
    neighbour steps straight instead (drop 1, the first of the tied two in D8
    order), and the first row steps along the row."""
    n = 12
    ii, jj = np.mgrid[:n, :n]
    h = (ii + jj).astype(np.float32)
    h[6, 6] = np.nan
    recv, step = d8_receivers(h)
    assert recv[4 * n + 4] == 3 * n + 3 and step[4 * n + 4] == np.sqrt(2)
    assert recv[7 * n + 7] == 6 * n + 7 and step[7 * n + 7] == 1
    assert recv[5] == 4 and step[5] == 1 and recv[0] == 0 and step[0] == 0
    c = d8_bounded_accumulation(h, 5.0)
    assert c[6, 6] == 0
    assert c[3, 3] == 3                       # (3, 3), (4, 4), (5, 5); (6, 6) has no data
    assert c[2, 2] == 4                       # (2, 2) .. (5, 5)
    assert c[8, 8] == 4                       # (8, 8) .. (11, 11)


def test_best_f1_matches_brute_force():
    rs = np.random.default_rng(3)
    for _ in range(20):
        score = rs.integers(0, 12, 300).astype(float)
        truth = rs.random(300) < 0.3 + 0.03 * score
        f, t = best_f1(score, truth)
        brute = max(2 * (p & truth).sum() / (p.sum() + truth.sum())
                    for p in (score >= u for u in np.unique(score)))
        assert abs(f - brute) < 1e-12
        p = score >= t
        assert abs(2 * (p & truth).sum() / (p.sum() + truth.sum()) - f) < 1e-12


def test_block_mean_by_hand():
    h = np.arange(20, dtype=np.float32).reshape(4, 5)
    h[3, 3] = np.nan
    b = block_mean(h)
    assert b.shape == (2, 2)
    assert b[0, 0] == (0 + 1 + 5 + 6) / 4 and b[0, 1] == (2 + 3 + 7 + 8) / 4
    assert b[1, 0] == (10 + 11 + 15 + 16) / 4 and np.isnan(b[1, 1])
