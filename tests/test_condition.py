"""Pins of the DTM-conditioning oracle (oracle/condition_ref.py) on the CPU.

The blockage maps are computed twice, by scipy's generic_filter running
literal restatements of the reference's callbacks (its own mechanism,
preprocess.py:647-698) and by an independent vectorised restatement; both
must agree on DTMs with ties, NaNs and edges.  Known answers: a single
diagonal outflow and the codes of its upstream neighbours, derived by hand
from preprocess.py:390-473; a sink field whose centre cells are loops and
whose far cells are not; the lowest-neighbour rule with its tie order; and a
plane, which has no blockages or loops, so its conditioned uv is the plain
uv of slm_amd.preprocess."""
import numpy as np
import pytest

from oracle import condition_ref as C


def _rough(n=40, m=33, seed=0, nan=True):
    rs = np.random.default_rng(seed)
    h = np.round(rs.standard_normal((n, m)) * 3).astype(np.float32)      # many ties
    if nan:
        h[:2, :] = np.nan
        h[5, 7] = np.nan
    return h


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_blockage_maps_two_restatements_agree(seed):
    h = _rough(seed=seed)
    b1, n1 = C.blockage_codes_filter(h)
    b2, n2 = C.blockage_codes(h)
    assert np.array_equal(b1, b2) and np.array_equal(n1, n2)
    assert b1.any() and n1.any()
    # blockage codes only hold diagonal bits
    assert not (b1 & 15).any()


def test_single_diagonal_outflow_known_answer():
    h = np.full((7, 7), 10.0, np.float32)
    h[3, 3] = 5.0
    h[2, 2] = 1.0              # the only lower neighbour of (3,3) is SW
    b, n = C.blockage_codes_filter(h)
    # pixels whose only lower neighbours are diagonal ones: (3,3) -> (2,2) SW;
    # around (2,2): (1,1) NE, (1,3) SE, (3,1) NW; around (3,3): (2,4) SE,
    # (4,2) NW, (4,4) SW.  (2,3) and (3,2) see (2,2) cardinally: no blockage.
    want = {(3, 3): 16, (1, 1): 128, (1, 3): 32, (3, 1): 64, (2, 4): 32, (4, 2): 64, (4, 4): 16}
    got = {tuple(int(c) for c in ij): int(b[tuple(ij)]) for ij in zip(*np.nonzero(b))}
    assert got == want
    # upstream of the SW blockage at (3,3), per upstream_of_diagonal_outflow:
    # N of it sees it at S (code 4), NE sees it at SW (16), E sees it at W (1)
    assert n[3, 4] == 4 and n[4, 4] == 16 and n[4, 3] == 1
    assert np.array_equal(n, C.blockage_codes(h)[1])


def test_reflect_border():
    """At the grid edge the missing neighbour is the edge pixel itself
    (scipy 'reflect'): a corner whose only lower neighbour is diagonal."""
    h = np.full((5, 5), 10.0, np.float32)
    h[1, 1] = 0.0
    b, _ = C.blockage_codes(h)
    assert b[0, 0] == 128        # NE of (0,0) is (1,1)
    assert np.array_equal(b, C.blockage_codes_filter(h)[0])


def test_loops_of_a_sink():
    n = 21
    x, y = np.meshgrid(np.arange(n) - 10.0, np.arange(n) - 10.0, indexing="ij")
    r = np.hypot(x, y)
    r[10, 10] = 1.0
    u = (-x / r).astype(np.float32)
    v = (-y / r).astype(np.float32)
    u[10, 10] = v[10, 10] = 0.0
    loops = C.loop_cells(u, v, 0.95, -0.5, 0.0)
    assert loops.shape == (n - 1, n - 1)
    assert loops[9:11, 9:11].all()                 # the four cells around the sink
    assert not loops[:5, :].any() and not loops[-5:, :].any()
    # uniform flow: no loops
    assert not C.loop_cells(np.ones((5, 5), np.float32), np.zeros((5, 5), np.float32), 0.95, -0.5, 0.0).any()


def test_break_out_of_loop_lowest_and_ties():
    h = np.full((5, 5), 9.0, np.float32)
    u = np.zeros((5, 5), np.float32)
    v = np.zeros((5, 5), np.float32)
    h[2, 3] = 1.0                                  # N of (2,2)
    C.break_out_of_loop(2, 2, u, v, h)
    assert (u[2, 2], v[2, 2]) == (0.0, 1.0)
    h[:] = 9.0
    h[1, 2] = h[3, 2] = 1.0                        # W and E tie: W (first in order) wins
    C.break_out_of_loop(2, 2, u, v, h)
    assert (u[2, 2], v[2, 2]) == (-1.0, 0.0)
    h[:] = 9.0
    h[3, 1] = 1.0                                  # SE: diagonal, unit length in float64
    C.break_out_of_loop(2, 2, u, v, h)
    assert (u[2, 2], v[2, 2]) == (np.float32(1 / np.sqrt(2)), np.float32(-1 / np.sqrt(2)))
    u0 = u.copy()
    C.break_out_of_loop(0, 2, u, v, h)             # not well inside the grid: unchanged
    assert np.array_equal(u, u0)


def test_plane_conditioned_equals_plain():
    from slm_amd import preprocess
    n = 30
    x, y = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    h = (-(0.3 * x + 0.1 * y)).astype(np.float32)
    uv, m, b, nb, loops = C.conditioned_uv(h, 2.0, 2)
    assert not b.any() and not nb.any() and not loops.any()
    uv0, m0 = preprocess.gradient_uv(h, 2.0, 2)
    assert np.array_equal(uv, uv0) and np.array_equal(m, m0)


def test_blockage_fix_values():
    """A pit at (2,2): blockage pixels point along their diagonal
    (1/sqrt2 each way, then renormalised to unit length)."""
    h = np.full((7, 7), 10.0, np.float32)
    h[3, 3] = 5.0
    h[2, 2] = 1.0
    uv, m, b, nb, loops = C.conditioned_uv(h, 1.0, 0)
    s = np.float32(1 / np.sqrt(2))
    assert tuple(uv[3, 3]) == (-s, -s)             # SW, unit length after renormalising
    assert tuple(uv[3, 4]) == (0.0, -1.0)          # blocked neighbour N of it: flows S
