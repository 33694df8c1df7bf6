"""Pins of the medial-axis restatement (oracle/medial_axis_ref.py) behind the
channel thinning skeletonize(medial_axis(.)) of Mapping.thin_channels
(mapping.py:707-719), on the CPU.

scikit-image is absent; its published algorithm is restated and pinned here:
table entries worked out by hand, the distance transform against brute
force, and shapes whose medial axes are known (a 1-pixel line is kept
whole; an odd-width bar keeps its centre row between the end caps; a digital
disk keeps its radii to its four tips), plus the invariants of the visit rule
(a subset, one pixel wide, the 8-connected component count kept)."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import medial_axis_ref as M


def _idx(*cells):
    """Table index of a pattern given as (row, col) cells of the 3x3 window."""
    return sum(1 << (r * 3 + c) for r, c in cells)


def test_keep_table_by_hand():
    t = M.keep_table()
    assert t.shape == (512,) and t.dtype == np.uint8
    assert not t[: 16].any() and all(t[i] == 0 for i in range(512) if not i & 16)   # centre unset: never kept
    assert t[_idx((1, 1))] == 1                          # isolated pixel: fewer than 3 set
    assert t[_idx((1, 1), (0, 1))] == 1                  # end of a line
    assert t[_idx((1, 1), (0, 1), (2, 1))] == 1          # middle of a line: removal splits it
    assert t[_idx((1, 1), (0, 0), (2, 2))] == 1          # diagonal line: splits too
    assert t[_idx((1, 1), (0, 1), (1, 2))] == 0          # N and E touch diagonally: removable
    assert t[511] == 0                                   # full window: removable
    assert t[_idx((1, 1), (0, 0), (0, 1), (0, 2))] == 0  # a lump on one side
    assert t[_idx((1, 1), (0, 1), (2, 1), (1, 0))] == 0  # T-junction arm: the rest stays joined


def test_squared_edt_brute_force():
    rs = np.random.default_rng(0)
    img = rs.random((23, 31)) < 0.8
    d2 = M.squared_edt(img)
    bx, by = np.nonzero(~img)
    for x in range(img.shape[0]):
        for y in range(img.shape[1]):
            want = 0 if not img[x, y] else int(((bx - x) ** 2 + (by - y) ** 2).min())
            assert d2[x, y] == want


def test_line_kept_whole():
    img = np.zeros((9, 20), bool)
    img[4, 3:17] = True
    assert np.array_equal(M.medial_axis(img), img)
    diag = np.eye(12, dtype=bool)
    assert np.array_equal(M.medial_axis(diag), diag)


@pytest.mark.parametrize("w", [3, 5, 7])
def test_bar_reduces_to_centre_row(w):
    img = np.zeros((w + 6, 40), bool)
    img[3:3 + w, 5:35] = True
    ma = M.medial_axis(img)
    c = 3 + w // 2
    assert ma[c, 5 + w // 2:35 - w // 2].all()                  # the centre row between the end caps
    off_axis = ma.copy()
    off_axis[c, :] = False
    # outside the centre row only the end-cap branches toward the corners remain
    assert not off_axis[:, 5 + w:35 - w].any()
    _check_invariants(img, ma)


def test_disk_radii_to_its_tips():
    """A digital disk of radius 15 has a one-pixel tip on each axis, a corner
    of its boundary: the axis keeps the four radii to the tips, which close
    around the centre pixel (removable, the four arms stay 8-connected)."""
    n = 41
    X, Y = np.mgrid[:n, :n]
    img = (X - 20) ** 2 + (Y - 20) ** 2 <= 15 ** 2
    ma = M.medial_axis(img)
    cross = np.zeros_like(img)
    cross[5:36, 20] = cross[20, 5:36] = True
    cross[20, 20] = False
    assert np.array_equal(ma, cross)
    _check_invariants(img, ma)


def _check_invariants(img, ma):
    assert not (ma & ~img).any()
    # one pixel wide: no 2x2 block of skeleton
    assert not (ma[:-1, :-1] & ma[1:, :-1] & ma[:-1, 1:] & ma[1:, 1:]).any()
    n_img = ndimage.label(img, M.EIGHT)[1]
    assert ndimage.label(ma, M.EIGHT)[1] == n_img


def test_random_shapes_keep_topology():
    rs = np.random.default_rng(3)
    img = ndimage.binary_dilation(rs.random((60, 70)) < 0.01, iterations=3)
    ma = M.medial_axis(img)
    _check_invariants(img, ma)
    assert ma.sum() < img.sum() / 2
