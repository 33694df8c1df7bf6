"""CPU restatement of scikit-image's medial_axis (2-D), for the channel
thinning skeletonize(medial_axis(channel_array)) of Mapping.thin_channels
(python/streamlines/mapping.py:707-719).

TEST INFRASTRUCTURE ONLY: the checker of slm_thin_channels' medial-axis step
(slm_amd/csrc/mapping.hip).  scikit-image is absent here; this follows its
published algorithm (skimage.morphology.medial_axis, the versions the
reference's requirements name, >= 0.13), built on the library routines it
calls (scipy.ndimage.distance_transform_edt and label, numpy's legacy
RandomState):
  1. a 512-entry table over 3x3 patterns (bit k of the index = pixel k of
     the pattern in row-major order, centre = bit 4): keep the centre when it
     is set and either removing it changes the number of 8-connected
     components of the pattern, or the pattern has fewer than 3 set pixels;
  2. the Euclidean distance of every foreground pixel to the background;
  3. a corner score per pixel: 9 minus the set pixels of its 3x3 window
     (outside the image counts as unset);
  4. a tie-breaker: RandomState(0).permutation of the foreground count,
     assigned to the foreground pixels in row-major order (the era's
     deterministic choice; later scikit-image versions draw it from a
     fresh generator, so their skeletons are not reproducible call to
     call);
  5. the foreground pixels visited once each in lexicographic order of
     (distance, corner score, tie-breaker); a visited pixel takes the table's
     value for its current 3x3 window (pixels outside the image count as
     unset, pixels not yet visited as they started).
Parity with scikit-image itself is unpinned (absent); the table and the
visit rule are pinned by hand-computed cases in tests/test_medial_axis.py.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

EIGHT = ndimage.generate_binary_structure(2, 2)


def pattern_of(index: int) -> np.ndarray:
    """The 3x3 pattern of a table index (bit k = row k // 3, column k % 3)."""
    return np.array([[index & 2 ** k != 0 for k in range(r * 3, r * 3 + 3)] for r in range(3)], bool)


def keep_table() -> np.ndarray:
    t = np.zeros(512, np.uint8)
    for index in range(512):
        if not index & 16:
            continue
        pat = pattern_of(index)
        n_with = ndimage.label(pat, EIGHT)[1]
        n_without = ndimage.label(pattern_of(index & ~16), EIGHT)[1]
        t[index] = 1 if (n_with != n_without or pat.sum() < 3) else 0
    return t


def window_index(img: np.ndarray) -> np.ndarray:
    """Table index of every pixel's 3x3 window (zero outside the image)."""
    a = np.pad(img.astype(np.int64), 1)
    nx, ny = img.shape
    idx = np.zeros((nx, ny), np.int64)
    k = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            idx += a[1 + dx:1 + dx + nx, 1 + dy:1 + dy + ny] << k
            k += 1
    return idx


def tiebreaker(n: int) -> np.ndarray:
    return np.random.RandomState(0).permutation(np.arange(n))


def medial_axis(image: np.ndarray) -> np.ndarray:
    img = np.asarray(image, bool)
    table = keep_table()
    corner = 9 - np.array([pattern_of(i).sum() for i in range(512)])[window_index(img)]
    d2 = ndimage.distance_transform_edt(img) ** 2 if img.any() else np.zeros(img.shape)
    fg = np.flatnonzero(img)
    tie = tiebreaker(fg.size)
    order = np.lexsort((tie, corner.reshape(-1)[fg], np.rint(d2.reshape(-1)[fg])))
    result = img.astype(np.uint8).copy()
    nx, ny = img.shape
    for o in order:
        p = fg[o]
        x, y = divmod(int(p), ny)
        acc = 16
        k = 0
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                if (dx or dy) and 0 <= x + dx < nx and 0 <= y + dy < ny and result[x + dx, y + dy]:
                    acc |= 1 << k
                k += 1
        result[x, y] = table[acc]
    return result.astype(bool)


def squared_edt(image: np.ndarray) -> np.ndarray:
    """Squared Euclidean distance to the nearest unset pixel, as integers."""
    img = np.asarray(image, bool)
    return np.rint(ndimage.distance_transform_edt(img) ** 2).astype(np.int64)
