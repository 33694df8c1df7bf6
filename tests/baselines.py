"""Channel maps made without the surrogate, as the baselines that show what
the trained network gains: the cheaper computations a user without it would
run on the same DTM.  Test infrastructure; the surrogate has no counterpart
in the reference.

  d8_bounded_accumulation   D8 flow paths capped at the streamline length
  block_mean                a 2x coarser DTM (for a coarser trace)
  f1                        F1 of a boolean map against the truth
  best_f1                   the highest F1 any threshold of a score reaches
"""
import numpy as np

D8 = ((-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1))


def d8_receivers(h):
    """(receiver, step) per pixel, flat: the neighbour of steepest descent
    (height drop per unit distance, over the 8 neighbours with data, the
    first in D8 order on a tie) and the distance to it in pixels; a pixel
    with no lower neighbour, or no data, is its own receiver (step 0)."""
    nx, ny = h.shape
    h = h.astype(np.float64)
    hp = np.pad(h, 1, constant_values=np.nan)
    ii, jj = np.mgrid[:nx, :ny]
    best = np.zeros((nx, ny))
    recv = ii * ny + jj
    step = np.zeros((nx, ny))
    for di, dj in D8:
        d = float(np.hypot(di, dj))
        drop = (h - hp[1 + di:1 + di + nx, 1 + dj:1 + dj + ny]) / d
        better = np.nan_to_num(drop, nan=-np.inf) > best
        best = np.where(better, drop, best)
        recv = np.where(better, (ii + di) * ny + (jj + dj), recv)
        step = np.where(better, d, step)
    return recv.reshape(-1), step.reshape(-1)


def d8_bounded_accumulation(h, max_len):
    """Number of D8 paths through each pixel: every pixel with data starts a
    path that follows the receivers until the next step would take its length
    past max_len (pixels) or it reaches a pixel without a lower neighbour;
    each pixel on a path (its start included) counts that path once.  The D8
    counterpart of the downstream streamline density with one seed per pixel
    and streamlines capped at max_len.  Paths descend strictly, so none
    returns to a pixel."""
    recv, step = d8_receivers(h)
    n = h.size
    pos = np.flatnonzero(np.isfinite(h.reshape(-1)))
    count = np.bincount(pos, minlength=n)
    length = np.zeros(pos.size)
    while pos.size:
        nxt, s = recv[pos], step[pos]
        go = (nxt != pos) & (length + s <= max_len)
        pos, length = nxt[go], length[go] + s[go]
        count += np.bincount(pos, minlength=n)
    return count.reshape(h.shape)


def block_mean(h, f=2):
    """f x f block means of h (trailing rows/columns that fill no block are
    dropped); a block with a pixel without data (NaN) has none."""
    nx, ny = h.shape[0] // f * f, h.shape[1] // f * f
    b = h[:nx, :ny].astype(np.float64).reshape(nx // f, f, ny // f, f)
    return b.mean(axis=(1, 3)).astype(np.float32)


def f1(pred, truth):
    """F1 of the boolean map pred against the boolean truth."""
    pred, truth = np.asarray(pred, bool), np.asarray(truth, bool)
    tp = int((pred & truth).sum())
    return 2.0 * tp / max(1, int(pred.sum()) + int(truth.sum()))


def best_f1(score, truth):
    """(F1, threshold): the highest F1 of (score >= threshold) against the
    boolean truth over every threshold."""
    score = np.asarray(score, np.float64).reshape(-1)
    truth = np.asarray(truth, bool).reshape(-1)
    order = np.argsort(-score, kind="stable")
    s, t = score[order], truth[order]
    tp = np.cumsum(t)
    k = np.arange(1, s.size + 1)
    f1 = 2.0 * tp / (k + t.sum())
    cut = np.r_[s[1:] != s[:-1], True]          # a threshold keeps every equal score
    f1 = np.where(cut, f1, -1.0)
    i = int(np.argmax(f1))
    return float(f1[i]), float(s[i])
