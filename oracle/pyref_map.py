"""Pure-Python restatement of the channel-mapping and segmentation kernels.

TEST INFRASTRUCTURE ONLY.  It pins the C oracle (oracle/slm_oracle_map.c):
each function follows one reference kernel line by line, running the work
items one after another in a given order (``order``, default seed order) --
one legal schedule of the reference.  Where the reference's result is the
same for every schedule (the tests check this by running two orders), the C
oracle must agree with it bit for bit.  Float arithmetic is numpy float32
scalars via oracle/pyref.py; use small grids.

Sources restated (opencl/):
  connect.cl:26-108            connect_channels (+ rungekutta.cl:280-309)
  channelheads.cl:32-93        map_channel_heads (+ rungekutta.cl:335-353)
  countlink.cl:34-114, :134-210  count_downchannels, flag_downchannels
  label.cl:21-134              label_confluences
  segment.cl:38-107, :133-188, :195-334, :360-414, :441-504, :530-594
  hillslopes.cl:31-91          hillslopes (link_hillslopes)
"""
from __future__ import annotations

import numpy as np

from .pyref import K, array_index, fast_length, f32, step, _dt_update

IS_CHANNEL, IS_THINCHANNEL, IS_INTERCHANNEL, IS_CHANNELHEAD, IS_CHANNELTAIL = 1, 2, 4, 8, 16
IS_MAJORCONFLUENCE, IS_MAJORINFLOW, IS_MINORINFLOW = 32, 128, 256
IS_LEFTFLANK, IS_SUBSEGMENTHEAD = 512, 16384
LFA = 0x80000000
U32 = 0xffffffff


def _order(k: K, order):
    return list(range(k.OFF, k.N)) if order is None else list(order)


def _rk(k: K, uv, dt, x, y):
    """compute_step_vec then the shared head of the mapping RK steps
    (rungekutta.cl:280-459): returns (x, y, idx, dl, err, probe_idx)."""
    _, d1, d2, nxt, probe, _ = step(k, uv, dt, x, y, 0, False)
    with np.errstate(divide="ignore", invalid="ignore"):
        err = fast_length(k, f32((d2[0] - d1[0]) / k.GRID_SCALE), f32((d2[1] - d1[1]) / k.GRID_SCALE))
    dl = fast_length(k, d2[0], d2[1])
    if dl < k.HALT:
        nxt = (f32(x + d1[0]), f32(y + d1[1]))
        dl = fast_length(k, d1[0], d1[1])
    return nxt[0], nxt[1], array_index(k, nxt[0], nxt[1]), dl, err, probe


def _char(v) -> int:
    v = f32(v)
    if np.isnan(v):
        return 0
    i = int(np.trunc(np.float64(v)))
    i = max(-(1 << 31), min((1 << 31) - 1, i))
    return ((i & 0xff) ^ 0x80) - 0x80


def connect_channels(p, seeds, mask, uv, mapping, order=None):
    k = K(p)
    IMAX = int(p.interchannel_max_n_steps)
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        px, py = x, y
        dt = k.DT_MAX
        idx = array_index(k, x, y)
        prev_idx = idx
        n = 0
        traj = {}

        def body():
            nonlocal x, y, px, py, dt, idx, n
            x2, y2, idx2, dl, err, probe = _rk(k, uv, dt, x, y)
            idx = probe
            if not mask[idx]:
                x, y, idx = x2, y2, idx2
                traj[n] = (_char(f32(f32(x - px) * k.RES)), _char(f32(f32(y - py) * k.RES)))
                n += 1
                if dl < k.HALT:
                    return
                dt = _dt_update(k, dt, err)
                px, py = x, y

        # first loop; its record index is bounded here at IMAX (the reference
        # writes past its private array instead, and then flags nothing)
        while prev_idx == idx and not mask[idx] and n != k.MAX_N_STEPS and n < IMAX:
            body()
        while (mapping[idx] & IS_CHANNEL) == 0 and not mask[idx] and n != IMAX and n < IMAX:
            body()
        if 2 < n < IMAX:
            vx, vy = f32(seeds[g, 0]), f32(seeds[g, 1])
            idx = array_index(k, vx, vy)
            s = 0
            while not mask[idx] and s < n - 1:
                if mapping[idx] == 0:
                    mapping[idx] |= IS_INTERCHANNEL
                vx = f32(vx + f32(f32(traj[s][0]) / k.RES))
                vy = f32(vy + f32(f32(traj[s][1]) / k.RES))
                idx = array_index(k, vx, vy)
                s += 1


def map_channel_heads(p, seeds, mask, uv, mapping, order=None):
    k = K(p)
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        dt, n = k.DT_MAX, 0
        idx = array_index(k, x, y)
        prev_idx = idx
        while prev_idx == idx and not mask[idx] and n < k.MAX_N_STEPS:
            x, y, idx, dl, err, _ = _rk(k, uv, dt, x, y)
            n += 1
            dt = _dt_update(k, dt, err)
        if n >= k.MAX_N_STEPS:
            continue
        if not mask[idx]:
            idx = array_index(k, x, y)
            if ((~mapping[idx]) & IS_THINCHANNEL) or (mapping[prev_idx] & IS_THINCHANNEL):
                mapping[idx] &= ~IS_CHANNELHEAD & U32


def count_downchannels(p, seeds, mask, uv, mapping, count, link, order=None):
    k = K(p)
    cap = (1000 * k.MAX_N_STEPS) & U32
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        dt, n, counter = k.DT_MAX, 0, 1
        idx = array_index(k, x, y)
        prev_idx = idx
        count[idx] = counter
        counter += 1
        while not mask[idx] and n < cap:
            x, y, idx, dl, err, _ = _rk(k, uv, dt, x, y)
            if dl < k.HALT:
                break
            dt = _dt_update(k, dt, err)
            n += 1
            if prev_idx != idx:
                mapping[idx] &= ~IS_CHANNELHEAD & U32
                mapping[idx] |= IS_THINCHANNEL
                link[prev_idx] = idx
                counter += 1
                if counter < count[idx]:
                    break
                count[idx] = counter
                prev_idx = idx
        link[prev_idx] = idx


def flag_downchannels(p, seeds, mask, mapping, count, link, order=None):
    k = K(p)
    cap = (1000 * k.MAX_N_STEPS) & U32
    for g in _order(k, order):
        head = array_index(k, f32(seeds[g, 0]), f32(seeds[g, 1]))
        idx = head
        prev_idx = idx + 1
        counter = 1
        count[idx] = counter
        mapping[idx] |= IS_THINCHANNEL
        while not mask[idx] and prev_idx != idx and counter < cap:
            prev_idx = idx
            idx = int(link[idx])
            counter += 1
            if not mask[idx]:
                if ((mapping[idx] & IS_MINORINFLOW) > 0 or (counter == 2 and count[idx] > 2)) and counter < 20:
                    mapping[head] &= ~IS_CHANNELHEAD & U32
                mapping[idx] |= IS_THINCHANNEL
                if counter >= count[idx]:
                    count[idx] = counter
            else:
                break
        if not mask[prev_idx] and counter < cap:
            mapping[prev_idx] |= IS_CHANNELTAIL


def label_confluences(p, seeds, mask, slt, mapping, count, link, order=None):
    k = K(p)
    for g in _order(k, order):
        vx, vy = f32(seeds[g, 0]), f32(seeds[g, 1])
        idx = array_index(k, vx, vy)
        inflows = []
        for dx, dy in ((0, 1), (0, -1), (1, 0), (-1, 0), (1, 1), (1, -1), (-1, 1), (-1, -1)):
            nbr = array_index(k, f32(vx + f32(dx)), f32(vy + f32(dy)))
            if not mask[nbr] and (mapping[nbr] & IS_THINCHANNEL) and link[nbr] == idx:
                inflows.append(nbr)
        if len(inflows) > 1:
            mapping[idx] |= IS_MAJORCONFLUENCE
            equal, dom_idx, dom = [], 0, f32(-3.4028234663852886e38)
            for i in inflows:
                if int(count[idx]) != (int(count[i]) + 1) & U32:
                    mapping[i] |= IS_MINORINFLOW
                else:
                    equal.append(i)
                    if f32(slt[i]) > dom:
                        dom_idx, dom = i, f32(slt[i])
            for i in equal:
                mapping[i] |= IS_MAJORINFLOW if i == dom_idx else IS_MINORINFLOW


def segment_downchannels(p, seeds, mask, mapping, count, link, label, order=None):
    k = K(p)
    T = int(p.segmentation_threshold)
    for g in _order(k, order):
        prev_idx = array_index(k, f32(seeds[g, 0]), f32(seeds[g, 1]))
        segment_label = prev_idx
        seg_counter = 0
        label[prev_idx] = segment_label
        mapping[prev_idx] |= IS_SUBSEGMENTHEAD
        idx = int(link[prev_idx])
        while not mask[idx] and prev_idx != idx and (mapping[idx] & IS_SUBSEGMENTHEAD) == 0:
            if int(count[idx]) >= seg_counter + T // 2:
                segment_label = idx
                seg_counter = int(count[idx]) + T
                mapping[idx] |= IS_SUBSEGMENTHEAD
            label[idx] = segment_label
            prev_idx = idx
            idx = int(link[idx])


def link_hillslopes(p, seeds, mask, uv, link, order=None):
    k = K(p)
    cap = (100 * k.MAX_N_STEPS) & U32
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        dt, n = k.DT_MAX, 0
        idx = array_index(k, x, y)
        prev_idx = idx
        while prev_idx == idx and not mask[idx] and n < cap:
            prev_idx = idx
            x, y, idx, dl, err, _ = _rk(k, uv, dt, x, y)
            if dl < k.HALT:
                break
            dt = _dt_update(k, dt, err)
            n += 1
        if prev_idx != idx and not mask[idx]:
            link[prev_idx] = idx


def segment_hillslopes(p, seeds, mask, uv, mapping, label, order=None):
    k = K(p)
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        dt, n = k.DT_MAX, 0
        idx = array_index(k, x, y)
        hs = idx
        while not mask[idx] and ((~mapping[idx]) & IS_THINCHANNEL) and n < k.MAX_N_STEPS:
            x, y, idx, dl, err, _ = _rk(k, uv, dt, x, y)
            n += 1
            if dl < k.HALT:
                break
            dt = _dt_update(k, dt, err)
        if mapping[idx] & IS_THINCHANNEL:
            label[hs] = label[idx]


def _rotate_left(k: K, px, py, dx, dy):
    def c8(v):
        return ((v & 0xff) ^ 0x80) - 0x80
    rx, ry = c8(dx - dy), c8(dx + dy)
    ax, ay = c8(abs(rx) & 0xff), c8(abs(ry) & 0xff)   # abs(char) is a uchar
    ax, ay = min(2, max(1, ax)), min(2, max(1, ay))
    ndx = c8(int(rx / ax))                               # C division truncates
    ndy = c8(int(ry / ay))
    return (((px + ndx) & U32) * k.NYP + ((py + ndy) & U32)) & U32, ndx, ndy


def subsegment_channel_edges(p, seeds, mask, mapping, link, label, order=None):
    k = K(p)
    masked = lambda i: True if i >= k.NXY else bool(mask[i])
    for g in _order(k, order):
        prev_idx = array_index(k, f32(seeds[g, 0]), f32(seeds[g, 1]))
        idx = int(link[prev_idx])
        n = 0
        while prev_idx != idx:
            ok = n < k.MAX_N_STEPS
            n += 1
            if not ok:
                break
            px, py = prev_idx // k.NYP, prev_idx % k.NYP
            x, y = idx // k.NYP, idx % k.NYP
            c8 = lambda v: ((v & 0xff) ^ 0x80) - 0x80
            dx, dy = c8(x - px), c8(y - py)
            left, turns, do_left = idx, 0, True
            while left != prev_idx:
                turns += 1
                if turns > 7:
                    break
                left, dx, dy = _rotate_left(k, px, py, dx, dy)
                if turns >= 2 and not masked(left) and (mapping[left] & IS_THINCHANNEL):
                    if left == link[idx]:
                        do_left = False
                    break
            if do_left:
                left, turns = idx, 0
                dx, dy = c8(x - px), c8(y - py)
                while left != prev_idx:
                    turns += 1
                    if turns > 4:
                        break
                    left, dx, dy = _rotate_left(k, px, py, dx, dy)
                    if turns >= 2 and not masked(left):
                        if mapping[left] & IS_THINCHANNEL:
                            break
                        mapping[left] |= IS_LEFTFLANK
                        label[left] |= LFA
            prev_idx = idx
            idx = int(link[idx])
            if mask[idx] or (mapping[prev_idx] & IS_SUBSEGMENTHEAD):
                break


def subsegment_flanks(p, seeds, mask, uv, mapping, label, order=None):
    k = K(p)
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        dt, n = k.DT_MAX, 0
        idx = array_index(k, x, y)
        hs = idx
        while (not mask[idx] and ((~mapping[idx]) & IS_LEFTFLANK) and ((~mapping[idx]) & IS_THINCHANNEL)
               and n < k.MAX_N_STEPS):
            x, y, idx, dl, err, _ = _rk(k, uv, dt, x, y)
            n += 1
            if dl < k.HALT:
                break
            dt = _dt_update(k, dt, err)
        if mapping[idx] & IS_LEFTFLANK:
            label[hs] |= LFA


def _fix_flanks(p, seeds, mask, uv, mapping, label, order, left):
    k = K(p)
    for g in _order(k, order):
        x, y = f32(seeds[g, 0]), f32(seeds[g, 1])
        dt, n = k.DT_MAX, 0
        idx = array_index(k, x, y)
        hs = idx

        def side(i):
            return (int(label[i]) & LFA) != 0 if left else (int(label[i]) & LFA) == 0

        while not mask[idx] and side(idx) and ((~mapping[idx]) & IS_THINCHANNEL) and n < k.MAX_N_STEPS:
            x2, y2, idx2, dl, err, probe = _rk(k, uv, dt, x, y)
            idx = probe
            if (int(label[hs]) & 0x7fffffff) != (int(label[idx]) & 0x7fffffff):
                label[hs] = label[idx]
            x, y, idx = x2, y2, idx2
            n += 1
            if dl < k.HALT:
                break
            dt = _dt_update(k, dt, err)
        if (int(label[hs]) & 0x7fffffff) != (int(label[idx]) & 0x7fffffff):
            label[hs] = label[idx]
        if left:
            if (int(label[idx]) & LFA) == 0 and ((~mapping[idx]) & IS_THINCHANNEL):
                label[hs] = int(label[hs]) & 0x7fffffff
        else:
            if int(label[idx]) & LFA:
                label[hs] |= LFA


def fix_right_flanks(p, seeds, mask, uv, mapping, label, order=None):
    _fix_flanks(p, seeds, mask, uv, mapping, label, order, False)


def fix_left_flanks(p, seeds, mask, uv, mapping, label, order=None):
    _fix_flanks(p, seeds, mask, uv, mapping, label, order, True)
