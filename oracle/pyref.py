"""Pure-Python restatement of the fields and trajectory kernels, for tiny grids.

TEST INFRASTRUCTURE ONLY.  It exists to pin the C oracle (oracle/slm_oracle.c):
it is written separately, straight from the reference's OpenCL C, with numpy
float32 scalars (every operation rounded to float32, as OpenCL float is) and
an FMA rounded exactly from rational arithmetic.  Agreement of the two on
small grids, bit for bit, is the evidence that the C oracle restates the
reference; it is slow (pure Python loops) and only meant for a few hundred
streamlines.

Sources restated:
  rng.cl:37-45, :62-67         lehmer_rand_uint, lehmer_rand_vec
  integratefields.cl:60-114    seeding, sub-pixel loop
  jittertrajectory.cl:42-88    loop, including the else that binds to
                               `if (idx<NXY_PADDED)` (always true)
  computestep.cl:42-57, :86-106, essentials.cl:24-129
  rungekutta.cl:49-77, :104-120, :163-186
  updatetraj.cl:40-55, :77-86, :165-172, writearray.cl:31-40, :70-82
  trajectory.cl:41-80, integratetraj.cl:54-97
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

f32 = np.float32
M = 0xfffffffb


def fma32(a, b, c) -> np.float32:
    """float32 fma(a,b,c), correctly rounded (ties to even) via exact rationals."""
    a, b, c = f32(a), f32(b), f32(c)
    if not (np.isfinite(a) and np.isfinite(b) and np.isfinite(c)):
        return f32(np.float64(a) * np.float64(b) + np.float64(c))
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    g = f32(float(exact))  # within one float32 ulp of exact
    best = None
    for cand in (np.nextafter(g, f32(-np.inf)), g, np.nextafter(g, f32(np.inf))):
        if not np.isfinite(cand):
            continue
        err = abs(Fraction(float(cand)) - exact)
        if best is None or err < best[0] or (err == best[0] and (cand.view(np.uint32) & 1) == 0):
            best = (err, cand)
    r = best[1]
    if r == 0:
        # sign of an exact zero sum follows IEEE: +0 unless both addends are -0
        prod_neg = (np.signbit(a) != np.signbit(b))
        r = f32(-0.0) if (exact == 0 and prod_neg and np.signbit(c)) else f32(0.0)
    return f32(r)


class K:
    """Kernel constants from an SlmParams-like object."""

    def __init__(self, p):
        self.NXP = int(p.nx_padded)
        self.NYP = int(p.ny_padded)
        self.NXY = int(p.nxy_padded)
        self.PAD = f32(p.pad_width)
        self.PP5 = f32(p.pad_width_pp5)
        self.NXF_MP5 = f32(p.nxf_mp5)
        self.NYF_MP5 = f32(p.nyf_mp5)
        self.GRID_SCALE = f32(p.grid_scale)
        self.COMBO = f32(p.combo_factor)
        self.DT_MAX = f32(p.dt_max)
        self.ADJ = f32(p.adjusted_max_error)
        self.HALT = f32(p.integration_halt_threshold)
        self.MAX_LENGTH = f32(p.max_length)
        self.MAX_N_STEPS = int(p.max_n_steps)
        self.RES = f32(p.trajectory_resolution)
        self.D = int(p.subpixel_seed_point_density)
        self.HALFSPAN = f32(p.subpixel_seed_halfspan)
        self.STEP = f32(p.subpixel_seed_step)
        self.J = f32(p.jitter_magnitude)
        self.SIGN = f32(p.downup_sign)
        self.N = int(p.n_seed_points)
        self.OFF = int(p.seeds_chunk_offset)
        self.contract = bool(p.fp_contract)
        self.wrap = bool(getattr(p, "approx_wrap", 0))

    # a*b + c as clang forms it: fused under FP_CONTRACT ON
    def muladd(self, a, b, c):
        return fma32(a, b, c) if self.contract else f32(f32(f32(a) * f32(b)) + f32(c))


def lehmer(state: int):
    s = ((state + 1) & 0xffffffff) * 279470273 % M
    return s, s


def rand_vec(state: int):
    a, state = lehmer(state)
    b, state = lehmer(state)
    two32 = f32(float(0xfffffffb))  # (float)0xfffffffb
    return (f32(f32(f32(a) / two32) - f32(0.5)), f32(f32(f32(b) / two32) - f32(0.5))), state


def byte_reverse(v: int) -> int:
    return int.from_bytes(int(v).to_bytes(4, "little"), "big")


def rng_seed(seed_idx: int) -> int:
    s = byte_reverse(seed_idx)
    s, _ = lehmer(s)
    return byte_reverse(s)


def _u(x) -> int:
    """(uint)max(0.0f, x) with x finite and >= 0 (truncation)."""
    x = max(f32(0.0), f32(x)) if not np.isnan(x) else f32(0.0)
    return int(np.floor(np.float64(x)))


def array_index(k: K, x, y) -> int:
    return min(k.NYP - 1, _u(f32(y + k.PP5))) + k.NYP * min(k.NXP - 1, _u(f32(x + k.PP5)))


def to_char(v) -> int:
    v = f32(v)
    if np.isnan(v):
        return 0
    i = int(np.trunc(np.float64(v)))
    i = max(-(1 << 31), min((1 << 31) - 1, i))
    return ((i & 0xff) ^ 0x80) - 0x80


def approximate(k: K, v):
    return f32(f32(to_char(f32(v * k.RES))) / k.RES)


def fast_length(k: K, x, y):
    return f32(np.sqrt(k.muladd(x, x, f32(y * y))))


def interpolate(k: K, uv, x, y):
    x_lft = min(k.NXP - 1, _u(f32(x + k.PP5)))
    y_dwn = min(k.NYP - 1, _u(f32(y + k.PP5)))
    x_rgt = min(k.NXP - 1, x_lft + 1)
    y_upp = min(k.NYP - 1, y_dwn + 1)
    rx = f32(f32(x - f32(x_lft)) + k.PAD)
    lx = f32(f32(1.0) - rx)
    uy = f32(f32(y - f32(y_dwn)) + k.PAD)
    dy = f32(f32(1.0) - uy)
    g = lambda i, j, c: f32(uv[i, j, c])
    dwn = [k.muladd(g(x_lft, y_dwn, c), lx, f32(g(x_rgt, y_dwn, c) * rx)) for c in (0, 1)]
    upp = [k.muladd(g(x_lft, y_upp, c), lx, f32(g(x_rgt, y_upp, c) * rx)) for c in (0, 1)]
    s = [k.muladd(dwn[c], dy, f32(upp[c] * uy)) for c in (0, 1)]
    with np.errstate(divide="ignore", invalid="ignore"):
        r = f32(f32(1.0) / f32(np.sqrt(k.muladd(s[0], s[0], f32(s[1] * s[1])))))
        return f32(s[0] * r), f32(s[1] * r)


def step(k: K, uv, dt, x, y, rng, jitter: bool):
    """compute_step_vec[_jittered]: returns uv1, dxy1, dxy2, next, idx, rng."""
    u1 = interpolate(k, uv, x, y)
    with np.errstate(divide="ignore", invalid="ignore"):
        if jitter:
            r, rng = rand_vec(rng)
            u1 = (k.muladd(r[0], k.J, u1[0]), k.muladd(r[1], k.J, u1[1]))
            ln = fast_length(k, u1[0], u1[1])
            u1 = (f32(u1[0] / ln), f32(u1[1] / ln))
        d1 = (approximate(k, f32(f32(u1[0] * k.COMBO) * dt)), approximate(k, f32(f32(u1[1] * k.COMBO) * dt)))
        u2 = interpolate(k, uv, f32(x + d1[0]), f32(y + d1[1]))
        if jitter:
            r, rng = rand_vec(rng)
            u2 = (k.muladd(r[0], k.J, u2[0]), k.muladd(r[1], k.J, u2[1]))
            ln = fast_length(k, u2[0], u2[1])
            u2 = (f32(u2[0] / ln), f32(u2[1] / ln))
        d2 = tuple(approximate(k, f32(f32(0.5) * k.muladd(f32(u2[c] * k.COMBO), dt, d1[c]))) for c in (0, 1))
    nxt = (f32(x + d2[0]), f32(y + d2[1]))
    return u1, d1, d2, nxt, array_index(k, nxt[0], nxt[1]), rng


def _dt_update(k: K, dt, err):
    if err == f32(0.0):
        return k.DT_MAX
    with np.errstate(divide="ignore", invalid="ignore"):
        return f32(min(k.DT_MAX, f32(f32(k.ADJ * dt) / err))) if not np.isnan(f32(f32(k.ADJ * dt) / err)) else k.DT_MAX


def fields(p, seeds, mask2d, uv3d):
    """integrate_fields for all seeds: returns (slc, slt) uint32 [nxp*nyp], steps."""
    k = K(p)
    mask = mask2d.reshape(-1).astype(bool)
    slc = np.zeros(k.NXY, dtype=np.uint64)
    slt = np.zeros(k.NXY, dtype=np.uint64)
    total = 0

    def write(idx, l):
        slt[idx] += int(np.floor(np.float64(f32(l + f32(0.5)))))
        slc[idx] += 1

    for gid in range(0, k.N - k.OFF):
        seed_idx = k.OFF + gid
        sx, sy = f32(seeds[seed_idx, 0]), f32(seeds[seed_idx, 1])
        rng0 = rng_seed(seed_idx)
        for j in range(k.D):
            for i in range(k.D):
                ox = k.muladd(f32(i), k.STEP, -k.HALFSPAN)
                oy = k.muladd(f32(j), k.STEP, -k.HALFSPAN)
                x, y = f32(sx + ox), f32(sy + oy)
                rng, n, l, dt = rng0, 0, f32(0.0), k.DT_MAX
                idx = array_index(k, x, y)
                if not mask[idx]:
                    write(idx, l)
                while idx < k.NXY and not mask[idx] and (l < k.MAX_LENGTH and n < k.MAX_N_STEPS):
                    u1, d1, d2, nxt, idx, rng = step(k, uv3d, dt, x, y, rng, True)
                    if idx < k.NXY:
                        if not mask[idx]:
                            with np.errstate(divide="ignore", invalid="ignore"):
                                err = fast_length(k, f32((d2[0] - d1[0]) / k.GRID_SCALE),
                                                  f32((d2[1] - d1[1]) / k.GRID_SCALE))
                            dl = fast_length(k, d2[0], d2[1])
                            x, y = nxt
                            stuck = dl < k.HALT
                            if not stuck:
                                dt = _dt_update(k, dt, err)
                            l = f32(l + dl)
                            n += 1
                            idx = array_index(k, x, y)
                            if idx < k.NXY and not mask[idx]:
                                write(idx, l)
                            if stuck:
                                break
                    else:  # never reached: idx is clamped into the grid
                        raise AssertionError("index outside the padded grid")
                total += n
    return (slc & 0xffffffff).astype(np.uint32), (slt & 0xffffffff).astype(np.uint32), total


def paths(p, seeds, mask2d, uv3d):
    """integrate_trajectory: returns (traj int8 [n, max_n_steps, 2], nsteps, length)."""
    k = K(p)
    mask = mask2d.reshape(-1).astype(bool)
    ns = k.N - k.OFF
    traj = np.zeros((ns, k.MAX_N_STEPS, 2), dtype=np.int8)
    nsteps = np.zeros(ns, dtype=np.uint16)
    length = np.zeros(ns, dtype=np.float32)
    for gid in range(ns):
        x, y = f32(seeds[k.OFF + gid, 0]), f32(seeds[k.OFF + gid, 1])
        px, py = x, y
        n, l, dt = 0, f32(0.0), k.DT_MAX
        idx = array_index(k, x, y)
        while not mask[idx] and (l < k.MAX_LENGTH and n < k.MAX_N_STEPS):
            u1, d1, d2, nxt, idx, _ = step(k, uv3d, dt, x, y, 0, False)
            if not mask[idx]:
                with np.errstate(divide="ignore", invalid="ignore"):
                    err = fast_length(k, f32((d2[0] - d1[0]) / k.GRID_SCALE), f32((d2[1] - d1[1]) / k.GRID_SCALE))
                dl = fast_length(k, d2[0], d2[1])
                if dl < k.HALT:
                    nxt = (f32(x + d1[0]), f32(y + d1[1]))
                    dl = fast_length(k, d1[0], d1[1])
                x, y = nxt
                idx = array_index(k, x, y)
                if dl < k.HALT:
                    l = f32(l + dl)
                    traj[gid, n] = (to_char(f32(f32(x - px) * k.RES)), to_char(f32(f32(y - py) * k.RES)))
                    n += 1
                    break
                dt = _dt_update(k, dt, err)
                l = f32(l + dl)
                traj[gid, n] = (to_char(f32(f32(x - px) * k.RES)), to_char(f32(f32(y - py) * k.RES)))
                n += 1
                px, py = x, y
            else:
                su, sv = f32(u1[0] * k.SIGN), f32(u1[1] * k.SIGN)

                def edge(xx, u):
                    d = f32(0.0)
                    with np.errstate(divide="ignore", invalid="ignore"):
                        if u > 0:
                            d = f32(f32(f32(np.trunc(np.float64(f32(xx + f32(1.5))))) - f32(xx + f32(0.5))) / u)
                        if u < 0:
                            d = f32(-f32(f32(xx + f32(0.5)) - f32(np.trunc(np.float64(f32(xx + f32(0.5)))))) / u)
                    return d
                tx, ty = edge(x, su), edge(y, sv)
                if abs(tx) < abs(ty):
                    dt = tx
                elif abs(ty) < abs(tx):
                    dt = ty
                else:
                    dt = f32(min(tx, ty))
                x, y = k.muladd(su, dt, x), k.muladd(sv, dt, y)
                if k.wrap:
                    x, y = approximate(k, x), approximate(k, y)
                else:
                    # (char) out of range keeps the int32 truncation: the lattice value
                    x = f32(f32(np.trunc(np.float64(f32(x * k.RES)))) / k.RES)
                    y = f32(f32(np.trunc(np.float64(f32(y * k.RES)))) / k.RES)
                x = f32(min(max(x, f32(-0.5)), k.NXF_MP5))
                y = f32(min(max(y, f32(-0.5)), k.NYF_MP5))
                dl = fast_length(k, f32(x - px), f32(y - py))
                l = f32(l + dl)
                traj[gid, n] = (to_char(f32(f32(x - px) * k.RES)), to_char(f32(f32(y - py) * k.RES)))
                n += 1
                break
        nsteps[gid] = n
        length[gid] = l
    return traj, nsteps, length


def hillslope_lengths(p, seeds, mask2d, uv3d, mapping2d, thin_flag, from_flag):
    """hillslope_lengths (lengths.cl:37-116) with lengths_runge_kutta_step
    (rungekutta.cl:492-516); unreached seeds stay 0."""
    k = K(p)
    mask = mask2d.reshape(-1).astype(bool)
    mapping = mapping2d.reshape(-1).astype(np.int64)
    out = np.zeros(k.N, dtype=np.float32)
    for gid in range(k.OFF, k.N):
        x, y = f32(seeds[gid, 0]), f32(seeds[gid, 1])
        n, l, dt, moved_off = 0, f32(0.0), k.DT_MAX, False
        idx = array_index(k, x, y)
        abandoned = False
        while ((~mapping[idx]) & thin_flag) and n < k.MAX_N_STEPS:
            u1, d1, d2, nxt, idx, _ = step(k, uv3d, dt, x, y, 0, False)
            if mask[idx]:
                abandoned = True
                break
            if not moved_off and ((~mapping[idx]) & from_flag):
                moved_off = True
            elif moved_off and (mapping[idx] & from_flag):
                abandoned = True
                break
            with np.errstate(divide="ignore", invalid="ignore"):
                err = fast_length(k, f32((d2[0] - d1[0]) / k.GRID_SCALE), f32((d2[1] - d1[1]) / k.GRID_SCALE))
            dl = fast_length(k, d2[0], d2[1])
            if dl < k.HALT:
                nxt = (f32(x + d1[0]), f32(y + d1[1]))
                dl = fast_length(k, d1[0], d1[1])
            x, y = nxt
            idx = array_index(k, x, y)
            if dl < k.HALT:
                abandoned = True
                break
            dt = _dt_update(k, dt, err)
            l = f32(l + dl)
            n += 1
        if not abandoned and (mapping[idx] & thin_flag):
            out[gid] = f32(l * f32(p.pixel_size))
    return out
