// hsl.hip -- FMM hillslope lengths on the GPU.
//
// The reference measures hillslope lengths per subsegment with an eikonal
// distance (python/streamlines/mapping.py:841-900, Mapping.fmm_lengths): for
// each selected subsegment s, phi = 1 on the 4-dilation D_s of the pixels
// with |label| == |s|, -1 on its thin-channel pixels, masked elsewhere;
// distance = skfmm.distance(phi, dx); hsl = 2 * mean(distance).  It loops
// over thousands of subsegments, one full-grid fast march each.
//
// Here every subsegment is solved at once.  Two subsegments whose dilated
// domains share a pixel have pixels at most 2 apart (Manhattan); slm_fmm_pairs
// lists such label pairs and the host colours the labels so that the domains
// of one colour are disjoint.  Per colour, slm_fmm_hsl:
//   k_fmm_setup   each pixel's problem (its own label's, or the one whose
//                 domain its 4-neighbours put it in) and its phi sign;
//   k_fmm_init    points next to the zero contour get the interface distance
//                 1 / sqrt(sum over crossing axes of 1 / (dx/2)^2) (the
//                 linear-interpolation initialisation of scikit-fmm);
//   k_fmm_relax_tiles  the first-order upwind update u = F(min x-neighbour,
//                 min y-neighbour) applied in place until nothing decreases:
//                 the fixed point of the discrete eikonal system, which the
//                 fast marching method computes by accepting points in order
//                 (oracle/slm_oracle_fmm.c, the same F in double precision);
//   k_fmm_reduce  per problem: the signed distances summed in fixed point
//                 (round(d * 2^24), int64, so the sum is order-free), the
//                 point count, the channel-pixel and own-pixel counts.
// slm_hsl_remap writes each selected label's value onto its pixels
// (remap_hsl, mapping.py:953-985).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <mutex>
#include "../../include/slm_hip.h"
#include "common.hpp"

namespace {

typedef unsigned long long u64;
constexpr int kBlock = 256;
constexpr double kFix = 16777216.0;   // 2^24

// One colour's state in the workspace; a group of colours is solved together,
// slot g (blockIdx.y) at base + g * stride.
struct FmmLayout {
    uint8_t* base;
    size_t stride, off_neg, off_u, off_fixed, off_act0, off_act1;
    // second-order march: per-label point lists, heap positions and heaps;
    // the lists of subsegments marched with LDS heaps (counters, then two
    // lists of big_cap keys)
    size_t off_lcnt, off_lcur, off_pts, off_pos, off_heap, off_big, big_cap;
};
__device__ __forceinline__ uint8_t* slot_base(const FmmLayout& L) { return L.base + blockIdx.y * L.stride; }

__device__ __forceinline__ uint32_t abs_label(int32_t l) {
    // |label|; the most negative int32 (a flank bit with no label) is no label
    return (l == INT32_MIN) ? 0u : (uint32_t)(l < 0 ? -l : l);
}

// Label pairs whose dilated domains can share a pixel: pixels at Manhattan
// distance 1 or 2 with different non-zero |label|.  Half of the offsets, so
// each unordered pixel pair is seen once; duplicates of a label pair remain.
__global__ __launch_bounds__(256) void k_fmm_pairs(const int32_t* __restrict__ label, uint32_t nx, uint32_t ny,
                                                   const uint8_t* __restrict__ colour, uint32_t n_labels,
                                                   u64* __restrict__ pairs, uint32_t cap,
                                                   uint32_t* __restrict__ count) {
    // slots are handed out per block: LDS counter, then one global atomic
    __shared__ uint32_t s_cnt, s_base;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t a = 0;
    int x = 0, y = 0;
    if (i < nx * ny) {
        a = abs_label(label[i]);
        if (a >= n_labels || colour[a] == 0xffu) a = 0;
        x = (int)(i / ny);
        y = (int)(i % ny);
    }
    const int off[6][2] = {{0, 1}, {0, 2}, {1, -1}, {1, 0}, {1, 1}, {2, 0}};
    u64 key[6];
    uint32_t slot[6];
    uint32_t has = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        key[k] = 0;
        slot[k] = 0;
        bool h = false;
        if (a) {
            const int xx = x + off[k][0], yy = y + off[k][1];
            if (xx >= 0 && yy >= 0 && xx < (int)nx && yy < (int)ny) {
                const uint32_t b = abs_label(label[(size_t)xx * ny + yy]);
                if (b && b != a && b < n_labels && colour[b] != 0xffu) {
                    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
                    key[k] = ((u64)lo << 32) | hi;
                    h = true;
                }
            }
        }
        const u64 bal = __ballot(h);
        if (!bal) continue;
        const uint32_t leader = (uint32_t)__ffsll((long long)bal) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&s_cnt, (uint32_t)__popcll(bal));
        base = __shfl(base, (int)leader);
        if (h) {
            slot[k] = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            has |= 1u << k;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_cnt ? atomicAdd(count, s_cnt) : 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; k++)
        if ((has >> k) & 1u) {
            const uint32_t g = s_base + slot[k];
            if (g < cap) pairs[g] = key[k];
        }
}

// problem key (|label| + 0 = none) and phi sign (1 = negative) of colour c
__global__ void k_fmm_setup(const int32_t* __restrict__ label, const uint32_t* __restrict__ mapping,
                            uint32_t thin_flag, const uint8_t* __restrict__ colour, uint32_t n_labels, uint32_t c0,
                            uint32_t nx, uint32_t ny, FmmLayout L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const uint32_t c = c0 + blockIdx.y;
    uint32_t* __restrict__ dom = (uint32_t*)slot_base(L);
    uint8_t* __restrict__ neg = slot_base(L) + L.off_neg;
    const uint32_t a = abs_label(label[i]);
    uint32_t key = 0;
    if (a && a < n_labels && colour[a] == c) {
        key = a;
    } else {
        const int x = (int)(i / ny), y = (int)(i % ny);
        const int off[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        for (int k = 0; k < 4 && !key; k++) {
            const int xx = x + off[k][0], yy = y + off[k][1];
            if (xx < 0 || yy < 0 || xx >= (int)nx || yy >= (int)ny) continue;
            const uint32_t b = abs_label(label[(size_t)xx * ny + yy]);
            if (b && b < n_labels && colour[b] == c) key = b;
        }
    }
    dom[i] = key;
    neg[i] = (key && a == key && (mapping[i] & thin_flag) == thin_flag) ? 1 : 0;
}

__device__ __forceinline__ double fmm_update(double a, double b, double dx) {
    if (a > b) {
        const double t = a;
        a = b;
        b = t;
    }
    if (isinf(b) || b - a >= dx) return a + dx;
    const double d = a - b;
    return (a + b + sqrt(2.0 * dx * dx - d * d)) * 0.5;
}

__global__ void k_fmm_init(uint32_t nx, uint32_t ny, double dx, FmmLayout L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const uint32_t* __restrict__ dom = (const uint32_t*)slot_base(L);
    const uint8_t* __restrict__ neg = slot_base(L) + L.off_neg;
    double* __restrict__ u = (double*)(slot_base(L) + L.off_u);
    uint8_t* __restrict__ fixed = slot_base(L) + L.off_fixed;
    const uint32_t key = dom[i];
    double v = INFINITY;
    uint8_t f = 0;
    if (key) {
        const int x = (int)(i / ny), y = (int)(i % ny);
        const double pi = neg[i] ? -1.0 : 1.0;
        double inv = 0.0;
        bool any = false;
        for (int axis = 0; axis < 2; axis++) {
            double best = INFINITY;
            for (int s = -1; s <= 1; s += 2) {
                const int xx = axis == 0 ? x + s : x, yy = axis == 1 ? y + s : y;
                if (xx < 0 || yy < 0 || xx >= (int)nx || yy >= (int)ny) continue;
                const size_t j = (size_t)xx * ny + yy;
                if (dom[j] != key || neg[j] == neg[i]) continue;
                const double pj = neg[j] ? -1.0 : 1.0;
                const double fr = pi / (pi - pj) * dx;
                if (fr < best) best = fr;
            }
            if (best < INFINITY) {
                inv += 1.0 / (best * best);
                any = true;
            }
        }
        if (any) {
            v = 1.0 / sqrt(inv);
            f = 1;
        }
    }
    u[i] = v;
    fixed[i] = f;
}

// Relaxation in LDS tiles: a 256-thread block loads a 32x32 tile of u with a
// one-pixel halo, applies the update in place to its pixels until none
// decreases (the halo is held fixed), and writes back what decreased.  Values
// only decrease, so in-place updates, and tiles reading neighbours that
// another block is writing, read valid upper bounds: repeated launches
// converge to the same fixed point as plain sweeps, in far fewer launches.
constexpr int kT = 32;
constexpr int kTH = kT + 2;
constexpr int kMaxLocal = 4 * kT;

__global__ __launch_bounds__(256) void k_fmm_relax_tiles(uint32_t nx, uint32_t ny, double dx, FmmLayout L,
                                                         uint32_t parity, uint32_t* __restrict__ changed) {
    const uint32_t* __restrict__ dom = (const uint32_t*)slot_base(L);
    const uint8_t* __restrict__ fixed = slot_base(L) + L.off_fixed;
    double* __restrict__ u = (double*)(slot_base(L) + L.off_u);
    const uint8_t* __restrict__ act_in = slot_base(L) + (parity ? L.off_act1 : L.off_act0);
    uint8_t* __restrict__ act_out = slot_base(L) + (parity ? L.off_act0 : L.off_act1);
    __shared__ double su[kTH * kTH];
    __shared__ uint32_t sk[kTH * kTH];
    __shared__ uint8_t sfix[kTH * kTH];
    const uint32_t nty = (ny + kT - 1) / kT, ntx = (nx + kT - 1) / kT;
    const int tx = (int)(blockIdx.x / nty), ty = (int)(blockIdx.x % nty);
    const int x0 = tx * kT - 1, y0 = ty * kT - 1;
    if (threadIdx.x == 0) act_out[blockIdx.x] = 0;
    // a tile can only decrease if it or a neighbouring tile changed in the
    // previous launch (act_in; all tiles on the first launch)
    {
        bool need = false;
        if (threadIdx.x < 9) {
            const int ax = tx + (int)threadIdx.x / 3 - 1, ay = ty + (int)threadIdx.x % 3 - 1;
            if (ax >= 0 && ay >= 0 && ax < (int)ntx && ay < (int)nty) need = act_in[(size_t)ax * nty + ay] != 0;
        }
        if (!__syncthreads_or(need)) return;
    }
    bool any_key = false;
    for (int k = threadIdx.x; k < kTH * kTH; k += blockDim.x) {
        const int lx = k / kTH, ly = k % kTH, gx = x0 + lx, gy = y0 + ly;
        uint32_t key = 0;
        double v = INFINITY;
        uint8_t f = 1;
        if (gx >= 0 && gy >= 0 && gx < (int)nx && gy < (int)ny) {
            const size_t g = (size_t)gx * ny + gy;
            key = dom[g];
            v = __hip_atomic_load(&u[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool halo = lx == 0 || ly == 0 || lx == kTH - 1 || ly == kTH - 1;
            f = (halo || fixed[g]) ? 1 : 0;
            if (!halo && key && !f) any_key = true;
        }
        su[k] = v;
        sk[k] = key;
        sfix[k] = f;
    }
    if (!__syncthreads_or(any_key)) return;
    bool grew = false;
    for (int it = 0; it < kMaxLocal; it++) {
        bool dec = false;
        for (int k = threadIdx.x; k < kT * kT; k += blockDim.x) {
            const int c = (k / kT + 1) * kTH + (k % kT + 1);
            const uint32_t key = sk[c];
            if (!key || sfix[c]) continue;
            double m[2] = {INFINITY, INFINITY};
            const int nb[4] = {c - kTH, c + kTH, c - 1, c + 1};   // x-1, x+1, y-1, y+1
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (sk[nb[j]] != key) continue;
                const double q = __hip_atomic_load(&su[nb[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (q < m[j >> 1]) m[j >> 1] = q;
            }
            if (m[0] < INFINITY || m[1] < INFINITY) {
                const double v = fmm_update(m[0], m[1], dx);
                if (v < su[c]) {
                    __hip_atomic_store(&su[c], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    dec = true;
                }
            }
        }
        if (!__syncthreads_or(dec)) break;
        grew = true;
    }
    if (!grew) return;
    for (int k = threadIdx.x; k < kT * kT; k += blockDim.x) {
        const int lx = k / kT + 1, ly = k % kT + 1, gx = x0 + lx, gy = y0 + ly;
        const int c = lx * kTH + ly;
        if (gx >= (int)nx || gy >= (int)ny || !sk[c] || sfix[c]) continue;
        const size_t g = (size_t)gx * ny + gy;
        if (su[c] < __hip_atomic_load(&u[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            __hip_atomic_store(&u[g], su[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) {
        act_out[blockIdx.x] = 1;
        atomicAdd(changed, 1u);
    }
}

// ---------------------------------------------------------------- second order
//
// scikit-fmm's default (distanceMarcher::updatePointOrderTwo, as published):
// per axis the smaller frozen neighbour v1 and, when the point beyond it in
// the same direction is frozen with a value not above v1, that value v2
// (negated when it lies across the zero contour from the neighbour or from
// the point) give (9/4)(u - t)^2/dx^2 with t = (4 v1 - v2)/3, otherwise
// (u - v1)^2/dx^2; u is the larger root of their sum = 1.  v2 is not reset
// when the other direction of the axis gives a smaller v1 without a usable
// second point (as published).  Which points are frozen when a point is
// updated depends on the order of acceptance, so the march runs in that
// order: one lane per subsegment (the subsegments of a colour are disjoint),
// a binary heap of (value, index) pairs in the workspace, a point's value
// replaced at each update.  Pops come in (value, index) order, which is the
// oracle's order, so the results are its bits.  A negative discriminant (on
// which scikit-fmm raises, and the reference then leaves the subsegment
// unmeasured) stops the lane and marks the subsegment failed.

struct HeapEnt {
    double v;
    int32_t i, pad;
};

__device__ __forceinline__ bool hless(const HeapEnt& a, const HeapEnt& b) {
    return a.v < b.v || (a.v == b.v && a.i < b.i);
}

__device__ void heap_up(HeapEnt* h, int32_t* pos, int k) {
    const HeapEnt e = h[k];
    while (k > 0) {
        const int p = (k - 1) >> 1;
        const HeapEnt hp = h[p];
        if (!hless(e, hp)) break;
        h[k] = hp;
        pos[hp.i] = k;
        k = p;
    }
    h[k] = e;
    pos[e.i] = k;
}

__device__ void heap_down(HeapEnt* h, int32_t* pos, int n, int k) {
    const HeapEnt e = h[k];
    for (;;) {
        int c = 2 * k + 1;
        if (c >= n) break;
        HeapEnt hc = h[c];
        if (c + 1 < n) {
            const HeapEnt h2 = h[c + 1];
            if (hless(h2, hc)) {
                hc = h2;
                c++;
            }
        }
        if (!hless(hc, e)) break;
        h[k] = hc;
        pos[hc.i] = k;
        k = c;
    }
    h[k] = e;
    pos[e.i] = k;
}

__global__ void k_fmm_lcount(uint32_t nx, uint32_t ny, FmmLayout L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t key = i < nx * ny ? ((const uint32_t*)slot_base(L))[i] : 0u;
    wave_group_add(key, 0u, (uint32_t*)(slot_base(L) + L.off_lcnt));
}

// Exclusive scan of the per-label counts [0, n] in place (one 1024-thread
// block per colour slot); the cursor copy feeds the scatter.
__global__ __launch_bounds__(1024) void k_fmm_lscan(uint32_t n, FmmLayout L) {
    uint32_t* __restrict__ cnt = (uint32_t*)(slot_base(L) + L.off_lcnt);
    uint32_t* __restrict__ cur = (uint32_t*)(slot_base(L) + L.off_lcur);
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = min(n, t * per), e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += cnt[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = (t == 0) ? 0u : part[t - 1];
    for (uint32_t i = b; i < e; i++) {
        const uint32_t c = cnt[i];
        cnt[i] = run;
        cur[i] = run;
        run += c;
    }
}

// Each label's points in any order: the march sorts them by pixel index, and
// the lane kernel's heap pops in (value, pixel index) order whatever order
// they were pushed in.
__global__ void k_fmm_lscatter(uint32_t nx, uint32_t ny, FmmLayout L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t key = i < nx * ny ? ((const uint32_t*)slot_base(L))[i] : 0u;
    const uint32_t at = wave_group_add(key, 0u, (uint32_t*)(slot_base(L) + L.off_lcur));
    if (key) ((uint32_t*)(slot_base(L) + L.off_pts))[at] = i;
}

// The second-order update of point (x, y) of problem key (state >= 2 frozen).
// Every value it may read (the first and second neighbours along each axis,
// both sides) is loaded first, in one round of independent loads; the
// arithmetic then follows the march's rules: per axis the smaller first
// neighbour, the second neighbour only behind it and not larger, its sign
// flipped across the contour, and v2 kept from the first side when the
// second side improves v1 alone.
__device__ double fmm_update2(const uint32_t* __restrict__ dom, const uint8_t* __restrict__ neg,
                              const uint8_t* __restrict__ state, const double* __restrict__ u, uint32_t key, int x,
                              int y, int nx, int ny, double dx, bool* bad) {
    const double idx2 = 1.0 / (dx * dx), aa = 9.0 / 4.0, third = 1.0 / 3.0;
    const size_t i = (size_t)x * ny + y;
    bool ok1[4], ok2[4], flip[4];
    double u1[4], u2[4];
    const uint8_t ni = neg[i];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int axis = q >> 1, sd = (q & 1) ? 1 : -1;
        const int x1 = axis == 0 ? x + sd : x, y1 = axis == 1 ? y + sd : y;
        const int x2 = axis == 0 ? x + 2 * sd : x, y2 = axis == 1 ? y + 2 * sd : y;
        const bool in1 = x1 >= 0 && y1 >= 0 && x1 < nx && y1 < ny;
        const bool in2 = x2 >= 0 && y2 >= 0 && x2 < nx && y2 < ny;
        const size_t n1 = in1 ? (size_t)x1 * ny + y1 : i, n2 = in2 ? (size_t)x2 * ny + y2 : i;
        const uint32_t d1 = dom[n1], d2 = dom[n2];
        const uint8_t s1 = state[n1], s2 = state[n2], e1 = neg[n1], e2 = neg[n2];
        u1[q] = u[n1];
        u2[q] = u[n2];
        ok1[q] = in1 && d1 == key && s1 >= 2;
        ok2[q] = in2 && d2 == key && s2 >= 2;
        flip[q] = e2 != e1 || e2 != ni;
    }
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int axis = 0; axis < 2; axis++) {
        double v1 = INFINITY, v2 = INFINITY;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int q = axis * 2 + side;
            if (!ok1[q] || !(u1[q] < v1)) continue;
            v1 = u1[q];
            if (!ok2[q] || !(u2[q] <= v1)) continue;
            v2 = flip[q] ? -u2[q] : u2[q];
        }
        if (v2 < INFINITY) {
            const double tp = third * (4.0 * v1 - v2);
            a += idx2 * aa;
            b -= idx2 * 2.0 * aa * tp;
            c += idx2 * aa * tp * tp;
        } else if (v1 < INFINITY) {
            a += idx2;
            b -= idx2 * 2.0 * v1;
            c += idx2 * v1 * v1;
        }
    }
    c -= 1.0;
    const double det = b * b - 4.0 * a * c;
    if (det < 0.0) {
        *bad = true;
        return INFINITY;
    }
    return (-b + sqrt(det)) / 2.0 / a;
}

// One lane per label of the slot's colour: the fast march of that subsegment.
// Subsegments with lds_lo < m <= lds_hi points are left to k_fmm_march_lds.
__global__ __launch_bounds__(64) void k_fmm_march(uint32_t nx, uint32_t ny, double dx, FmmLayout L,
                                                  const uint8_t* __restrict__ colour, uint32_t n_labels, uint32_t c0,
                                                  uint8_t* __restrict__ failed, uint32_t lds_lo, uint32_t lds_hi) {
    const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key == 0 || key >= n_labels || colour[key] != c0 + blockIdx.y) return;
    uint8_t* sb = slot_base(L);
    const uint32_t* __restrict__ dom = (const uint32_t*)sb;
    const uint8_t* __restrict__ neg = sb + L.off_neg;
    double* __restrict__ u = (double*)(sb + L.off_u);
    uint8_t* __restrict__ state = sb + L.off_fixed;    // k_fmm_init: 1 next to the contour
    const uint32_t* __restrict__ lst = (const uint32_t*)(sb + L.off_lcnt);
    const uint32_t beg = lst[key], end = lst[key + 1];
    if (beg == end || (end - beg > lds_lo && end - beg <= lds_hi)) return;
    const uint32_t* __restrict__ pts = (const uint32_t*)(sb + L.off_pts) + beg;
    int32_t* __restrict__ pos = (int32_t*)(sb + L.off_pos);
    HeapEnt* __restrict__ h = (HeapEnt*)(sb + L.off_heap) + beg;
    const uint32_t m = end - beg;
    int hn = 0;
    // states: 0 far, 1 trial, 2 next to the contour (frozen), 3 accepted
    for (uint32_t k = 0; k < m; k++) {
        const uint32_t p = pts[k];
        pos[p] = -1;
        if (state[p]) {
            state[p] = 2;
            h[hn].v = u[p];
            h[hn].i = (int32_t)p;
            heap_up(h, pos, hn);
            hn++;
        }
    }
    bool bad = false;
    while (hn > 0 && !bad) {
        const HeapEnt top = h[0];
        hn--;
        if (hn > 0) {
            h[0] = h[hn];
            heap_down(h, pos, hn, 0);
        }
        const int32_t i = top.i;
        pos[i] = -1;
        state[i] = 3;
        const int x = i / (int)ny, y = i % (int)ny;
        const int nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
        for (int k = 0; k < 4 && !bad; k++) {
            const int xx = nb[k][0], yy = nb[k][1];
            if (xx < 0 || yy < 0 || xx >= (int)nx || yy >= (int)ny) continue;
            const int32_t j = xx * (int)ny + yy;
            if (dom[j] != key || state[j] >= 2) continue;
            const double v = fmm_update2(dom, neg, state, u, key, xx, yy, (int)nx, (int)ny, dx, &bad);
            if (bad || v == 0.0) continue;     // scikit-fmm skips a zero update
            const double old = u[j];
            u[j] = v;
            state[j] = 1;
            if (pos[j] >= 0) {
                h[pos[j]].v = v;
                if (v < old || (v == old)) heap_up(h, pos, pos[j]);
                else heap_down(h, pos, hn, pos[j]);
            } else {
                h[hn].v = v;
                h[hn].i = j;
                heap_up(h, pos, hn);
                hn++;
            }
        }
    }
    if (bad) failed[key] = 1;
}

// The subsegments of each slot's colour with lds_lo < m <= cap_a points go to
// the slot's list 0, those with cap_a < m <= cap_b to its list 1 (counters
// big[0] and big[2]).
__global__ void k_fmm_classify(FmmLayout L, const uint8_t* __restrict__ colour, uint32_t n_labels, uint32_t c0,
                               uint32_t lds_lo, uint32_t cap_a, uint32_t cap_b) {
    const uint32_t key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key == 0 || key >= n_labels || colour[key] != c0 + blockIdx.y) return;
    uint8_t* sb = slot_base(L);
    const uint32_t* __restrict__ lst = (const uint32_t*)(sb + L.off_lcnt);
    const uint32_t m = lst[key + 1] - lst[key];
    if (m <= lds_lo || m > cap_b) return;
    uint32_t* big = (uint32_t*)(sb + L.off_big);
    const uint32_t cls = m > cap_a ? 1u : 0u;
    const uint32_t at = atomicAdd(&big[2 * cls], 1u);
    big[4 + cls * L.big_cap + at] = key;
}

// The march of a subsegment held whole in LDS by one 64-lane wave.  Its
// points are sorted by pixel index, so local index order is pixel order and
// the lane kernel's heap order (value, pixel index) is (value, local index).
// Per point: u (double), flags (bits 0-1 the lane kernel's state, bit 2 the
// channel side, bits 3-4 whether the points before and after in the row,
// local index -1 and +1, belong to the subsegment) and the local indices of
// the pixels one row before and after (0xFFFF: not in the subsegment).  The
// trial front is split by local index mod 64: lane L lists its own points
// that are in the front (at most ceil(cap / 64)) and keeps their minimum, so
// the next point to accept is a reduction over the 64 lane minima, and only
// the lanes that owned it, or whose minimum grew, scan their lists again
// (dropping accepted points), all at once.
constexpr uint32_t kFmmNone = 0xFFFFu;
constexpr uint8_t kFmmNeg = 4, kFmmYm = 8, kFmmYp = 16;

struct FmmLds {
    double* u;
    uint16_t* xn;    // [2 * cap]: row before, row after
    uint16_t* list;  // [64][per_lane]
    uint8_t* f;
    uint32_t per_lane;
};

__host__ __device__ constexpr uint32_t fmm_lds_per_lane(uint32_t cap) { return (cap + 63) / 64; }
__host__ __device__ constexpr size_t fmm_lds_bytes(uint32_t cap) {
    return ((size_t)cap * 13 + (size_t)64 * fmm_lds_per_lane(cap) * 2 + 15) & ~(size_t)15;
}

__device__ __forceinline__ bool fless(double va, uint32_t la, double vb, uint32_t lb) {
    return va < vb || (va == vb && la < lb);
}

// fmm_update2 on the LDS copy: the same values, by local index, read in three
// rounds of independent LDS loads (the point; its first neighbours; their
// neighbours beyond), then the same arithmetic.
__device__ double fmm_update2_lds(const FmmLds& S, uint32_t l, double dx, bool* bad) {
    const double idx2 = 1.0 / (dx * dx), aa = 9.0 / 4.0, third = 1.0 / 3.0;
    const uint8_t fl = S.f[l];
    const uint16_t xa = S.xn[2 * l], xb = S.xn[2 * l + 1];
    const bool nl = (fl & kFmmNeg) != 0;
    uint32_t n1[4], n2[4];
    n1[0] = xa;
    n1[1] = xb;
    n1[2] = (fl & kFmmYm) ? l - 1 : kFmmNone;
    n1[3] = (fl & kFmmYp) ? l + 1 : kFmmNone;
    uint8_t f1[4], f2[4];
    double u1[4], u2[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t a1 = n1[q] != kFmmNone ? n1[q] : l;
        f1[q] = S.f[a1];
        u1[q] = S.u[a1];
        if (q < 2) n2[q] = S.xn[2 * a1 + q];
    }
    if (n1[0] == kFmmNone) n2[0] = kFmmNone;
    if (n1[1] == kFmmNone) n2[1] = kFmmNone;
    n2[2] = (n1[2] != kFmmNone && (f1[2] & kFmmYm)) ? l - 2 : kFmmNone;
    n2[3] = (n1[3] != kFmmNone && (f1[3] & kFmmYp)) ? l + 2 : kFmmNone;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t a2 = n2[q] != kFmmNone ? n2[q] : l;
        f2[q] = S.f[a2];
        u2[q] = S.u[a2];
    }
    double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
    for (int axis = 0; axis < 2; axis++) {
        double v1 = INFINITY, v2 = INFINITY;
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int q = axis * 2 + side;
            if (n1[q] == kFmmNone || (f1[q] & 3) < 2 || !(u1[q] < v1)) continue;
            v1 = u1[q];
            if (n2[q] == kFmmNone || (f2[q] & 3) < 2 || !(u2[q] <= v1)) continue;
            const bool e1 = (f1[q] & kFmmNeg) != 0, e2 = (f2[q] & kFmmNeg) != 0;
            v2 = (e2 != e1 || e2 != nl) ? -u2[q] : u2[q];
        }
        if (v2 < INFINITY) {
            const double tp = third * (4.0 * v1 - v2);
            a += idx2 * aa;
            b -= idx2 * 2.0 * aa * tp;
            c += idx2 * aa * tp * tp;
        } else if (v1 < INFINITY) {
            a += idx2;
            b -= idx2 * 2.0 * v1;
            c += idx2 * v1 * v1;
        }
    }
    c -= 1.0;
    const double det = b * b - 4.0 * a * c;
    if (det < 0.0) {
        *bad = true;
        return INFINITY;
    }
    return (-b + sqrt(det)) / 2.0 / a;
}

// The block is one wave: its LDS accesses complete in order, so a fence at
// wavefront scope (which keeps the compiler from moving them) orders one
// lane's writes before other lanes' later reads.
__device__ __forceinline__ void fmm_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One step of a min-reduction of (value, local index): each lane takes the
// smaller of its own and the one CTRL (a DPP pattern) brings it.
template <int CTRL>
__device__ __forceinline__ void fmm_dpp_min(double& v, uint32_t& l) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)bits, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(bits >> 32), CTRL, 0xF, 0xF, false);
    const uint32_t ol = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)l, CTRL, 0xF, 0xF, false);
    const double ov = __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    if (fless(ov, ol, v, l)) {
        v = ov;
        l = ol;
    }
}

__device__ __forceinline__ double fmm_readlane(double v, int lane) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)bits, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The minimum (value, local index) over the wave, in every lane (all 64 lanes
// active); empty lanes hold (inf, ~0).  Within each row of 16 lanes by DPP
// (quad swaps, then rotations by 4 and 8), then across the four rows.
__device__ __forceinline__ void fmm_wave_min(double& bv, uint32_t& bl) {
    fmm_dpp_min<0xB1>(bv, bl);      // quad_perm [1, 0, 3, 2]
    fmm_dpp_min<0x4E>(bv, bl);      // quad_perm [2, 3, 0, 1]
    fmm_dpp_min<0x124>(bv, bl);     // row_ror:4
    fmm_dpp_min<0x128>(bv, bl);     // row_ror:8
    double v = fmm_readlane(bv, 0);
    uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)bl, 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16) {
        const double ov = fmm_readlane(bv, r);
        const uint32_t ol = (uint32_t)__builtin_amdgcn_readlane((int)bl, r);
        if (fless(ov, ol, v, l)) {
            v = ov;
            l = ol;
        }
    }
    bv = v;
    bl = l;
}

// Lane's scan of its front list (count in LDS): drops accepted points and
// returns the minimum (value, local index).  Eight entries at a time: the
// group's list reads, then their states and values, are issued together,
// and the survivors are written back in place (every write goes to a
// position the group has already read).
__device__ void fmm_lane_scan(const FmmLds& S, uint32_t lane, uint32_t* cnt_lds, double& mv, uint32_t& ml) {
    uint16_t* lst = S.list + lane * S.per_lane;
    const uint32_t cnt = cnt_lds[lane];
    mv = INFINITY;
    ml = 0xFFFFFFFFu;
    uint32_t w = 0;
    for (uint32_t j0 = 0; j0 < cnt; j0 += 8) {
        const uint32_t left = cnt - j0;
        uint16_t e[8];
        uint8_t f[8];
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) e[k] = lst[j0 + ((uint32_t)k < left ? (uint32_t)k : 0u)];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            f[k] = S.f[e[k]];
            v[k] = S.u[e[k]];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if ((uint32_t)k >= left || (f[k] & 3) == 3) continue;
            lst[w++] = e[k];
            if (fless(v[k], e[k], mv, ml)) {
                mv = v[k];
                ml = e[k];
            }
        }
    }
    cnt_lds[lane] = w;
}

// The march of k_fmm_march for the subsegments of list cls of the k slots,
// one 64-lane block each (blocks take the next entry of the slots' lists,
// counted by next[cls], until they are done), with the subsegment held in LDS
// (fmm_lds_bytes(cap) bytes, at most cap points).  Set-up: the point list is
// sorted (bitonic, in LDS) and written back sorted, pos maps each pixel to
// its local index, and the neighbour indices, flags and u are gathered.  Per
// accepted point, lanes 0-3 compute the four neighbour updates at once (an
// update reads only frozen values, and the updates of one point freeze
// nothing) and write those before the first negative discriminant, as the
// lane kernel stops there; a point new to the front joins its owner's list
// (lengths in LDS), and the owner lanes fold the values into their minima in
// neighbour order.  Global memory is only read until the march ends; then u
// is written back.
__global__ __launch_bounds__(64) void k_fmm_march_lds(uint32_t nx, uint32_t ny, double dx, FmmLayout L, uint32_t k_slots,
                                                      uint32_t* __restrict__ next, uint8_t* __restrict__ failed,
                                                      uint32_t cls, uint32_t cap) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_fmm[];
    __shared__ uint32_t cur;
    __shared__ uint32_t s_cnt[64];   // each lane's front list length
    FmmLds S;
    S.u = (double*)lds_fmm;
    S.xn = (uint16_t*)(S.u + cap);
    S.per_lane = fmm_lds_per_lane(cap);
    S.list = S.xn + 2 * cap;
    S.f = (uint8_t*)(S.list + 64 * S.per_lane);
    uint32_t* scratch = (uint32_t*)lds_fmm;          // 2 * cap entries, before u is filled
    const uint32_t lane = threadIdx.x;
    for (;;) {
        if (lane == 0) cur = atomicAdd(&next[cls], 1u);
        __syncthreads();
        uint32_t e = cur;
        __syncthreads();
        // the entry's slot: the slots' lists follow one another
        uint32_t g = 0;
        for (; g < k_slots; g++) {
            const uint32_t c = ((const uint32_t*)(L.base + g * L.stride + L.off_big))[2 * cls];
            if (e < c) break;
            e -= c;
        }
        if (g == k_slots) break;
        uint8_t* sb = L.base + g * L.stride;
        const uint32_t* __restrict__ dom = (const uint32_t*)sb;
        const uint8_t* __restrict__ neg = sb + L.off_neg;
        double* __restrict__ u = (double*)(sb + L.off_u);
        const uint8_t* __restrict__ state = sb + L.off_fixed;
        const uint32_t* __restrict__ lst = (const uint32_t*)(sb + L.off_lcnt);
        int32_t* __restrict__ pos = (int32_t*)(sb + L.off_pos);
        const uint32_t key = ((const uint32_t*)(sb + L.off_big))[4 + cls * L.big_cap + e];
        uint32_t* __restrict__ pts = (uint32_t*)(sb + L.off_pts) + lst[key];
        const uint32_t m = lst[key + 1] - lst[key];
        // sort the points by pixel index
        uint32_t p2 = 1;
        while (p2 < m) p2 <<= 1;
        for (uint32_t k = lane; k < p2; k += 64) scratch[k] = k < m ? pts[k] : 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t kk = 2; kk <= p2; kk <<= 1) {
            for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
                for (uint32_t i = lane; i < p2; i += 64) {
                    const uint32_t ixj = i ^ jj;
                    if (ixj > i) {
                        const uint32_t x0 = scratch[i], x1 = scratch[ixj];
                        if (((i & kk) == 0) == (x0 > x1)) {
                            scratch[i] = x1;
                            scratch[ixj] = x0;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t k = lane; k < m; k += 64) {
            const uint32_t p = scratch[k];
            pts[k] = p;
            pos[p] = (int32_t)k;
        }
        __syncthreads();
        // neighbours, flags, u; each lane lists its own frozen points
        uint32_t cnt = 0;
        for (uint32_t k = lane; k < m; k += 64) {
            const uint32_t p = pts[k];
            const uint32_t x = p / ny, y = p % ny;
            uint16_t up = kFmmNone, dn = kFmmNone;
            if (x > 0 && dom[p - ny] == key) up = (uint16_t)pos[p - ny];
            if (x + 1 < nx && dom[p + ny] == key) dn = (uint16_t)pos[p + ny];
            uint8_t fl = neg[p] ? kFmmNeg : 0;
            if (y > 0 && dom[p - 1] == key) fl |= kFmmYm;
            if (y + 1 < ny && dom[p + 1] == key) fl |= kFmmYp;
            if (state[p]) {
                fl |= 2;
                S.list[lane * S.per_lane + cnt++] = (uint16_t)k;
            }
            S.u[k] = u[p];
            S.xn[2 * k] = up;
            S.xn[2 * k + 1] = dn;
            S.f[k] = fl;
        }
        s_cnt[lane] = cnt;
        __syncthreads();
        double mv;
        uint32_t ml;
        fmm_lane_scan(S, lane, s_cnt, mv, ml);
        fmm_wave_sync();
        bool bad = false;
        for (;;) {
            // the next point: the smallest (value, local index) of the lane minima
            double bv = mv;
            uint32_t bl = ml;
            fmm_wave_min(bv, bl);
            if (bad || bl == 0xFFFFFFFFu) break;
            const uint32_t i = bl;
            if (lane == 0) S.f[i] = (uint8_t)(S.f[i] | 3u);
            bool rescan = (i & 63u) == lane;         // its owner scans again (dropping it)
            fmm_wave_sync();
            // lanes 0-3: the update of neighbour k (x before, x after, y before,
            // y after) of a point not frozen
            int32_t flag = 0;
            double v = 0.0;
            uint32_t l = kFmmNone;
            uint8_t fl = 0;
            if (lane < 4) {
                const uint8_t fi = S.f[i];
                if (lane < 2) l = S.xn[2 * i + lane];
                else if (lane == 2 && (fi & kFmmYm)) l = i - 1;
                else if (lane == 3 && (fi & kFmmYp)) l = i + 1;
                if (l != kFmmNone) {
                    fl = S.f[l];
                    if ((fl & 3) < 2) {
                        bool bd = false;
                        v = fmm_update2_lds(S, l, dx, &bd);
                        flag = bd ? 2 : (v == 0.0 ? 0 : 1);     // scikit-fmm skips a zero update
                    }
                }
            }
            // the lane kernel applies them in neighbour order and stops at the
            // first negative discriminant; the four points are distinct and an
            // update reads only frozen values, so lane k writes update k itself
            // when no earlier one failed, and appends a point new to the front
            // to its owner's list
            const uint64_t negs = __ballot(flag == 2) & 0xFull;
            const uint32_t kstop = negs ? (uint32_t)__ffsll((unsigned long long)negs) - 1u : 4u;
            if (negs) bad = true;
            const bool apply = flag == 1 && lane < kstop;
            if (apply) {
                S.u[l] = v;
                S.f[l] = (uint8_t)((fl & ~3u) | 1u);
                if ((fl & 3) != 1) {
                    const uint32_t o = l & 63u;
                    S.list[o * S.per_lane + atomicAdd(&s_cnt[o], 1u)] = (uint16_t)l;
                }
            }
            // each owner lane folds its applied values into its minimum, in
            // neighbour order; when a lane's minimum is q its value is q's
            // value before the update, so a larger one means the minimum grew
            const uint32_t la = apply ? l : kFmmNone;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)la, k);
                const double vq = fmm_readlane(v, k);
                if (q == kFmmNone || (q & 63u) != lane || rescan) continue;
                if (fless(vq, q, mv, ml)) {
                    mv = vq;
                    ml = q;
                } else if (ml == q && vq > mv) {
                    rescan = true;
                }
            }
            fmm_wave_sync();
            // the lanes that rescan do so together, each its own list
            if (rescan) fmm_lane_scan(S, lane, s_cnt, mv, ml);
            fmm_wave_sync();
        }
        for (uint32_t k = lane; k < m; k += 64) u[pts[k]] = S.u[k];
        if (lane == 0 && bad) failed[key] = 1;
        __syncthreads();
    }
}

// Per problem: fixed-point sum of the signed distances of the points a
// channel reaches, their count, the channel pixels and the own pixels.  The
// lanes of a wave that share a key add once.
__global__ void k_fmm_reduce(FmmLayout L, const int32_t* __restrict__ label, uint32_t nx, uint32_t ny,
                             u64* __restrict__ sums, uint32_t* __restrict__ cnt, uint32_t* __restrict__ nch,
                             uint32_t* __restrict__ nphi) {
    const uint32_t* __restrict__ dom = (const uint32_t*)slot_base(L);
    const uint8_t* __restrict__ neg = slot_base(L) + L.off_neg;
    const double* __restrict__ u = (const double*)(slot_base(L) + L.off_u);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = 0;
    long long q = 0;
    uint32_t c = 0, ch = 0, own = 0;
    if (i < nx * ny) {
        key = dom[i];
        if (key) {
            own = abs_label(label[i]) == key ? 1u : 0u;
            ch = neg[i];
            const double v = u[i];
            if (!isinf(v)) {
                q = __double2ll_rn((neg[i] ? -v : v) * kFix);
                c = 1;
            }
        }
    }
    // per distinct non-zero key of the wave, the sums of its lanes (the
    // others contribute zero) reach every lane, and the lowest lane of the
    // key adds them: integer sums, so the grouping changes nothing
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t todo = __ballot(key != 0u);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
        const bool in = key == lk;
        long long gq = in ? q : 0;
        uint32_t gc = in ? c : 0u, gch = in ? ch : 0u, gown = in ? own : 0u;
        for (int o = 32; o > 0; o >>= 1) {
            gq += __shfl_xor(gq, o);
            gc += __shfl_xor(gc, o);
            gch += __shfl_xor(gch, o);
            gown += __shfl_xor(gown, o);
        }
        if ((int)lane == leader) {
            if (gq) atomicAdd(&sums[lk], (u64)gq);
            if (gc) atomicAdd(&cnt[lk], gc);
            if (gch) atomicAdd(&nch[lk], gch);
            if (gown) atomicAdd(&nphi[lk], gown);
        }
        todo &= ~__ballot(in);
    }
}

__global__ void k_hsl_remap(const int32_t* __restrict__ label, uint32_t n, const int32_t* __restrict__ hsl,
                            const uint8_t* __restrict__ sel, uint32_t n_labels, float* __restrict__ grid) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t l = label[i];
    const uint32_t a = abs_label(l);
    float v = 0.0f;
    if (a && a < n_labels && (sel[a] & (l < 0 ? 2u : 1u))) v = (float)hsl[a];
    grid[i] = v;
}

// ----------------------------------------------------------------- map_hsl
// Mapping.map_hsl (mapping.py:987-1020): uint16 rescaling, grey_dilation
// (scipy: flat dfw x dfw window, offsets -(s-1)/2 .. s/2 along each axis,
// 'reflect' borders) into the masked pixels only, mean over disk(mdr)
// (skimage rank.mean: float64 sum / in-image count, truncated), rescale.
// Float32 arithmetic in numpy's order, no contraction.

// order-preserving uint key of a float (min/max through integer atomics)
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fval(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_minmax(const float* __restrict__ a, size_t n, uint32_t* __restrict__ mm) {
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t k = fkey(a[i]);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_down(lo, o));
        hi = max(hi, (uint32_t)__shfl_down(hi, o));
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void k_hsl_clip(const float* __restrict__ hsl, size_t n, const uint32_t* __restrict__ mm,
                           uint16_t* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mn = fval(mm[0]), mx = fval(mm[1]);
    const float t = (65535.0f * (hsl[i] - mn)) / (mx - mn);
    out[i] = (t >= 0.0f && t < 65536.0f) ? (uint16_t)(uint32_t)t : (uint16_t)0;
}

__device__ __forceinline__ int reflect_idx(int j, int n) {
    while (j < 0 || j >= n) j = (j < 0) ? (-j - 1) : (2 * n - j - 1);
    return j;
}

template <int AXIS>
__global__ void k_max_axis(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, uint32_t nx, uint32_t ny,
                           int lo, int hi) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)nx * ny) return;
    const int x = (int)(i / ny), y = (int)(i % ny);
    uint16_t m = 0;
    for (int k = -lo; k <= hi; k++) {
        const uint16_t v = AXIS == 0 ? in[(size_t)reflect_idx(x + k, (int)nx) * ny + y]
                                     : in[(size_t)x * ny + reflect_idx(y + k, (int)ny)];
        m = v > m ? v : m;
    }
    out[i] = m;
}

// hsl_dilated[~mask] = hsl_clipped[~mask]
__global__ void k_keep_unmasked(uint16_t* __restrict__ dil, const uint16_t* __restrict__ clip,
                                const uint8_t* __restrict__ mask, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n && !mask[i]) dil[i] = clip[i];
}

// Prefix sums along y, stored transposed so that neighbouring threads (x)
// write neighbouring addresses: P[y][x] = sum of in[x][0..y-1], y in 0..ny.
template <typename T, typename S>
__global__ void k_row_prefix(const T* __restrict__ in, uint32_t nx, uint32_t ny, S* __restrict__ P) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nx) return;
    S acc = 0;
    const T* row = in + (size_t)x * ny;
    P[x] = 0;
    for (uint32_t y = 0; y < ny; y++) {
        acc += (S)row[y];
        P[(size_t)(y + 1) * nx + x] = acc;
    }
}

// half-width of disk(r) at row offset dx: the largest w with w^2 <= r^2 - dx^2
__device__ __forceinline__ int disk_w(int r, int dx) {
    const int t = r * r - dx * dx;
    int w = (int)sqrtf((float)t);
    while (w * w > t) w--;
    while ((w + 1) * (w + 1) <= t) w++;
    return w;
}

__global__ void k_disk_mean(const u64* __restrict__ P, uint32_t nx, uint32_t ny, int r, uint16_t* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)nx * ny) return;
    const int x = (int)(i / ny), y = (int)(i % ny);
    u64 sum = 0, pop = 0;
    for (int dx = -r; dx <= r; dx++) {
        const int xx = x + dx;
        if (xx < 0 || xx >= (int)nx) continue;
        const int w = disk_w(r, dx);
        const int y0 = max(0, y - w), y1 = min((int)ny - 1, y + w);
        sum += P[(size_t)(y1 + 1) * nx + xx] - P[(size_t)y0 * nx + xx];
        pop += (u64)(y1 - y0 + 1);
    }
    out[i] = pop ? (uint16_t)(uint32_t)((double)sum / (double)pop) : (uint16_t)0;
}

__global__ void k_hsl_unclip(const uint16_t* __restrict__ m, size_t n, const uint32_t* __restrict__ mm,
                             float* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mn = fval(mm[0]), mx = fval(mm[1]);
    out[i] = (((float)m[i]) / 65535.0f) * (mx - mn) + mn;
}

// -------------------------------------------------------------- map_aspect
// Mapping.map_aspect (mapping.py:1022-1053): channel slopes zeroed, slope
// quantised to uint8 by sf = max/255 (float32) and median-filtered over
// disk(r_slope); u and v quantised by 2/255 and median-filtered over
// disk(r_uv); aspect = arctan2(v, u) (float64, rounded to float32); mask =
// slope < threshold, or channel.  Median (skimage rank.median): the smallest
// value v with 2 * #(values <= v) >= in-image count (k_disk_median).

__global__ void k_slope_zero_channels(const float* __restrict__ slope, const uint32_t* __restrict__ mapping,
                                      uint32_t is_channel, size_t n, float* __restrict__ out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = ((mapping[i] & is_channel) == is_channel) ? 0.0f : slope[i];
}

__global__ void k_quant_slope(const float* __restrict__ s, size_t n, const uint32_t* __restrict__ mm,
                              uint8_t* __restrict__ q) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sf = fval(mm[1]) / 255.0f;
    const float t = s[i] / sf;
    q[i] = (t >= 0.0f && t < 256.0f) ? (uint8_t)(uint32_t)t : (uint8_t)0;
}

__global__ void k_quant_uv(const float2* __restrict__ uv, size_t n, uint8_t* __restrict__ qu,
                           uint8_t* __restrict__ qv) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float sf = (float)(2.0 / 255.0);
    const float tu = (uv[i].x + 1.0f) / sf, tv = (uv[i].y + 1.0f) / sf;
    qu[i] = (tu >= 0.0f && tu < 256.0f) ? (uint8_t)(uint32_t)tu : (uint8_t)0;
    qv[i] = (tv >= 0.0f && tv < 256.0f) ? (uint8_t)(uint32_t)tv : (uint8_t)0;
}

// Median over disk(r) by a sliding histogram (Huang): a thread owns one
// column y and a run of kRun pixels along x; it builds the histogram of the
// footprint at the first pixel, then slides it one pixel along x (for every
// row offset dy: out goes (x - w, y + dy), in comes (x + 1 + w, y + dy)),
// reading the median from 16 coarse and 256 fine counts.  Neighbouring
// threads hold neighbouring y, so each footprint load is one row segment.
// Histograms live in LDS, bin-major (bin b of thread t at b * 64 + t).
constexpr int kRun = 64;
constexpr int kMedThreads = 64;

__global__ __launch_bounds__(kMedThreads) void k_disk_median(const uint8_t* __restrict__ img, uint32_t nx,
                                                             uint32_t ny, int r, uint8_t* __restrict__ out) {
    __shared__ uint16_t fine[256 * kMedThreads];
    __shared__ uint16_t coarse[16 * kMedThreads];
    const uint32_t t = threadIdx.x;
    const uint32_t y = blockIdx.x * kMedThreads + t;
    const uint32_t x0 = blockIdx.y * kRun;
    for (int b = 0; b < 256; b++) fine[b * kMedThreads + t] = 0;
    for (int c = 0; c < 16; c++) coarse[c * kMedThreads + t] = 0;
    if (y >= ny || x0 >= nx) return;                     // this thread's histogram is never read by others
    auto add = [&](uint8_t v, int d) {
        fine[v * kMedThreads + t] = (uint16_t)(fine[v * kMedThreads + t] + d);
        coarse[(v >> 4) * kMedThreads + t] = (uint16_t)(coarse[(v >> 4) * kMedThreads + t] + d);
    };
    uint32_t pop = 0;
    for (int dy = -r; dy <= r; dy++) {
        const int yy = (int)y + dy;
        if (yy < 0 || yy >= (int)ny) continue;
        const int w = disk_w(r, dy);
        for (int xx = max(0, (int)x0 - w); xx <= min((int)nx - 1, (int)x0 + w); xx++) {
            add(img[(size_t)xx * ny + yy], 1);
            pop++;
        }
    }
    const uint32_t x1 = min(nx, x0 + (uint32_t)kRun);
    for (uint32_t x = x0;; x++) {
        // the smallest v with 2 * #(values <= v) >= pop
        uint32_t acc = 0;
        int c = 0;
        while (c < 15 && 2u * (acc + coarse[c * kMedThreads + t]) < pop) {
            acc += coarse[c * kMedThreads + t];
            c++;
        }
        int v = c * 16;
        while (v < 255 && 2u * (acc + fine[v * kMedThreads + t]) < pop) {
            acc += fine[v * kMedThreads + t];
            v++;
        }
        out[(size_t)x * ny + y] = (uint8_t)v;
        if (x + 1 >= x1) break;
        for (int dy = -r; dy <= r; dy++) {
            const int yy = (int)y + dy;
            if (yy < 0 || yy >= (int)ny) continue;
            const int w = disk_w(r, dy);
            const int xo = (int)x - w, xi = (int)x + 1 + w;
            if (xo >= 0) {
                add(img[(size_t)xo * ny + yy], -1);
                pop--;
            }
            if (xi < (int)nx) {
                add(img[(size_t)xi * ny + yy], 1);
                pop++;
            }
        }
    }
}

__global__ void k_aspect(const uint8_t* __restrict__ mu, const uint8_t* __restrict__ mv,
                         const uint8_t* __restrict__ ms, const uint32_t* __restrict__ mm, int do_median,
                         const float* __restrict__ slope0, const uint32_t* __restrict__ mapping, uint32_t is_channel,
                         float slope_threshold, size_t n, float* __restrict__ aspect, uint8_t* __restrict__ mask) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double sf = 2.0 / 255.0;
    const double ux = sf * (double)mu[i] - 1.0, vy = sf * (double)mv[i] - 1.0;
    aspect[i] = (float)atan2(vy, ux);
    float sl;
    if (do_median) sl = (fval(mm[1]) / 255.0f) * (float)ms[i];
    else sl = slope0[i];
    const bool chan = (mapping[i] & is_channel) == is_channel;
    mask[i] = (sl < slope_threshold || chan) ? 1 : 0;
}

__global__ void k_fill_mm(uint32_t* __restrict__ mm) {
    if (threadIdx.x == 0) {
        mm[0] = 0xffffffffu;
        mm[1] = 0u;
    }
}

__global__ void k_fill_bytes(uint8_t* __restrict__ p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_zero_bytes(uint8_t* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

inline dim3 grid_for(size_t n) { return dim3((uint32_t)((n + kBlock - 1) / kBlock)); }

inline size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

void zero(void* p, size_t bytes, hipStream_t s) {
    const size_t blocks = std::min<size_t>((bytes + kBlock - 1) / kBlock, 65535);
    hipLaunchKernelGGL(k_zero_bytes, dim3((uint32_t)std::max<size_t>(blocks, 1)), dim3(kBlock), 0, s, (uint8_t*)p,
                       bytes);
}

}  // namespace

extern "C" int slm_fmm_pairs(const int32_t* label, uint32_t nx, uint32_t ny, const uint8_t* colour,
                             uint32_t n_labels, uint64_t* pairs, uint32_t cap, uint32_t* count, void* stream,
                             uint32_t flags) {
    if (!label || !colour || !count || (!pairs && cap) || !nx || !ny || (uint64_t)nx * ny >= 0x80000000ull)
        return slm_fail(SLM_E_INVALID_ARG, "bad fmm_pairs arguments");
    hipStream_t s = (hipStream_t)stream;
    zero(count, sizeof(uint32_t), s);
    hipLaunchKernelGGL(k_fmm_pairs, grid_for((size_t)nx * ny), dim3(kBlock), 0, s, label, nx, ny, colour, n_labels,
                       (u64*)pairs, cap, count);
    return slm_finish(stream, flags);
}

namespace {
// subsegments marched with LDS heaps have more than kLdsLo points; those of
// at most kCapA points are marched from 32 KiB of LDS, larger ones from the
// device's LDS per block
constexpr uint32_t kLdsLo = 512;
constexpr uint32_t kCapA = 2048;

// capacity of each LDS-march list of a slot: its subsegments are disjoint
size_t fmm_big_cap(uint32_t nx, uint32_t ny) { return (size_t)nx * ny / kLdsLo + 2; }

// bytes of one colour slot, and the slot layout
size_t fmm_slot_bytes(uint32_t nx, uint32_t ny) {
    const size_t n = (size_t)nx * ny;
    const size_t n_tiles = (size_t)((nx + kT - 1) / kT) * ((ny + kT - 1) / kT);
    // dom, neg, u, state/fixed, two tile-activity maps; then, for the
    // second-order march, label counts and cursors (labels < n + 1), point
    // lists, heap positions and heaps
    return al256(n * 4) + al256(n) + al256(n * 8) + al256(n) + 2 * al256(n_tiles) + 2 * al256((n + 2) * 4) +
           al256(n * 4) + al256(n * 4) + al256(n * sizeof(HeapEnt)) + al256(16 + 8 * fmm_big_cap(nx, ny));
}

// A second stream per device (default priority) with fork and join events.
struct SideStream {
    hipStream_t s;
    hipEvent_t fork, join;
};

int fmm_side_stream(SideStream** out) {
    static std::mutex mu;
    static SideStream side[64];
    static bool made[64];
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return slm_fail(SLM_E_INVALID_ARG, "device %d", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!made[dev]) {
        HIP_TRY(hipStreamCreateWithFlags(&side[dev].s, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&side[dev].fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&side[dev].join, hipEventDisableTiming));
        made[dev] = true;
    }
    *out = &side[dev];
    return SLM_OK;
}

// Largest subsegment marched in LDS: fmm_lds_bytes(cap) within the device's
// LDS per block (less 512 bytes for the kernel's static LDS), local indices
// below 0xFFFF.
uint32_t fmm_lds_cap() {
    static uint32_t cap[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if (!cap[dev]) {
        int bytes = 0;
        if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) bytes = 65536;
        uint32_t c = (uint32_t)std::min<long>(65534, std::max<long>(0, ((long)bytes - 512) / 13));
        while (c && fmm_lds_bytes(c) + 512 > (size_t)bytes) c--;
        cap[dev] = c;
    }
    return cap[dev];
}
}  // namespace

extern "C" size_t slm_fmm_workspace_size_k(uint32_t nx, uint32_t ny, uint32_t k) {
    if (!nx || !ny || !k) return 0;
    return 512 + (size_t)k * fmm_slot_bytes(nx, ny);
}

extern "C" size_t slm_fmm_workspace_size(uint32_t nx, uint32_t ny) { return slm_fmm_workspace_size_k(nx, ny, 1); }

extern "C" int slm_fmm_hsl(const int32_t* label, const uint32_t* mapping, uint32_t thin_flag, const uint8_t* colour,
                           uint32_t n_labels, uint32_t n_colours, uint32_t nx, uint32_t ny, double dx, uint32_t order,
                           uint64_t* sums, uint32_t* cnt, uint32_t* nch, uint32_t* nphi, uint8_t* failed,
                           void* workspace, size_t workspace_bytes, void* stream, uint32_t flags) {
    if (!label || !mapping || !colour || !sums || !cnt || !nch || !nphi || !workspace || !nx || !ny ||
        (uint64_t)nx * ny >= 0x80000000ull || !(dx > 0.0) || n_colours > 255 || (order != 1 && order != 2) ||
        (order == 2 && !failed) || n_labels > (uint64_t)nx * ny + 1)
        return slm_fail(SLM_E_INVALID_ARG, "bad fmm_hsl arguments");
    if (workspace_bytes < slm_fmm_workspace_size(nx, ny))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes, need %zu", workspace_bytes,
                        slm_fmm_workspace_size(nx, ny));
    const size_t n = (size_t)nx * ny;
    const size_t slot = fmm_slot_bytes(nx, ny);
    // colours relaxed together: as many slots as the workspace holds
    const uint32_t group = (uint32_t)std::min<size_t>(std::max<uint32_t>(n_colours, 1u), (workspace_bytes - 512) / slot);
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    uint32_t* ctr = (uint32_t*)w;
    const uint32_t n_tiles = ((nx + kT - 1) / kT) * ((ny + kT - 1) / kT);
    FmmLayout L;
    L.base = w + 256;
    L.stride = slot;
    L.off_neg = al256(n * 4);
    L.off_u = L.off_neg + al256(n);
    L.off_fixed = L.off_u + al256(n * 8);
    L.off_act0 = L.off_fixed + al256(n);
    L.off_act1 = L.off_act0 + al256(n_tiles);
    L.off_lcnt = L.off_act1 + al256(n_tiles);
    L.off_lcur = L.off_lcnt + al256((n + 2) * 4);
    L.off_pts = L.off_lcur + al256((n + 2) * 4);
    L.off_pos = L.off_pts + al256(n * 4);
    L.off_heap = L.off_pos + al256(n * 4);
    L.off_big = L.off_heap + al256(n * sizeof(HeapEnt));
    L.big_cap = fmm_big_cap(nx, ny);
    hipStream_t s = (hipStream_t)stream;
    zero(sums, (size_t)n_labels * 8, s);
    zero(cnt, (size_t)n_labels * 4, s);
    zero(nch, (size_t)n_labels * 4, s);
    zero(nphi, (size_t)n_labels * 4, s);
    if (failed) zero(failed, (size_t)n_labels, s);
    // a bound on the launches keeps a bad input from looping
    const uint32_t max_launches = 2u * (nx + ny) + 64u;
    for (uint32_t c0 = 0; c0 < n_colours; c0 += group) {
        const uint32_t k = std::min(group, n_colours - c0);
        const dim3 g((uint32_t)((n + kBlock - 1) / kBlock), k);
        hipLaunchKernelGGL(k_fmm_setup, g, dim3(kBlock), 0, s, label, mapping, thin_flag, colour, n_labels, c0, nx,
                           ny, L);
        hipLaunchKernelGGL(k_fmm_init, g, dim3(kBlock), 0, s, nx, ny, dx, L);
        if (order == 2) {
            for (uint32_t j = 0; j < k; j++) zero(L.base + j * slot + L.off_lcnt, (size_t)(n_labels + 1) * 4, s);
            hipLaunchKernelGGL(k_fmm_lcount, g, dim3(kBlock), 0, s, nx, ny, L);
            hipLaunchKernelGGL(k_fmm_lscan, dim3(1, k), dim3(1024), 0, s, n_labels + 1, L);
            hipLaunchKernelGGL(k_fmm_lscatter, g, dim3(kBlock), 0, s, nx, ny, L);
            // subsegments of more than kLdsLo points march in LDS (test hook: all one lane each)
            const uint32_t cap_b = (flags & SLM_FLAG_FMM_LANES) ? 0u : fmm_lds_cap();
            const uint32_t lo = cap_b > kLdsLo ? kLdsLo : 0u, hi = cap_b > kLdsLo ? cap_b : 0u;
            if (!hi)
                hipLaunchKernelGGL(k_fmm_march, dim3((n_labels + 63) / 64, k), dim3(64), 0, s, nx, ny, dx, L, colour,
                                   n_labels, c0, failed, lo, hi);
            if (hi) {
                uint32_t* next = ctr + 1;              // two queue cursors beside the relaxation counter
                zero(next, 2 * sizeof(uint32_t), s);
                for (uint32_t j = 0; j < k; j++) zero(L.base + j * slot + L.off_big, 16, s);
                const uint32_t cap_a = std::min(kCapA, hi);
                hipLaunchKernelGGL(k_fmm_classify, dim3((n_labels + 255) / 256, k), dim3(256), 0, s, L, colour,
                                   n_labels, c0, lo, cap_a, hi);
                const size_t lds_a = fmm_lds_bytes(cap_a), lds_b = fmm_lds_bytes(hi);
                static bool attr_set = false;
                if (!attr_set) {
                    HIP_TRY(hipFuncSetAttribute((const void*)k_fmm_march_lds,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
                    attr_set = true;
                }
                // The largest subsegments set the pace (one wave each, one per CU for
                // their LDS): they start first on the caller's stream, while the small
                // LDS marches and the lane kernel run beside them on a second stream.
                SideStream* side = nullptr;
                int rc = fmm_side_stream(&side);
                if (rc) return rc;
                HIP_TRY(hipEventRecord(side->fork, s));
                HIP_TRY(hipStreamWaitEvent(side->s, side->fork, 0));
                if (hi > cap_a)
                    hipLaunchKernelGGL(k_fmm_march_lds, dim3(256), dim3(64), lds_b, s, nx, ny, dx, L, k, next,
                                       failed, 1u, hi);
                // persistent blocks: several per CU for the small subsegments
                hipLaunchKernelGGL(k_fmm_march_lds, dim3(1280), dim3(64), lds_a, side->s, nx, ny, dx, L, k, next,
                                   failed, 0u, cap_a);
                hipLaunchKernelGGL(k_fmm_march, dim3((n_labels + 63) / 64, k), dim3(64), 0, side->s, nx, ny, dx, L,
                                   colour, n_labels, c0, failed, lo, hi);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipEventRecord(side->join, side->s));
                HIP_TRY(hipStreamWaitEvent(s, side->join, 0));
            }
            hipLaunchKernelGGL(k_fmm_reduce, g, dim3(kBlock), 0, s, L, label, nx, ny, (u64*)sums, cnt, nch, nphi);
            continue;
        }
        for (uint32_t j = 0; j < k; j++)
            hipLaunchKernelGGL(k_fill_bytes, grid_for(n_tiles), dim3(kBlock), 0, s, L.base + j * slot + L.off_act0,
                               (size_t)n_tiles, (uint8_t)1);
        for (uint32_t launches = 0;; launches++) {
            if (launches > max_launches) return slm_fail(SLM_E_HIP, "eikonal relaxation did not converge");
            zero(ctr, sizeof(uint32_t), s);
            hipLaunchKernelGGL(k_fmm_relax_tiles, dim3(n_tiles, k), dim3(kBlock), 0, s, nx, ny, dx, L,
                               launches & 1u, ctr);
            uint32_t dec = 0;
            HIP_TRY(hipMemcpyAsync(&dec, ctr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (!dec) break;
        }
        hipLaunchKernelGGL(k_fmm_reduce, g, dim3(kBlock), 0, s, L, label, nx, ny, (u64*)sums, cnt, nch, nphi);
    }
    return slm_finish(stream, flags);
}

extern "C" int slm_hsl_remap(const int32_t* label, uint32_t n, const int32_t* hsl, const uint8_t* selected,
                             uint32_t n_labels, float* grid, void* stream, uint32_t flags) {
    if (!label || !hsl || !selected || !grid) return slm_fail(SLM_E_INVALID_ARG, "bad hsl_remap arguments");
    if (!n) return slm_finish(stream, flags);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_hsl_remap, grid_for(n), dim3(kBlock), 0, s, label, n, hsl, selected, n_labels, grid);
    return slm_finish(stream, flags);
}

extern "C" size_t slm_map_hsl_workspace_size(uint32_t nx, uint32_t ny) {
    const size_t n = (size_t)nx * ny;
    return 3 * al256(n * 2) + al256((size_t)nx * (ny + 1) * 8) + 256 + 256;
}

extern "C" int slm_map_hsl(const float* hsl, const uint8_t* mask, uint32_t nx, uint32_t ny, int dfw, int mdr,
                           float* out, void* workspace, size_t workspace_bytes, void* stream, uint32_t flags) {
    if (!hsl || !mask || !out || !workspace || !nx || !ny || dfw < 0 || mdr < 0 ||
        (uint64_t)nx * ny >= 0x80000000ull || dfw > 2 * (int)std::min(nx, ny))
        return slm_fail(SLM_E_INVALID_ARG, "bad map_hsl arguments");
    if (workspace_bytes < slm_map_hsl_workspace_size(nx, ny))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes, need %zu", workspace_bytes,
                        slm_map_hsl_workspace_size(nx, ny));
    const size_t n = (size_t)nx * ny;
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    uint16_t* clip = (uint16_t*)w;
    uint16_t* tmp = (uint16_t*)(w + al256(n * 2));
    uint16_t* dil = (uint16_t*)(w + 2 * al256(n * 2));
    u64* P = (u64*)(w + 3 * al256(n * 2));
    uint32_t* mm = (uint32_t*)((uint8_t*)P + al256((size_t)nx * (ny + 1) * 8));
    hipStream_t s = (hipStream_t)stream;
    const dim3 g = grid_for(n);
    hipLaunchKernelGGL(k_fill_mm, dim3(1), dim3(64), 0, s, mm);
    hipLaunchKernelGGL(k_minmax, dim3((uint32_t)std::min<size_t>((n + kBlock - 1) / kBlock, 1024)), dim3(kBlock), 0, s,
                       hsl, n, mm);
    hipLaunchKernelGGL(k_hsl_clip, g, dim3(kBlock), 0, s, hsl, n, (const uint32_t*)mm, clip);
    const uint16_t* src = clip;
    if (dfw > 0) {
        const int lo = (dfw - 1) / 2, hi = dfw / 2;
        hipLaunchKernelGGL(k_max_axis<0>, g, dim3(kBlock), 0, s, (const uint16_t*)clip, tmp, nx, ny, lo, hi);
        hipLaunchKernelGGL(k_max_axis<1>, g, dim3(kBlock), 0, s, (const uint16_t*)tmp, dil, nx, ny, lo, hi);
        hipLaunchKernelGGL(k_keep_unmasked, g, dim3(kBlock), 0, s, dil, (const uint16_t*)clip, mask, n);
        src = dil;
    }
    hipLaunchKernelGGL((k_row_prefix<uint16_t, u64>), dim3((nx + 63) / 64), dim3(64), 0, s, src, nx, ny, P);
    hipLaunchKernelGGL(k_disk_mean, g, dim3(kBlock), 0, s, (const u64*)P, nx, ny, mdr, tmp);
    hipLaunchKernelGGL(k_hsl_unclip, g, dim3(kBlock), 0, s, (const uint16_t*)tmp, n, (const uint32_t*)mm, out);
    return slm_finish(stream, flags);
}

extern "C" size_t slm_map_aspect_workspace_size(uint32_t nx, uint32_t ny) {
    const size_t n = (size_t)nx * ny;
    return al256(n * 4) + 4 * al256(n) + 256 + 256;
}

extern "C" int slm_map_aspect(const float* slope, const float* uv, const uint32_t* mapping, uint32_t is_channel,
                              uint32_t nx, uint32_t ny, int r_slope, int r_uv, int do_median, float slope_threshold,
                              float* aspect, uint8_t* mask, void* workspace, size_t workspace_bytes, void* stream,
                              uint32_t flags) {
    if (!slope || !uv || !mapping || !aspect || !mask || !workspace || !nx || !ny || r_slope < 0 || r_uv < 0 ||
        (uint64_t)nx * ny >= 0x80000000ull || r_slope > 32767 || r_uv > 32767)
        return slm_fail(SLM_E_INVALID_ARG, "bad map_aspect arguments");
    if (workspace_bytes < slm_map_aspect_workspace_size(nx, ny))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes, need %zu", workspace_bytes,
                        slm_map_aspect_workspace_size(nx, ny));
    const size_t n = (size_t)nx * ny;
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    float* s0 = (float*)w;
    uint8_t* qs = w + al256(n * 4);
    uint8_t* qu = qs + al256(n);
    uint8_t* qv = qu + al256(n);
    uint8_t* med = qv + al256(n);
    uint32_t* mm = (uint32_t*)(med + al256(n));
    uint8_t* ms = med;   // the slope median (do_median) lands in med
    hipStream_t s = (hipStream_t)stream;
    const dim3 g = grid_for(n);
    // footprint counts fit uint16 for r <= 100 ((2r+1)^2 < 65536)
    if (std::max(r_slope, r_uv) > 100) return slm_fail(SLM_E_INVALID_ARG, "median disk radius above 100 pixels");
    const dim3 gm((ny + kMedThreads - 1) / kMedThreads, (nx + kRun - 1) / kRun);
    hipLaunchKernelGGL(k_slope_zero_channels, g, dim3(kBlock), 0, s, slope, mapping, is_channel, n, s0);
    hipLaunchKernelGGL(k_fill_mm, dim3(1), dim3(64), 0, s, mm);
    if (do_median) {
        hipLaunchKernelGGL(k_minmax, dim3((uint32_t)std::min<size_t>((n + kBlock - 1) / kBlock, 1024)), dim3(kBlock), 0,
                           s, (const float*)s0, n, mm);
        hipLaunchKernelGGL(k_quant_slope, g, dim3(kBlock), 0, s, (const float*)s0, n, (const uint32_t*)mm, qs);
        hipLaunchKernelGGL(k_disk_median, gm, dim3(kMedThreads), 0, s, (const uint8_t*)qs, nx, ny, r_slope, med);
        ms = med;
    }
    hipLaunchKernelGGL(k_quant_uv, g, dim3(kBlock), 0, s, (const float2*)uv, n, qu, qv);
    hipLaunchKernelGGL(k_disk_median, gm, dim3(kMedThreads), 0, s, (const uint8_t*)qu, nx, ny, r_uv, qs);
    hipLaunchKernelGGL(k_disk_median, gm, dim3(kMedThreads), 0, s, (const uint8_t*)qv, nx, ny, r_uv, qu);
    hipLaunchKernelGGL(k_aspect, g, dim3(kBlock), 0, s, (const uint8_t*)qs, (const uint8_t*)qu, (const uint8_t*)ms,
                       (const uint32_t*)mm, do_median, (const float*)s0, mapping, is_channel, slope_threshold, n,
                       aspect, mask);
    return slm_finish(stream, flags);
}
