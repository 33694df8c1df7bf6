// mapping.hip -- channel mapping and segmentation kernels of libslm_hip.so.
//
// Reference kernels (opencl/): connect.cl, channelheads.cl, countlink.cl,
// label.cl, hillslopes.cl, segment.cl, launched from
// python/streamlines/mapping.py:627-776 through connect.py, channelheads.py,
// countlink.py, label.py, hillslopes.py and segment.py.
//
// One thread per seed pixel: the work is a short un-jittered RK2 trace
// (compute_step_vec, computestep.cl:42-57) or an integer walk along link.
// Where the reference reads values that other work items write in the same
// launch (count_downchannels, flag_downchannels, segment_downchannels,
// fix_right/left_flanks) the work is split into launches with a fixed order
// of phases and order-free updates (atomic max / or / and), so every run
// gives the same bits; the phases are documented at each entry point.
//
// Workspace (slm_map_workspace_size): two 64-bit per-pixel arrays A and B
// (16 bytes per pixel) plus 256 bytes of counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "common.hpp"
#include "trace_common.hpp"

using namespace slm;

namespace {

constexpr uint32_t F_CHANNEL = SLM_IS_CHANNEL;
constexpr uint32_t F_THIN = SLM_IS_THINCHANNEL;
constexpr uint32_t F_INTER = SLM_IS_INTERCHANNEL;
constexpr uint32_t F_HEAD = SLM_IS_CHANNELHEAD;
constexpr uint32_t F_TAIL = SLM_IS_CHANNELTAIL;
constexpr uint32_t F_MAJCONF = SLM_IS_MAJORCONFLUENCE;
constexpr uint32_t F_MAJIN = SLM_IS_MAJORINFLOW;
constexpr uint32_t F_MININ = SLM_IS_MINORINFLOW;
constexpr uint32_t F_LEFT = SLM_IS_LEFTFLANK;
constexpr uint32_t F_SSH = SLM_IS_SUBSEGMENTHEAD;
constexpr uint32_t LFA = SLM_LEFT_FLANK_ADDITION;
constexpr int kBlock = 256;

typedef unsigned long long u64;

__device__ __forceinline__ u64 key2(uint32_t hi, uint32_t lo) { return ((u64)hi << 32) | (u64)lo; }

// Clears n 64-bit words (a kernel rather than a memset node, so that a
// captured graph replays it correctly).
__global__ void k_zero_u64(u64* __restrict__ a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0ull;
}

// link[i]; a link outside the grid reads as a self link (end of the chain).
__device__ __forceinline__ uint32_t lk(const SlmParams& P, const uint32_t* __restrict__ link, uint32_t i) {
    const uint32_t j = link[i];
    return j < P.nxy_padded ? j : i;
}

__device__ __forceinline__ uint32_t seed_pixel(const SlmParams& P, const float2* __restrict__ seeds, uint32_t k) {
    const float2 s = seeds[k];
    return array_idx(P, s.x, s.y);
}

// The shared head of the mapping RK steps (rungekutta.cl:280-459): dxy1
// retry of a short step, move, new pixel; returns the step length.
template <bool C>
__device__ __forceinline__ float map_move(const SlmParams& P, const StepOut& o, float& x, float& y,
                                          uint32_t& idx) {
    float2 nx = o.next;
    float dl = fast_length_cr<C>(o.dxy2.x, o.dxy2.y);
    if (dl < P.integration_halt_threshold) {
        nx = make_float2(x + o.dxy1.x, y + o.dxy1.y);
        dl = fast_length_cr<C>(o.dxy1.x, o.dxy1.y);
    }
    x = nx.x;
    y = nx.y;
    idx = array_idx(P, x, y);
    return dl;
}

// --------------------------------------------------------------- connect_channels

// connect_channels (connect.cl:26-108).  Pass 0 traces and counts the
// recorded steps n; if 2 < n < INTERCHANNEL_MAX_N_STEPS, pass 1 traces again
// (the trace is deterministic) and replays each record compress(vec-prev_vec)
// as it is produced, flagging visited pixels whose mapping is 0.  The first
// loop is bounded by INTERCHANNEL_MAX_N_STEPS as well (such a trace flags
// nothing in the reference either).
template <bool C>
__global__ void k_connect(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                          const float2* __restrict__ uv, uint32_t* __restrict__ mapping) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t cap = P.interchannel_max_n_steps;
    const float res = P.trajectory_resolution;
    const float2 s0 = seeds[k];
    const bool fixed = dt_is_fixed(P);
    uint32_t n_total = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && !(n_total > 2 && n_total < cap)) return;
        float x = s0.x, y = s0.y, px = x, py = y, dt = P.dt_max;
        uint32_t n = 0, rng = 0, idx = array_idx(P, x, y);
        const uint32_t seed_idx = idx;
        float rx = s0.x, ry = s0.y;
        uint32_t ridx = idx;
        for (int loop = 0; loop < 2; loop++) {
            for (;;) {
                if (loop == 0) {
                    if (!(seed_idx == idx && !mask[idx] && n != P.max_n_steps && n < cap)) break;
                } else {
                    if (!((mapping[idx] & F_CHANNEL) == 0 && !mask[idx] && n < cap)) break;
                }
                StepOut o;
                compute_step<C, false, false, true>(P, uv, dt, x, y, rng, o);
                idx = o.idx;
                if (mask[idx]) continue;
                const float dl = map_move<C>(P, o, x, y, idx);
                const bool stuck = dl < P.integration_halt_threshold;
                if (!stuck && !fixed) dt = next_dt(P, dt, step_error<C>(P, o));
                if (pass == 1) {
                    // replay record n: the flag loop's step n
                    if (mask[ridx] || n >= n_total - 1u) return;
                    if (mapping[ridx] == 0u) atomicOr(&mapping[ridx], F_INTER);
                    rx = rx + fdiv((float)compress1(x - px, res), res);
                    ry = ry + fdiv((float)compress1(y - py, res), res);
                    ridx = array_idx(P, rx, ry);
                }
                n++;
                if (!stuck) {
                    px = x;
                    py = y;
                }
            }
        }
        n_total = n;
    }
}

// ----------------------------------------------------------------- thinning

__global__ void k_zs_init(uint32_t nxy, const uint32_t* __restrict__ mapping, uint8_t* __restrict__ a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nxy) a[i] = (mapping[i] & (F_CHANNEL | F_INTER)) ? 1 : 0;
}

// The Zhang-Suen deletion test of a set pixel from its 8 neighbours
// q = P2..P9 = N, NE, E, SE, S, SW, W, NW (N = row x-1), subiteration sub.
__device__ __forceinline__ bool zs_delete(const int q[8], int sub) {
    int B = 0, A = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        B += q[j];
        A += (q[j] == 0 && q[(j + 1) & 7] == 1);
    }
    const int c1 = sub == 0 ? q[0] * q[2] * q[4] : q[0] * q[2] * q[6];
    const int c2 = sub == 0 ? q[2] * q[4] * q[6] : q[0] * q[4] * q[6];
    return B >= 2 && B <= 6 && A == 1 && c1 == 0 && c2 == 0;
}

// Zhang-Suen iterations in LDS tiles: a 256-thread block loads a 64x64 tile
// with a 2K-pixel halo (outside the grid reads 0) and runs K iterations (2K
// subiterations) on it; after s subiterations the outer s pixels of the
// region are stale, so the tile's own pixels are exact after all 2K.  It
// writes them to b and ORs into *iter_mask bit j when iteration j deleted one
// of them.  a -> b, ping-pong across launches.
constexpr int kZsT = 64;
constexpr int kZsK = 8;
constexpr int kZsH = 2 * kZsK;
constexpr int kZsR = kZsT + 2 * kZsH;

__global__ __launch_bounds__(256) void k_zs_tiles(uint32_t nx, uint32_t ny, const uint8_t* __restrict__ a,
                                                  uint8_t* __restrict__ b, uint32_t* __restrict__ iter_mask) {
    __shared__ uint8_t buf[2][kZsR * kZsR];
    __shared__ uint32_t s_mask;
    const uint32_t nty = (ny + kZsT - 1) / kZsT;
    const int gx0 = (int)(blockIdx.x / nty) * kZsT - kZsH, gy0 = (int)(blockIdx.x % nty) * kZsT - kZsH;
    if (threadIdx.x == 0) s_mask = 0;
    bool any = false;
    for (int k = threadIdx.x; k < kZsR * kZsR; k += blockDim.x) {
        const int gx = gx0 + k / kZsR, gy = gy0 + k % kZsR;
        uint8_t v = 0;
        if (gx >= 0 && gy >= 0 && gx < (int)nx && gy < (int)ny) v = a[(size_t)gx * ny + gy];
        buf[0][k] = v;
        any |= v != 0;
    }
    if (!__syncthreads_or(any)) {
        // nothing set in the region: the tile's pixels stay 0
        for (int k = threadIdx.x; k < kZsT * kZsT; k += blockDim.x) {
            const int gx = gx0 + kZsH + k / kZsT, gy = gy0 + kZsH + k % kZsT;
            if (gx < (int)nx && gy < (int)ny) b[(size_t)gx * ny + gy] = 0;
        }
        return;
    }
    int cur = 0;
    for (int sub = 0; sub < 2 * kZsK; sub++) {
        bool del_in = false;
        for (int k = threadIdx.x; k < kZsR * kZsR; k += blockDim.x) {
            const int lx = k / kZsR, ly = k % kZsR;
            uint8_t v = buf[cur][k];
            if (v && lx > 0 && ly > 0 && lx < kZsR - 1 && ly < kZsR - 1) {
                const uint8_t* c = &buf[cur][k];
                const int q[8] = {c[-kZsR], c[-kZsR + 1], c[1], c[kZsR + 1], c[kZsR], c[kZsR - 1], c[-1], c[-kZsR - 1]};
                if (zs_delete(q, sub & 1)) {
                    v = 0;
                    if (lx >= kZsH && ly >= kZsH && lx < kZsH + kZsT && ly < kZsH + kZsT &&
                        gx0 + lx < (int)nx && gy0 + ly < (int)ny)
                        del_in = true;
                }
            }
            buf[cur ^ 1][k] = v;
        }
        if (__syncthreads_or(del_in) && threadIdx.x == 0) s_mask |= 1u << (sub >> 1);
        cur ^= 1;
    }
    for (int k = threadIdx.x; k < kZsT * kZsT; k += blockDim.x) {
        const int lx = kZsH + k / kZsT, ly = kZsH + k % kZsT;
        const int gx = gx0 + lx, gy = gy0 + ly;
        if (gx < (int)nx && gy < (int)ny) b[(size_t)gx * ny + gy] = buf[cur][lx * kZsR + ly];
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_mask) atomicOr(iter_mask, s_mask);
}

__global__ void k_zs_final(uint32_t nxy, const uint8_t* __restrict__ a, uint32_t* __restrict__ mapping) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nxy && a[i]) mapping[i] |= F_THIN;
}

// ---------------------------------------------------------- map_channel_heads

// map_channel_heads (channelheads.cl:32-93) with channelheads_runge_kutta_step
// (rungekutta.cl:335-353: no stuck exit, dt always updated).
template <bool C>
__global__ __launch_bounds__(256) void k_heads(SlmParams P, const float2* __restrict__ seeds,
                                               const uint8_t* __restrict__ mask, const float2* __restrict__ uv,
                                               uint32_t* __restrict__ mapping) {
    float x = 0.f, y = 0.f, dt = 0.f;
    uint32_t n = 0, rng = 0, idx = 0, prev_idx = 0;
    const bool fixed = dt_is_fixed(P);
    CycleSkip cyc;
    auto finish = [&]() {
        if (n >= P.max_n_steps) return;
        if (!mask[idx]) {
            if (((~mapping[idx]) & F_THIN) || (mapping[prev_idx] & F_THIN)) atomicAnd(&mapping[idx], ~F_HEAD);
        }
    };
    auto start = [&](uint32_t k) -> bool {
        const float2 s = seeds[k];
        x = s.x;
        y = s.y;
        dt = P.dt_max;
        n = 0;
        idx = prev_idx = array_idx(P, x, y);
        cyc.reset();
        return true;
    };
    auto step = [&]() -> bool {
        if (!(prev_idx == idx && !mask[idx] && n < P.max_n_steps)) {
            finish();
            return false;
        }
        StepOut o;
        compute_step<C, false, false, true>(P, uv, dt, x, y, rng, o);
        idx = o.idx;
        map_move<C>(P, o, x, y, idx);
        n++;
        if (!fixed) dt = next_dt(P, dt, step_error<C>(P, o));
        // a repeated state inside the seed pixel never leaves it: the trace
        // runs to MAX_N_STEPS and changes nothing
        if (prev_idx == idx && cyc.check(x, y, dt, 0u, n, P.max_n_steps)) return false;
        return true;
    };
    persist_seeds(P.seeds_chunk_offset, P.n_seed_points - P.seeds_chunk_offset, start, step);
}

// --------------------------------------------------------- count_downchannels

// countlink_runge_kutta_step (rungekutta.cl:382-406): true when stuck.
// fixed: dt_is_fixed(P), so dt stays DT_MAX and the step error is not needed.
// R2: TRAJECTORY_RESOLUTION is a power of two and inv_res its reciprocal.
template <bool C, bool R2>
__device__ __forceinline__ bool countlink_step(const SlmParams& P, const float2* __restrict__ uv, float& dt,
                                               float& x, float& y, uint32_t& idx, bool fixed, float inv_res) {
    StepOut o;
    uint32_t rng = 0;
    compute_step<C, false, R2, true>(P, uv, dt, x, y, rng, o, inv_res);
    idx = o.idx;
    const float dl = map_move<C>(P, o, x, y, idx);
    if (dl < P.integration_halt_threshold) return true;
    if (!fixed) dt = next_dt(P, dt, step_error<C>(P, o));
    return false;
}

// count_downchannels (countlink.cl:34-114).  PASS 0: full traces rank each
// pixel by max (counter, seed) in keyA.  PASS 1: traces stop on entering a
// pixel ranked higher by another trace, flag it, count = max, and offer
// link keys (counter at prev, next pixel) into keyL.
template <bool C, int PASS, bool R2>
__global__ void k_count(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                        const float2* __restrict__ uv, uint32_t* __restrict__ mapping,
                        uint32_t* __restrict__ count, u64* __restrict__ keyA, u64* __restrict__ keyL) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t cap = 1000u * P.max_n_steps;
    const bool fixed = dt_is_fixed(P);
    const float inv_res = 1.0f / P.trajectory_resolution;
    const float2 s = seeds[k];
    float x = s.x, y = s.y, dt = P.dt_max;
    uint32_t idx = array_idx(P, x, y), prev = idx, n = 0, counter = 2, cprev = 1;
    if (PASS == 0)
        atomicMax(&keyA[idx], key2(1u, k));
    else
        atomicMax(&count[idx], 1u);
    // Brent cycle detection on (x, y, dt, prev); see k_flag_walk for the skip rule
    float sx = x, sy = y, sdt = dt;
    uint32_t sprev = prev, sn = n, scounter = counter, power = 1;
    bool ff = true;
    while (!mask[idx] && n < cap) {
        if (countlink_step<C, R2>(P, uv, dt, x, y, idx, fixed, inv_res)) break;
        n++;
        if (prev != idx) {
            ++counter;
            if (PASS == 0) {
                atomicMax(&keyA[idx], key2(counter, k));
            } else {
                atomicAnd(&mapping[idx], ~F_HEAD);
                atomicOr(&mapping[idx], F_THIN);
                atomicMax(&keyL[prev], key2(cprev, idx));
                const u64 a = keyA[idx];
                if (counter < (uint32_t)(a >> 32) && (uint32_t)a != k) break;
                atomicMax(&count[idx], counter);
                cprev = counter;
            }
            prev = idx;
        }
        if (ff) {
            if (x == sx && y == sy && dt == sdt && prev == sprev) {
                const uint32_t lam = n - sn, clap = counter - scounter, laps = (cap - n) / lam;
                if (laps >= 2) {
                    n += (laps - 1) * lam;
                    counter += (laps - 1) * clap;
                    cprev += (laps - 1) * clap;
                }
                ff = false;
            } else if (n - sn == power) {
                sx = x;
                sy = y;
                sdt = dt;
                sprev = prev;
                sn = n;
                scounter = counter;
                power <<= 1;
            }
        }
    }
    if (PASS == 1) atomicMax(&keyL[prev], key2(cprev, idx));
}

__global__ void k_link_final(uint32_t nxy, const u64* __restrict__ keyL, uint32_t* __restrict__ link) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nxy) {
        const u64 v = keyL[i];
        if (v) link[i] = (uint32_t)v;
    }
}

// ---------------------------------------------------------- flag_downchannels

__global__ void k_flag_init(SlmParams P, const float2* __restrict__ seeds, uint32_t* __restrict__ mapping,
                            uint32_t* __restrict__ count) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t h = seed_pixel(P, seeds, k);
    count[h] = 1u;
    atomicOr(&mapping[h], F_THIN);
}

// flag_downchannels (countlink.cl:134-210).  PHASE 2: walks (flags, count
// max, tail).  PHASE 3: the head test on the final counts.
template <int PHASE>
__global__ void k_flag_walk(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                            uint32_t* __restrict__ mapping, uint32_t* __restrict__ count,
                            const uint32_t* __restrict__ link) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t cap = 1000u * P.max_n_steps;
    const uint32_t h = seed_pixel(P, seeds, k);
    uint32_t idx = h, prev = h + 1u, counter = 1;
    // A walk that returns to a pixel is on a link cycle and would circle to the
    // hop cap.  Its updates are or and max of a counter that grows by the cycle
    // length each lap, so all laps but the last two are dominated: skip them
    // (Brent's cycle detection on the pixel).  Same bits as walking every hop.
    uint32_t sidx = idx, scounter = counter, power = 1;
    bool ff = PHASE == 2;
    while (!mask[idx] && prev != idx && counter < cap) {
        prev = idx;
        idx = lk(P, link, idx);
        counter++;
        if (mask[idx]) break;
        if (PHASE == 3) {
            if (counter >= 20u) break;
            if ((mapping[idx] & F_MININ) || (counter == 2u && count[idx] > 2u)) atomicAnd(&mapping[h], ~F_HEAD);
        } else {
            atomicOr(&mapping[idx], F_THIN);
            atomicMax(&count[idx], counter);
        }
        if (ff) {
            if (idx == sidx && counter != scounter) {
                const uint32_t lam = counter - scounter, laps = (cap - counter) / lam;
                if (laps >= 2) counter += (laps - 1) * lam;
                ff = false;
            } else if (counter - scounter == power) {
                sidx = idx;
                scounter = counter;
                power <<= 1;
            }
        }
    }
    if (PHASE == 2 && prev < P.nxy_padded && !mask[prev] && counter < cap) atomicOr(&mapping[prev], F_TAIL);
}

// ---------------------------------------------------------- label_confluences

// label_confluences (label.cl:21-134).
__global__ void k_confluences(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                              const float* __restrict__ slt, uint32_t* __restrict__ mapping,
                              const uint32_t* __restrict__ count, const uint32_t* __restrict__ link) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const float2 s = seeds[k];
    const uint32_t idx = array_idx(P, s.x, s.y);
    // N, S, E, W, NE, SE, NW, SW
    const float ox[8] = {0.f, 0.f, 1.f, -1.f, 1.f, 1.f, -1.f, -1.f};
    const float oy[8] = {1.f, -1.f, 0.f, 0.f, 1.f, -1.f, 1.f, -1.f};
    uint32_t in[8], eq[8], n_in = 0, n_eq = 0, dom_idx = 0;
    float dom = -3.402823466e+38f;
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const uint32_t nb = array_idx(P, s.x + ox[d], s.y + oy[d]);
        if (!mask[nb] && (mapping[nb] & F_THIN) && link[nb] == idx) in[n_in++] = nb;
    }
    if (n_in > 1) {
        atomicOr(&mapping[idx], F_MAJCONF);
        const uint32_t ci = count[idx];
        for (uint32_t i = 0; i < n_in; i++) {
            if (ci != count[in[i]] + 1u) {
                atomicOr(&mapping[in[i]], F_MININ);
            } else {
                eq[n_eq++] = in[i];
                if (slt[in[i]] > dom) {
                    dom_idx = in[i];
                    dom = slt[dom_idx];
                }
            }
        }
        for (uint32_t i = 0; i < n_eq; i++) atomicOr(&mapping[eq[i]], eq[i] == dom_idx ? F_MAJIN : F_MININ);
    }
}

// ------------------------------------------------------- segment_downchannels

__device__ __forceinline__ u64 own_key(const uint32_t* __restrict__ mapping, const uint32_t* __restrict__ count,
                                       uint32_t prev) {
    return ((u64)count[prev] << 32) | ((mapping[prev] & F_MAJIN) ? 0x80000000ull : 0ull) |
           (u64)(prev & 0x7fffffffu);
}

__global__ void k_seg_heads(SlmParams P, const float2* __restrict__ seeds, uint32_t* __restrict__ mapping,
                            uint8_t* __restrict__ mark) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t h = seed_pixel(P, seeds, k);
    atomicOr(&mapping[h], F_SSH);
    mark[h] = 1;
}

// A full walk that returns to a pixel is on a link cycle and would circle to
// its hop cap (a cycle of the medial-axis skeleton's links, for one).
// Brent's detection on the pixel fires once the walk has gone round the whole
// cycle from a point on it; every ownership offer after that repeats one
// already made, so the walk stops there: the same bits as walking to the cap.
// (The owned walks of the later phases cannot circle: the pixel where a walk
// entered a cycle is owned by one of its two inflows only.)
struct LinkCycle {
    uint32_t sidx, shop, power;
    __device__ LinkCycle(uint32_t idx) : sidx(idx), shop(0), power(1) {}
    // after the hop that reached idx (hop count, from 1): true on detection
    __device__ bool step(uint32_t idx, uint32_t hop) {
        if (idx == sidx) return true;
        if (hop - shop == power) {
            sidx = idx;
            shop = hop;
            power <<= 1;
        }
        return false;
    }
};

// Full walks: own[q] = max own_key(prev) over hops prev -> q.
__global__ void k_seg_own(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                          const uint32_t* __restrict__ mapping, const uint32_t* __restrict__ count,
                          const uint32_t* __restrict__ link, u64* __restrict__ own) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    uint32_t prev = seed_pixel(P, seeds, k), idx = lk(P, link, prev);
    LinkCycle cyc(prev);
    for (uint32_t hop = 0; hop < P.nxy_padded && !mask[idx] && prev != idx && !(mapping[idx] & F_SSH); hop++) {
        atomicMax(&own[idx], own_key(mapping, count, prev));
        if (cyc.step(idx, hop + 1u)) break;
        prev = idx;
        idx = lk(P, link, idx);
    }
}

// PHASE 2: owned walks mark segment starts (also where they join another
// walk's reach, then stop).  PHASE 3: owned walks label.
template <int PHASE>
__global__ void k_seg_walk(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                           const uint32_t* __restrict__ mapping, const uint32_t* __restrict__ count,
                           const uint32_t* __restrict__ link, const u64* __restrict__ own,
                           uint8_t* __restrict__ mark, uint32_t* __restrict__ label) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t T = P.segmentation_threshold;
    const uint32_t h = seed_pixel(P, seeds, k);
    uint32_t prev = h, idx = lk(P, link, h), seg_label = h, segc = 0;
    if (PHASE == 3) atomicMax(&label[h], h);
    for (uint32_t hop = 0; hop < P.nxy_padded && !mask[idx] && prev != idx && !(mapping[idx] & F_SSH); hop++) {
        const bool owned = own[idx] == own_key(mapping, count, prev);
        if (PHASE == 2) {
            const uint32_t c = count[idx];
            if (c >= segc + T / 2u) {
                mark[idx] = 1;
                segc = c + T;
            }
            if (!owned) break;
        } else {
            if (!owned) break;
            if (mark[idx]) seg_label = idx;
            atomicMax(&label[idx], seg_label);
        }
        prev = idx;
        idx = lk(P, link, idx);
    }
}

__global__ void k_seg_final(uint32_t nxy, const uint8_t* __restrict__ mark, uint32_t* __restrict__ mapping) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nxy && mark[i]) mapping[i] |= F_SSH;
}

// ------------------------------------------------------------ link_hillslopes

// hillslopes (hillslopes.cl:31-91).
template <bool C, bool R2>
__global__ void k_link_hs(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                          const float2* __restrict__ uv, uint32_t* __restrict__ link) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t cap = 100u * P.max_n_steps;
    const bool fixed = dt_is_fixed(P);
    const float inv_res = 1.0f / P.trajectory_resolution;
    const float2 s = seeds[k];
    float x = s.x, y = s.y, dt = P.dt_max;
    uint32_t idx = array_idx(P, x, y), prev = idx, n = 0;
    float sx = x, sy = y, sdt = dt;
    uint32_t sn = 0, power = 1;
    while (prev == idx && !mask[idx] && n < cap) {
        prev = idx;
        if (countlink_step<C, R2>(P, uv, dt, x, y, idx, fixed, inv_res)) break;
        n++;
        if (prev == idx) {
            // a repeated state inside the seed pixel never leaves it: no link
            if (x == sx && y == sy && dt == sdt) return;
            if (n - sn == power) {
                sx = x;
                sy = y;
                sdt = dt;
                sn = n;
                power <<= 1;
            }
        }
    }
    if (prev != idx && !mask[idx]) link[prev] = idx;
}

// ------------------------------------------------------ hillslope segmentation

// segment_runge_kutta_step (rungekutta.cl:435-459); probe = the pixel of
// compute_step_vec's next point, before the dxy1 retry.  fixed:
// dt_is_fixed(P), so dt stays DT_MAX and the step error is not needed.
template <bool C, bool R2>
__device__ __forceinline__ bool segment_step(const SlmParams& P, const float2* __restrict__ uv, float& dt,
                                             float& x, float& y, uint32_t& idx, uint32_t& n, uint32_t& probe,
                                             float inv_res, bool fixed) {
    StepOut o;
    uint32_t rng = 0;
    compute_step<C, false, R2, true>(P, uv, dt, x, y, rng, o, inv_res);
    probe = o.idx;
    const float dl = map_move<C>(P, o, x, y, idx);
    n += 1u;
    if (dl < P.integration_halt_threshold) return true;
    if (!fixed) dt = next_dt(P, dt, step_error<C>(P, o));
    return false;
}

// segment_hillslopes (segment.cl:133-188).  FLANKS: subsegment_flanks
// (:360-414) in the same walk, for seeds = the unmasked non-thin pixels once
// subsegment_channel_edges has flagged the left flanks.  The flank walk is
// this walk up to its first left-flank pixel (the same steps and the same
// cycle skips, and it also stops wherever this one does), so
// LEFT_FLANK_ADDITION is decided there and ORed into the label written at
// the end, as the reference writes the channel's label and ORs the flag
// after (a seed that is itself a left flank, flagged by the channel edges,
// stops its flank walk at once).
template <bool C, bool R2, bool FLANKS>
__global__ __launch_bounds__(256) void k_seg_hs(SlmParams P, const float2* __restrict__ seeds,
                                                const uint8_t* __restrict__ mask, const float2* __restrict__ uv,
                                                const uint32_t* __restrict__ mapping,
                                                uint32_t* __restrict__ label) {
    const float inv_res = 1.0f / P.trajectory_resolution;
    const bool fixed = dt_is_fixed(P);
    float x = 0.f, y = 0.f, dt = 0.f;
    uint32_t n = 0, probe = 0, idx = 0, hs = 0;
    bool open = false, lfa = false;   // the flank walk goes on; its flag
    CycleSkip cyc;
    auto finish = [&]() {
        if (FLANKS && open) lfa = (mapping[idx] & F_LEFT) != 0u;
        if (mapping[idx] & F_THIN) label[hs] = label[idx] | (lfa ? LFA : 0u);
        else if (lfa) label[hs] |= LFA;
    };
    auto start = [&](uint32_t k) -> bool {
        const float2 s = seeds[k];
        x = s.x;
        y = s.y;
        dt = P.dt_max;
        n = 0;
        cyc.reset();
        idx = hs = array_idx(P, x, y);
        open = FLANKS;
        lfa = false;
        return true;
    };
    auto step = [&]() -> bool {
        if (FLANKS && open &&
            (mask[idx] || (mapping[idx] & (F_LEFT | F_THIN)) || n >= P.max_n_steps)) {
            open = false;
            lfa = (mapping[idx] & F_LEFT) != 0u;
        }
        if (!(!mask[idx] && ((~mapping[idx]) & F_THIN) && n < P.max_n_steps) ||
            segment_step<C, R2>(P, uv, dt, x, y, idx, n, probe, inv_res, fixed)) {
            finish();
            return false;
        }
        cyc.check(x, y, dt, 0u, n, P.max_n_steps);
        return true;
    };
    persist_seeds(P.seeds_chunk_offset, P.n_seed_points - P.seeds_chunk_offset, start, step);
}

// rotate_left (segment.cl:195-200) in char arithmetic (abs of a char is a
// uchar); the index is computed in uint arithmetic like the reference's.
__device__ __forceinline__ uint32_t rotate_left(const SlmParams& P, uint32_t px, uint32_t py, int8_t& dx,
                                                int8_t& dy) {
    const int8_t rx = (int8_t)(dx - dy), ry = (int8_t)(dx + dy);
    int ax = (int8_t)(uint8_t)(rx < 0 ? -rx : rx), ay = (int8_t)(uint8_t)(ry < 0 ? -ry : ry);
    ax = ax < 1 ? 1 : (ax > 2 ? 2 : ax);
    ay = ay < 1 ? 1 : (ay > 2 ? 2 : ay);
    dx = (int8_t)(rx / ax);
    dy = (int8_t)(ry / ay);
    return (px + (uint32_t)(int32_t)dx) * P.ny_padded + (py + (uint32_t)(int32_t)dy);
}

// subsegment_channel_edges (segment.cl:226-334); an index outside the grid
// reads as masked.
__global__ void k_sub_edges(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                            uint32_t* __restrict__ mapping, const uint32_t* __restrict__ link,
                            uint32_t* __restrict__ label) {
    const uint32_t k = P.seeds_chunk_offset + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.n_seed_points) return;
    const uint32_t nyp = P.ny_padded, nxy = P.nxy_padded;
    uint32_t prev = seed_pixel(P, seeds, k), idx = lk(P, link, prev), n = 0;
    while (prev != idx && n++ < P.max_n_steps) {
        const uint32_t px = prev / nyp, py = prev % nyp, x = idx / nyp, y = idx % nyp;
        int8_t dx = (int8_t)(x - px), dy = (int8_t)(y - py);
        uint32_t left = idx, turns = 0;
        bool do_left = true;
        while (left != prev && ++turns <= 7) {
            left = rotate_left(P, px, py, dx, dy);
            if (turns >= 2 && left < nxy && !mask[left] && (mapping[left] & F_THIN)) {
                if (left == lk(P, link, idx)) do_left = false;
                break;
            }
        }
        if (do_left) {
            turns = 0;
            left = idx;
            dx = (int8_t)(x - px);
            dy = (int8_t)(y - py);
            while (left != prev && ++turns <= 4) {
                left = rotate_left(P, px, py, dx, dy);
                if (turns >= 2 && left < nxy && !mask[left]) {
                    if (mapping[left] & F_THIN) break;
                    atomicOr(&mapping[left], F_LEFT);
                    atomicOr(&label[left], LFA);
                }
            }
        }
        prev = idx;
        idx = lk(P, link, idx);
        if (mask[idx] || (mapping[prev] & F_SSH)) break;
    }
}

// subsegment_flanks (segment.cl:360-414).
template <bool C, bool R2>
__global__ __launch_bounds__(256) void k_sub_flanks(SlmParams P, const float2* __restrict__ seeds,
                                                    const uint8_t* __restrict__ mask,
                                                    const float2* __restrict__ uv,
                                                    const uint32_t* __restrict__ mapping,
                                                    uint32_t* __restrict__ label) {
    const float inv_res = 1.0f / P.trajectory_resolution;
    const bool fixed = dt_is_fixed(P);
    float x = 0.f, y = 0.f, dt = 0.f;
    uint32_t n = 0, probe = 0, idx = 0, hs = 0;
    CycleSkip cyc;
    auto start = [&](uint32_t k) -> bool {
        const float2 s = seeds[k];
        x = s.x;
        y = s.y;
        dt = P.dt_max;
        n = 0;
        cyc.reset();
        idx = hs = array_idx(P, x, y);
        return true;
    };
    auto step = [&]() -> bool {
        if (!(!mask[idx] && !(mapping[idx] & F_LEFT) && !(mapping[idx] & F_THIN) && n < P.max_n_steps) ||
            segment_step<C, R2>(P, uv, dt, x, y, idx, n, probe, inv_res, fixed)) {
            if (mapping[idx] & F_LEFT) label[hs] |= LFA;
            return false;
        }
        cyc.check(x, y, dt, 0u, n, P.max_n_steps);
        return true;
    };
    persist_seeds(P.seeds_chunk_offset, P.n_seed_points - P.seeds_chunk_offset, start, step);
}

// fix_right_flanks (segment.cl:441-504) / fix_left_flanks (:530-594); other
// pixels' labels are read from the snapshot taken before the launch.
template <bool C, bool LEFT, bool R2>
__global__ __launch_bounds__(256) void k_fix(SlmParams P, const float2* __restrict__ seeds,
                                             const uint8_t* __restrict__ mask, const float2* __restrict__ uv,
                                             const uint32_t* __restrict__ mapping,
                                             const uint32_t* __restrict__ snap, uint32_t* __restrict__ label) {
    const float inv_res = 1.0f / P.trajectory_resolution;
    const bool fixed = dt_is_fixed(P);
    float x = 0.f, y = 0.f, dt = 0.f;
    uint32_t n = 0, idx = 0, hs = 0, own = 0;
    auto lab = [&](uint32_t q) -> uint32_t { return q == hs ? own : snap[q]; };
    CycleSkip cyc;
    auto finish = [&]() {
        if ((own & ~LFA) != (lab(idx) & ~LFA)) own = lab(idx);
        if (LEFT) {
            if (!(lab(idx) & LFA) && !(mapping[idx] & F_THIN)) own &= ~LFA;
        } else {
            if (lab(idx) & LFA) own |= LFA;
        }
        label[hs] = own;
    };
    auto start = [&](uint32_t k) -> bool {
        const float2 s = seeds[k];
        x = s.x;
        y = s.y;
        dt = P.dt_max;
        n = 0;
        cyc.reset();
        idx = hs = array_idx(P, x, y);
        own = snap[hs];
        return true;
    };
    auto step = [&]() -> bool {
        const bool side = LEFT ? ((lab(idx) & LFA) != 0u) : ((lab(idx) & LFA) == 0u);
        if (!(!mask[idx] && side && !(mapping[idx] & F_THIN) && n < P.max_n_steps)) {
            finish();
            return false;
        }
        uint32_t probe;
        const bool stuck = segment_step<C, R2>(P, uv, dt, x, y, idx, n, probe, inv_res, fixed);
        if ((own & ~LFA) != (lab(probe) & ~LFA)) own = lab(probe);
        if (stuck) {
            finish();
            return false;
        }
        cyc.check(x, y, dt, own, n, P.max_n_steps);
        return true;
    };
    persist_seeds(P.seeds_chunk_offset, P.n_seed_points - P.seeds_chunk_offset, start, step);
}

// ------------------------------------------------------------------ host side

// ---------------------------------------------------------- mid-slopes and ridges
// Mapping.map_midslopes / map_ridges (mapping.py:779-817): the sla angle
// arctan2(dsla, usla) - pi/4 (float32), smoothed by scipy's gaussian_filter
// (separable, axis 0 then axis 1, float32 between the passes; each output a
// float64 sum: centre * w0, then (left + right) * w_j from the outermost pair
// in; 'reflect' borders), thresholded, for ridges dilated twice by the 3x3
// square and hole-filled, then thinned; the skeletons flag the mapping grid.

__global__ void k_sla_angle(const float2* __restrict__ sla, uint32_t nxy, float* __restrict__ ang) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    // np.arctan2 on float32, correctly rounded here (through float64), minus float32(pi/4)
    const float a = (float)atan2((double)sla[i].x, (double)sla[i].y);
    ang[i] = a - 0.785398185253143310546875f;
}

// scipy.ndimage 'reflect' extension of index k into [0, n)
__device__ __forceinline__ int reflect_idx(int k, int n) {
    const int p = 2 * n;
    k %= p;
    if (k < 0) k += p;
    return k < n ? k : p - 1 - k;
}

template <int AXIS>
__global__ void k_gauss(const float* __restrict__ in, float* __restrict__ out, uint32_t nx, uint32_t ny,
                        const double* __restrict__ w, int r) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const int x = (int)(i / ny), y = (int)(i % ny);
    const int n = AXIS == 0 ? (int)nx : (int)ny, pos = AXIS == 0 ? x : y;
    auto at = [&](int k) -> double {
        const int kk = reflect_idx(k, n);
        return (double)in[AXIS == 0 ? (size_t)kk * ny + y : (size_t)x * ny + kk];
    };
    double acc = at(pos) * w[r];
    for (int j = r; j >= 1; j--) acc += (at(pos - j) + at(pos + j)) * w[r - j];
    out[i] = (float)acc;
}

__global__ void k_slope_marks(const float* __restrict__ gm, const float* __restrict__ gr,
                              const uint8_t* __restrict__ mask, uint32_t nxy, float mid_thr, float ridge_level,
                              uint8_t* __restrict__ mid, uint8_t* __restrict__ ridge) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    const bool ok = !mask[i];
    mid[i] = (ok && fabsf(gm[i]) <= mid_thr) ? 1 : 0;
    ridge[i] = (ok && gr[i] <= ridge_level) ? 1 : 0;
}

// binary_dilation by the 3x3 square, border value 0
__global__ void k_dilate8(const uint8_t* __restrict__ a, uint8_t* __restrict__ b, uint32_t nx, uint32_t ny) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const int x = (int)(i / ny), y = (int)(i % ny);
    uint8_t v = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++) {
            const int xx = x + dx, yy = y + dy;
            if (xx >= 0 && yy >= 0 && xx < (int)nx && yy < (int)ny && a[(size_t)xx * ny + yy]) v = 1;
        }
    b[i] = v;
}

// binary_fill_holes: the background reached from outside the grid through
// 4-neighbours (scipy's fill dilates the inverted image from a border of 1s);
// the rest of the background is a hole.  Bit-packed rows: word x*nwy + w, bit
// b is pixel (x, 64w + b).  One wave owns a 64x64 tile, one row word per
// lane, and closes the tile with Kogge-Stone occluded fills (the shift-and-
// propagate doubling of bitboard flood fills): along a row within the word,
// across rows with lane shuffles, alternating until nothing grows.  The tile
// is seeded by the grid border and by its neighbours' current reach bits;
// launches repeat until no tile changes.  Reach only grows, so tiles may read
// neighbours that are being updated: the fixpoint, the set of background
// pixels 4-connected to an open border pixel, does not depend on the order.

// background bits; bits past ny are 0
__global__ void k_pack_open(const uint8_t* __restrict__ fat, uint32_t nx, uint32_t ny, uint32_t nwy,
                            u64* __restrict__ open, u64* __restrict__ reach) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * nwy) return;
    const uint32_t x = i / nwy, w = i % nwy;
    const uint32_t y0 = w * 64u, y1 = min(ny, y0 + 64u);
    const uint8_t* row = fat + (size_t)x * ny;
    u64 o = 0;
    for (uint32_t y = y0; y < y1; y++)
        if (!row[y]) o |= 1ull << (y - y0);
    open[i] = o;
    reach[i] = 0;
}

__device__ __forceinline__ u64 fill_row(u64 x, u64 o) {
    u64 a = x, g = o;                      // towards higher y
    a |= g & (a << 1);  g &= g << 1;
    a |= g & (a << 2);  g &= g << 2;
    a |= g & (a << 4);  g &= g << 4;
    a |= g & (a << 8);  g &= g << 8;
    a |= g & (a << 16); g &= g << 16;
    a |= g & (a << 32);
    u64 b = x, h = o;                      // towards lower y
    b |= h & (b >> 1);  h &= h >> 1;
    b |= h & (b >> 2);  h &= h >> 2;
    b |= h & (b >> 4);  h &= h >> 4;
    b |= h & (b >> 8);  h &= h >> 8;
    b |= h & (b >> 16); h &= h >> 16;
    b |= h & (b >> 32);
    return a | b;
}

// the same across the 64 lanes (rows) of the wave
__device__ __forceinline__ u64 fill_col(u64 x, u64 o, uint32_t lane) {
    u64 a = x, g = o, b = x, h = o;
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        u64 as = __shfl_up(a, d, 64), gs = __shfl_up(g, d, 64);
        if (lane < d) as = gs = 0;
        a |= g & as;
        g &= gs;
        u64 bs = __shfl_down(b, d, 64), hs = __shfl_down(h, d, 64);
        if (lane + d > 63u) bs = hs = 0;
        b |= h & bs;
        h &= hs;
    }
    return a | b;
}

__global__ __launch_bounds__(256) void k_fill_tiles(const u64* __restrict__ open, u64* __restrict__ reach,
                                                     uint32_t nx, uint32_t ny, uint32_t nwy, uint32_t n_tiles,
                                                     uint32_t* __restrict__ changed) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (t >= n_tiles) return;                       // whole waves leave together
    const uint32_t tx = t / nwy, w = t % nwy;
    const uint32_t x0 = tx * 64u, x = x0 + lane;
    const bool row_ok = x < nx;
    const size_t wi = (size_t)x * nwy + w;
    const u64 o = row_ok ? open[wi] : 0ull;
    const u64 r_in = row_ok ? reach[wi] : 0ull;
    u64 seed = 0;
    if (row_ok) {
        if (x == 0 || x + 1 == nx) seed = o;        // first and last grid rows
        if (w == 0) seed |= o & 1ull;               // y = 0
        if (w + 1 == nwy) seed |= o & (1ull << ((ny - 1u) & 63u));   // y = ny - 1
        if (w > 0 && (reach[wi - 1] >> 63)) seed |= o & 1ull;
        if (w + 1 < nwy && (reach[wi + 1] & 1ull)) seed |= o & (1ull << 63);
        if (lane == 0 && x0 > 0) seed |= o & reach[wi - nwy];
        if (lane == 63u && x + 1 < nx) seed |= o & reach[wi + nwy];
    }
    u64 r = (r_in | seed) & o;
    while (true) {
        const u64 prev = r;
        r = fill_col(fill_row(r, o), o, lane);
        if (!__any(r != prev)) break;
    }
    const bool grew = r != r_in;
    if (row_ok && grew) reach[wi] = r;
    if (__any(grew) && lane == 0) atomicAdd(changed, 1u);
}

__global__ void k_fill_final(uint8_t* __restrict__ fat, const u64* __restrict__ reach, uint32_t nx, uint32_t ny,
                             uint32_t nwy) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const uint32_t x = i / ny, y = i % ny;
    const bool reached = (reach[(size_t)x * nwy + (y >> 6)] >> (y & 63u)) & 1ull;
    fat[i] = (fat[i] || !reached) ? 1 : 0;
}

__global__ void k_flag_marks(const uint8_t* __restrict__ a, uint32_t nxy, uint32_t flag,
                             uint32_t* __restrict__ mapping) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nxy && a[i]) mapping[i] |= flag;
}

int check_grid(const SlmParams* p) {
    int rc = slm_check_params(p);
    if (rc) return rc;
    if (p->nxy_padded >= 0x80000000u) return slm_fail(SLM_E_INVALID_ARG, "nxy_padded must be below 2^31");
    return SLM_OK;
}

int check_map(const SlmParams* p, const void* seeds) {
    int rc = check_grid(p);
    if (rc) return rc;
    if (p->n_seed_points > p->seeds_chunk_offset && !seeds) return slm_fail(SLM_E_INVALID_ARG, "seeds is NULL");
    return SLM_OK;
}

// ------------------------------------------------------------------ medial axis
//
// scikit-image's medial_axis (the first step of the reference's channel
// thinning, skeletonize(medial_axis(.)), mapping.py:717), as published: every
// foreground pixel is visited once, in lexicographic order of (squared
// distance to the background, corner score = 9 - set pixels of its 3x3
// window, a tie-breaker permutation from the host), and takes the keep-table
// value of its current 3x3 window (pixels outside the grid unset, pixels not
// yet visited as they started).  A visit reads only the pixels visited before
// it, so a pixel can be decided as soon as every neighbour that precedes it
// has been: rounds decide all ready pixels at once, in any order, and give
// the sequential visit's result.  Each pixel's order key packs the three
// fields into one 64-bit integer.
constexpr uint32_t kMaInf = 0x3fffffffu;

// The keep table as 512 bits, passed by value.
struct MaTable {
    uint32_t w[16];
};

// Distance along x to the nearest unset pixel of the same column (y), both
// ways; kMaInf when the column has none.  One thread per column.
__global__ void k_ma_cols(const uint8_t* __restrict__ a, uint32_t nx, uint32_t ny, uint32_t* __restrict__ f) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ny) return;
    uint32_t last = kMaInf;
    for (uint32_t x = 0; x < nx; x++) {
        const size_t i = (size_t)x * ny + y;
        if (!a[i]) last = x;
        f[i] = (last == kMaInf) ? kMaInf : x - last;
    }
    last = kMaInf;
    for (uint32_t x = nx; x-- > 0;) {
        const size_t i = (size_t)x * ny + y;
        if (!a[i]) last = x;
        if (last != kMaInf && last - x < f[i]) f[i] = last - x;
    }
}

// Squared Euclidean distance of each set pixel to the nearest unset one:
// min over columns y' of (y - y')^2 + f[x][y']^2, searched outward from y
// until (y - y')^2 alone is no better.
__global__ void k_ma_d2(const uint8_t* __restrict__ a, const uint32_t* __restrict__ f, uint32_t nx, uint32_t ny,
                        u64* __restrict__ key) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    if (!a[i]) {
        key[i] = 0;
        return;
    }
    const uint32_t x = i / ny, y = i % ny;
    const size_t row = (size_t)x * ny;
    u64 best = (f[i] == kMaInf) ? ~0ull : (u64)f[i] * f[i];
    for (uint32_t dy = 1; dy < ny; dy++) {
        const u64 dd = (u64)dy * dy;
        if (dd >= best) break;
        if (y >= dy) {
            const uint32_t g = f[row + y - dy];
            if (g != kMaInf && dd + (u64)g * g < best) best = dd + (u64)g * g;
        }
        if (y + dy < ny) {
            const uint32_t g = f[row + y + dy];
            if (g != kMaInf && dd + (u64)g * g < best) best = dd + (u64)g * g;
        }
    }
    if (best > 0x0fffffffull) best = 0x0fffffffull;    // no unset pixel anywhere: last
    key[i] = best << 36;
}

// Set pixels per 256-pixel block, then their exclusive prefix (one block).
__global__ void k_ma_bcount(const uint8_t* __restrict__ a, uint32_t n, uint32_t* __restrict__ bsum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool set = i < n && a[i];
    const uint32_t c = __syncthreads_count(set);
    if (threadIdx.x == 0) bsum[blockIdx.x] = c;
}

__global__ __launch_bounds__(1024) void k_ma_bscan(uint32_t* __restrict__ bsum, uint32_t nb, uint32_t* __restrict__ total) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023u) / 1024u;
    const uint32_t b = min(nb, t * per), e = min(nb, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += bsum[i];
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
        const uint32_t c = bsum[i];
        bsum[i] = run;
        run += c;
    }
    if (t == 1023) *total = part[1023];
}

// Row-major rank of each set pixel: the order key's tie-breaker is
// tie[rank] (tie: the host's permutation of [0, n_set)); corner score; the
// list of set pixels in rank order; state 0 (to visit) or 3 (unset).
__global__ void k_ma_keys(const uint8_t* __restrict__ a, uint32_t nx, uint32_t ny,
                          const uint32_t* __restrict__ bsum, const uint32_t* __restrict__ tie, uint32_t n_set,
                          u64* __restrict__ key, uint32_t* __restrict__ list, uint8_t* __restrict__ state) {
    __shared__ uint32_t wsum[4];
    const uint32_t n = nx * ny;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool set = i < n && a[i];
    const u64 m = __ballot(set);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = bsum[blockIdx.x];
    for (uint32_t k = 0; k < w; k++) off += wsum[k];
    if (i >= n) return;
    if (!set) {
        state[i] = 3;
        return;
    }
    const uint32_t rank = off + below;
    if (rank >= n_set) {      // the host's count disagrees: leave the pixel unvisited (the call fails)
        state[i] = 3;
        return;
    }
    const int x = (int)(i / ny), y = (int)(i % ny);
    uint32_t cnt = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++) {
            const int xx = x + dx, yy = y + dy;
            if (xx >= 0 && yy >= 0 && xx < (int)nx && yy < (int)ny && a[(size_t)xx * ny + yy]) cnt++;
        }
    key[i] |= ((u64)(9u - cnt) << 32) | tie[rank];
    list[rank] = i;
    state[i] = 0;
}

// One round: every set pixel still to visit whose preceding neighbours are
// all decided takes table[window]; state 1 kept, 2 removed.  pending counts
// the pixels left undecided.
__global__ void k_ma_round(uint32_t nx, uint32_t ny, const uint32_t* __restrict__ list, uint32_t n_set,
                           const u64* __restrict__ key, uint8_t* __restrict__ state, MaTable table,
                           uint32_t* __restrict__ pending) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_set) return;
    const uint32_t i = list[r];
    if (state[i] != 0) return;
    const u64 k = key[i];
    const int x = (int)(i / ny), y = (int)(i % ny);
    uint32_t acc = 16u;
    int bit = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++, bit++) {
            if (!dx && !dy) continue;
            const int xx = x + dx, yy = y + dy;
            if (xx < 0 || yy < 0 || xx >= (int)nx || yy >= (int)ny) continue;
            const size_t q = (size_t)xx * ny + yy;
            const uint8_t st = __hip_atomic_load(&state[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (st == 0 && key[q] < k) {       // a preceding neighbour not visited yet
                atomicAdd(pending, 1u);
                return;
            }
            if (st == 0 || st == 1) acc |= 1u << bit;
        }
    const uint8_t v = (table.w[acc >> 5] >> (acc & 31u)) & 1u ? 1 : 2;
    __hip_atomic_store(&state[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The medial axis back into the byte image: set where kept.
__global__ void k_ma_final(uint32_t n, const uint8_t* __restrict__ state, uint8_t* __restrict__ a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = state[i] == 1 ? 1 : 0;
}

// Workspace: two 8-byte-per-pixel areas (the byte images of the thinning,
// the keys of the count and segment phases), the counters at 16 B per pixel,
// then the medial axis' order keys, column distances, states, visit list and
// block sums.
inline size_t ma_al(size_t v) { return (v + 255) & ~(size_t)255; }
struct MaWs {
    size_t key, f, state, list, bsum, total, end;
};
MaWs ma_ws(size_t n) {
    MaWs w;
    w.key = ma_al(16 * n + 256);
    w.f = w.key + ma_al(n * 8);
    w.state = w.f + ma_al(n * 4);
    w.list = w.state + ma_al(n);
    w.bsum = w.list + ma_al(n * 4);
    w.total = w.bsum + ma_al((n / 256 + 2) * 4);
    w.end = w.total + 256;
    return w;
}
size_t map_ws_bytes(const SlmParams* p) { return ma_ws(p->nxy_padded).end; }

int check_ws(const SlmParams* p, void* ws, size_t bytes) {
    if (!ws) return slm_fail(SLM_E_WORKSPACE, "workspace is NULL");
    if (bytes < map_ws_bytes(p))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes < %zu", bytes, map_ws_bytes(p));
    return SLM_OK;
}

inline dim3 seed_grid(const SlmParams* p) {
    return dim3((p->n_seed_points - p->seeds_chunk_offset + kBlock - 1) / kBlock);
}
inline dim3 pixel_grid(const SlmParams* p) { return dim3((p->nxy_padded + kBlock - 1) / kBlock); }
inline uint32_t n_work(const SlmParams* p) { return p->n_seed_points - p->seeds_chunk_offset; }


inline void zero_u64(u64* a, size_t n, hipStream_t s) {
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_zero_u64, dim3((uint32_t)std::max<size_t>(blocks, 1)), dim3(kBlock), 0, s, a, n);
}

// Zhang-Suen thinning of the binary image a (b: scratch of the same size) to
// convergence, kZsK iterations per launch (k_zs_tiles); one host
// synchronisation per launch.  The iteration count is the global loop's: the
// iterations that deleted a pixel plus the first one that deleted none.
int zs_thin(uint8_t* a, uint8_t* b, uint32_t nx, uint32_t ny, uint32_t* ctr, hipStream_t s, uint32_t* iters_out) {
    const uint32_t n_tiles = ((nx + kZsT - 1) / kZsT) * ((ny + kZsT - 1) / kZsT);
    const uint32_t max_iters = nx + ny + 2;
    uint8_t* src = a;
    uint8_t* dst = b;
    uint32_t iters = 0;
    for (uint32_t done = 0;; done += kZsK) {
        zero_u64((u64*)ctr, 1, s);
        hipLaunchKernelGGL(k_zs_tiles, dim3(n_tiles), dim3(256), 0, s, nx, ny, (const uint8_t*)src, dst, ctr);
        std::swap(src, dst);
        uint32_t mask = 0;
        HIP_TRY(hipMemcpyAsync(&mask, ctr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        // the deleting iterations of a launch are a prefix (an iteration that
        // deletes nothing leaves nothing to delete after it)
        if (mask != (1u << kZsK) - 1u) {
            iters = done + (mask ? 32u - (uint32_t)__builtin_clz(mask) : 0u) + 1u;
            break;
        }
        if (done + kZsK >= max_iters)
            return slm_fail(SLM_E_HIP, "thinning did not converge in %u iterations", done + kZsK);
    }
    if (src != a) HIP_TRY(hipMemcpyAsync(a, src, (size_t)nx * ny, hipMemcpyDeviceToDevice, s));
    if (iters_out) *iters_out = iters;
    return SLM_OK;
}

inline u64* ws_a(void* ws) { return (u64*)ws; }
inline u64* ws_b(void* ws, const SlmParams* p) { return (u64*)ws + p->nxy_padded; }
inline uint32_t* ws_ctr(void* ws, const SlmParams* p) { return (uint32_t*)((u64*)ws + 2 * (size_t)p->nxy_padded); }

}  // namespace

extern "C" size_t slm_map_workspace_size(const SlmParams* p) {
    if (slm_check_params(p)) return 0;
    return map_ws_bytes(p);
}

#define SEED_ARGS(p) (*p), (const float2*)seeds_xy
#define LAUNCH_C(kern, grid, ...)                                                           \
    do {                                                                                    \
        if (p->fp_contract)                                                                 \
            hipLaunchKernelGGL(kern<true>, grid, dim3(kBlock), 0, s, __VA_ARGS__);          \
        else                                                                                \
            hipLaunchKernelGGL(kern<false>, grid, dim3(kBlock), 0, s, __VA_ARGS__);         \
    } while (0)

extern "C" int slm_connect_channels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                    const float* uv, uint32_t* mapping, void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !uv || !mapping) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    LAUNCH_C(k_connect, seed_grid(p), SEED_ARGS(p), mask, (const float2*)uv, mapping);
    return slm_finish(stream, flags);
}

namespace {
// scikit-image's medial_axis keep table: bit (index) set when the centre of
// the 3x3 pattern `index` (bit k = row k / 3, column k % 3; centre bit 4) is
// set and either removing it changes the number of 8-connected components of
// the pattern, or fewer than 3 of its pixels are set.
int pattern_components(uint32_t idx) {
    int lab[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, n = 0;
    for (int k = 0; k < 9; k++) {
        if (!(idx >> k & 1u) || lab[k]) continue;
        n++;
        int stack[9], top = 0;
        stack[top++] = k;
        lab[k] = n;
        while (top) {
            const int c = stack[--top], r = c / 3, q = c % 3;
            for (int j = 0; j < 9; j++)
                if ((idx >> j & 1u) && !lab[j] && std::abs(j / 3 - r) <= 1 && std::abs(j % 3 - q) <= 1) {
                    lab[j] = n;
                    stack[top++] = j;
                }
        }
    }
    return n;
}

MaTable ma_keep_table() {
    MaTable t;
    for (int w = 0; w < 16; w++) t.w[w] = 0;
    for (uint32_t idx = 0; idx < 512; idx++) {
        if (!(idx & 16u)) continue;
        const bool keep = pattern_components(idx) != pattern_components(idx & ~16u) || __builtin_popcount(idx) < 3;
        if (keep) t.w[idx >> 5] |= 1u << (idx & 31u);
    }
    return t;
}

// The medial axis of the byte image a, in place (see k_ma_round).  tie: the
// host's tie-breaker permutation of the n_set set pixels.
int medial_axis(uint8_t* a, uint32_t nx, uint32_t ny, const uint32_t* tie, uint32_t n_set, uint8_t* ws,
                uint32_t* pending, hipStream_t s) {
    const size_t n = (size_t)nx * ny;
    const MaWs w = ma_ws(n);
    u64* key = (u64*)(ws + w.key);
    uint32_t* f = (uint32_t*)(ws + w.f);
    uint8_t* state = ws + w.state;
    uint32_t* list = (uint32_t*)(ws + w.list);
    uint32_t* bsum = (uint32_t*)(ws + w.bsum);
    uint32_t* total = (uint32_t*)(ws + w.total);
    const uint32_t nb = (uint32_t)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_ma_cols, dim3((ny + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (const uint8_t*)a, nx, ny, f);
    hipLaunchKernelGGL(k_ma_d2, dim3(nb), dim3(kBlock), 0, s, (const uint8_t*)a, (const uint32_t*)f, nx, ny, key);
    hipLaunchKernelGGL(k_ma_bcount, dim3(nb), dim3(kBlock), 0, s, (const uint8_t*)a, (uint32_t)n, bsum);
    hipLaunchKernelGGL(k_ma_bscan, dim3(1), dim3(1024), 0, s, bsum, nb, total);
    uint32_t count = 0;
    HIP_TRY(hipMemcpyAsync(&count, total, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (count != n_set)
        return slm_fail(SLM_E_INVALID_ARG, "the channel image has %u set pixels, the tie-breaker %u", count, n_set);
    if (!count) return SLM_OK;
    hipLaunchKernelGGL(k_ma_keys, dim3(nb), dim3(kBlock), 0, s, (const uint8_t*)a, nx, ny, (const uint32_t*)bsum,
                       tie, n_set, key, list, state);
    static const MaTable table = ma_keep_table();
    // rounds, 8 per host check; each round decides at least the pixel first
    // in the order, so n_set rounds always suffice
    const uint32_t rb = (n_set + kBlock - 1) / kBlock;
    for (uint32_t done = 0;; done += 8) {
        for (int r = 0; r < 8; r++) {
            zero_u64((u64*)pending, 1, s);
            hipLaunchKernelGGL(k_ma_round, dim3(rb), dim3(kBlock), 0, s, nx, ny, (const uint32_t*)list, n_set,
                               (const u64*)key, state, table, pending);
        }
        uint32_t left = 0;
        HIP_TRY(hipMemcpyAsync(&left, pending, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (!left) break;
        if (done > n_set) return slm_fail(SLM_E_HIP, "medial axis rounds did not finish");
    }
    hipLaunchKernelGGL(k_ma_final, dim3(nb), dim3(kBlock), 0, s, (uint32_t)n, (const uint8_t*)state, a);
    return SLM_OK;
}
}  // namespace

extern "C" int slm_thin_channels(const SlmParams* p, uint32_t* mapping, const uint32_t* tiebreak, uint32_t n_set,
                                 void* workspace, size_t workspace_bytes, uint32_t* n_iter, void* stream,
                                 uint32_t flags) {
    int rc = check_grid(p);
    if (rc) return rc;
    if ((rc = check_ws(p, workspace, workspace_bytes))) return rc;
    if (!mapping || (n_set && !tiebreak)) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t nxy = p->nxy_padded;
    uint8_t* a = (uint8_t*)ws_a(workspace);
    uint8_t* b = a + nxy;
    uint32_t* ctr = ws_ctr(workspace, p);
    hipLaunchKernelGGL(k_zs_init, pixel_grid(p), dim3(kBlock), 0, s, nxy, (const uint32_t*)mapping, a);
    if (tiebreak) {
        if ((rc = medial_axis(a, p->nx_padded, p->ny_padded, tiebreak, n_set, (uint8_t*)workspace, ctr, s)))
            return rc;
    }
    uint32_t iters = 0;
    if ((rc = zs_thin(a, b, p->nx_padded, p->ny_padded, ctr, s, &iters))) return rc;
    hipLaunchKernelGGL(k_zs_final, pixel_grid(p), dim3(kBlock), 0, s, nxy, (const uint8_t*)a, mapping);
    if (n_iter) *n_iter = iters;
    return slm_finish(stream, flags);
}

extern "C" int slm_map_channel_heads(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                     const float* uv, uint32_t* mapping, void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !uv || !mapping) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    if (p->fp_contract)
        hipLaunchKernelGGL(k_heads<true>, persist_grid(k_heads<true>, n_work(p)), dim3(kBlock), 0, s, SEED_ARGS(p),
                           mask, (const float2*)uv, mapping);
    else
        hipLaunchKernelGGL(k_heads<false>, persist_grid(k_heads<false>, n_work(p)), dim3(kBlock), 0, s, SEED_ARGS(p),
                           mask, (const float2*)uv, mapping);
    return slm_finish(stream, flags);
}

extern "C" int slm_count_downchannels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                      const float* uv, uint32_t* mapping, uint32_t* count, uint32_t* link,
                                      void* workspace, size_t workspace_bytes, void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if ((rc = check_ws(p, workspace, workspace_bytes))) return rc;
    if (!mask || !uv || !mapping || !count || !link) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    u64* keyA = ws_a(workspace);
    u64* keyL = ws_b(workspace, p);
    zero_u64(keyA, (size_t)2 * p->nxy_padded, s);
    const dim3 g = seed_grid(p);
#define SLM_LAUNCH_COUNT(CC, RR)                                                                          \
    do {                                                                                                  \
        hipLaunchKernelGGL((k_count<CC, 0, RR>), g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, (const float2*)uv, \
                           mapping, count, keyA, keyL);                                                   \
        hipLaunchKernelGGL((k_count<CC, 1, RR>), g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, (const float2*)uv, \
                           mapping, count, keyA, keyL);                                                   \
    } while (0)
    const bool r2 = slm_res_is_pow2(p);
    if (p->fp_contract) {
        if (r2) SLM_LAUNCH_COUNT(true, true);
        else SLM_LAUNCH_COUNT(true, false);
    } else {
        if (r2) SLM_LAUNCH_COUNT(false, true);
        else SLM_LAUNCH_COUNT(false, false);
    }
#undef SLM_LAUNCH_COUNT
    hipLaunchKernelGGL(k_link_final, pixel_grid(p), dim3(kBlock), 0, s, p->nxy_padded, (const u64*)keyL, link);
    return slm_finish(stream, flags);
}

extern "C" int slm_flag_downchannels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                     uint32_t* mapping, uint32_t* count, const uint32_t* link, void* stream,
                                     uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !mapping || !count || !link) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    const dim3 g = seed_grid(p);
    hipLaunchKernelGGL(k_flag_init, g, dim3(kBlock), 0, s, SEED_ARGS(p), mapping, count);
    hipLaunchKernelGGL(k_flag_walk<2>, g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, mapping, count, link);
    hipLaunchKernelGGL(k_flag_walk<3>, g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, mapping, count, link);
    return slm_finish(stream, flags);
}

extern "C" int slm_label_confluences(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                     const float* slt, uint32_t* mapping, const uint32_t* count,
                                     const uint32_t* link, void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !slt || !mapping || !count || !link) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_confluences, seed_grid(p), dim3(kBlock), 0, s, SEED_ARGS(p), mask, slt, mapping, count,
                       link);
    return slm_finish(stream, flags);
}

extern "C" int slm_segment_downchannels(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                        uint32_t* mapping, const uint32_t* count, const uint32_t* link,
                                        uint32_t* label, void* workspace, size_t workspace_bytes, void* stream,
                                        uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if ((rc = check_ws(p, workspace, workspace_bytes))) return rc;
    if (!mask || !mapping || !count || !link || !label) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    u64* own = ws_a(workspace);
    uint8_t* mark = (uint8_t*)ws_b(workspace, p);
    zero_u64(own, (size_t)p->nxy_padded + ((size_t)p->nxy_padded + 7) / 8, s);
    const dim3 g = seed_grid(p);
    hipLaunchKernelGGL(k_seg_heads, g, dim3(kBlock), 0, s, SEED_ARGS(p), mapping, mark);
    hipLaunchKernelGGL(k_seg_own, g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, (const uint32_t*)mapping, count, link,
                       own);
    hipLaunchKernelGGL(k_seg_walk<2>, g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, (const uint32_t*)mapping, count,
                       link, (const u64*)own, mark, label);
    hipLaunchKernelGGL(k_seg_walk<3>, g, dim3(kBlock), 0, s, SEED_ARGS(p), mask, (const uint32_t*)mapping, count,
                       link, (const u64*)own, mark, label);
    hipLaunchKernelGGL(k_seg_final, pixel_grid(p), dim3(kBlock), 0, s, p->nxy_padded, (const uint8_t*)mark,
                       mapping);
    return slm_finish(stream, flags);
}

extern "C" int slm_link_hillslopes(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                   const float* uv, uint32_t* link, void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !uv || !link) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    const bool r2 = slm_res_is_pow2(p);
    if (p->fp_contract) {
        if (r2) hipLaunchKernelGGL((k_link_hs<true, true>), seed_grid(p), dim3(kBlock), 0, s, SEED_ARGS(p), mask,
                                   (const float2*)uv, link);
        else hipLaunchKernelGGL((k_link_hs<true, false>), seed_grid(p), dim3(kBlock), 0, s, SEED_ARGS(p), mask,
                                (const float2*)uv, link);
    } else {
        if (r2) hipLaunchKernelGGL((k_link_hs<false, true>), seed_grid(p), dim3(kBlock), 0, s, SEED_ARGS(p), mask,
                                   (const float2*)uv, link);
        else hipLaunchKernelGGL((k_link_hs<false, false>), seed_grid(p), dim3(kBlock), 0, s, SEED_ARGS(p), mask,
                                (const float2*)uv, link);
    }
    return slm_finish(stream, flags);
}

namespace {
template <bool FLANKS>
int seg_hs_launch(const SlmParams* p, const float* seeds_xy, const uint8_t* mask, const float* uv,
                  const uint32_t* mapping, uint32_t* label, void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !uv || !mapping || !label) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
#define SLM_LAUNCH_SEG(CC, RR)                                                                              \
    hipLaunchKernelGGL((k_seg_hs<CC, RR, FLANKS>), persist_grid(k_seg_hs<CC, RR, FLANKS>, n_work(p)), dim3(kBlock), 0, \
                       s, SEED_ARGS(p), mask, (const float2*)uv, mapping, label)
    const bool r2 = slm_res_is_pow2(p);
    if (p->fp_contract) {
        if (r2) SLM_LAUNCH_SEG(true, true);
        else SLM_LAUNCH_SEG(true, false);
    } else {
        if (r2) SLM_LAUNCH_SEG(false, true);
        else SLM_LAUNCH_SEG(false, false);
    }
#undef SLM_LAUNCH_SEG
    return slm_finish(stream, flags);
}
}  // namespace

extern "C" int slm_segment_hillslopes(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                      const float* uv, const uint32_t* mapping, uint32_t* label, void* stream,
                                      uint32_t flags) {
    return seg_hs_launch<false>(p, seeds_xy, mask, uv, mapping, label, stream, flags);
}

extern "C" int slm_segment_hillslopes_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                             const float* uv, const uint32_t* mapping, uint32_t* label,
                                             void* stream, uint32_t flags) {
    return seg_hs_launch<true>(p, seeds_xy, mask, uv, mapping, label, stream, flags);
}

extern "C" int slm_subsegment_channel_edges(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                            uint32_t* mapping, const uint32_t* link, uint32_t* label,
                                            void* stream, uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !mapping || !link || !label) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sub_edges, seed_grid(p), dim3(kBlock), 0, s, SEED_ARGS(p), mask, mapping, link, label);
    return slm_finish(stream, flags);
}

extern "C" int slm_subsegment_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                     const float* uv, const uint32_t* mapping, uint32_t* label, void* stream,
                                     uint32_t flags) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if (!mask || !uv || !mapping || !label) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
#define SLM_LAUNCH_SUB(CC, RR)                                                                                 \
    hipLaunchKernelGGL((k_sub_flanks<CC, RR>), persist_grid(k_sub_flanks<CC, RR>, n_work(p)), dim3(kBlock), 0, s, SEED_ARGS(p), \
                       mask, (const float2*)uv, mapping, label)
    const bool r2 = slm_res_is_pow2(p);
    if (p->fp_contract) {
        if (r2) SLM_LAUNCH_SUB(true, true);
        else SLM_LAUNCH_SUB(true, false);
    } else {
        if (r2) SLM_LAUNCH_SUB(false, true);
        else SLM_LAUNCH_SUB(false, false);
    }
#undef SLM_LAUNCH_SUB
    return slm_finish(stream, flags);
}

static int fix_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask, const float* uv,
                      const uint32_t* mapping, uint32_t* label, void* workspace, size_t workspace_bytes,
                      void* stream, uint32_t flags, bool left) {
    int rc = check_map(p, seeds_xy);
    if (rc) return rc;
    if (!n_work(p)) return slm_finish(stream, flags);
    if ((rc = check_ws(p, workspace, workspace_bytes))) return rc;
    if (!mask || !uv || !mapping || !label) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    uint32_t* snap = (uint32_t*)ws_a(workspace);
    HIP_TRY(hipMemcpyAsync(snap, label, (size_t)4 * p->nxy_padded, hipMemcpyDeviceToDevice, s));
    const float2* sd = (const float2*)seeds_xy;
    const float2* u = (const float2*)uv;
#define SLM_LAUNCH_FIX(CC, LL, RR)                                                                          \
    hipLaunchKernelGGL((k_fix<CC, LL, RR>), persist_grid(k_fix<CC, LL, RR>, n_work(p)), dim3(kBlock), 0, s, *p, sd, mask, u, \
                       mapping, snap, label)
    const bool r2 = slm_res_is_pow2(p);
    if (p->fp_contract) {
        if (left) {
            if (r2) SLM_LAUNCH_FIX(true, true, true);
            else SLM_LAUNCH_FIX(true, true, false);
        } else {
            if (r2) SLM_LAUNCH_FIX(true, false, true);
            else SLM_LAUNCH_FIX(true, false, false);
        }
    } else {
        if (left) {
            if (r2) SLM_LAUNCH_FIX(false, true, true);
            else SLM_LAUNCH_FIX(false, true, false);
        } else {
            if (r2) SLM_LAUNCH_FIX(false, false, true);
            else SLM_LAUNCH_FIX(false, false, false);
        }
    }
#undef SLM_LAUNCH_FIX
    return slm_finish(stream, flags);
}

extern "C" int slm_fix_right_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                    const float* uv, const uint32_t* mapping, uint32_t* label, void* workspace,
                                    size_t workspace_bytes, void* stream, uint32_t flags) {
    return fix_flanks(p, seeds_xy, mask, uv, mapping, label, workspace, workspace_bytes, stream, flags, false);
}

extern "C" int slm_fix_left_flanks(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                   const float* uv, const uint32_t* mapping, uint32_t* label, void* workspace,
                                   size_t workspace_bytes, void* stream, uint32_t flags) {
    return fix_flanks(p, seeds_xy, mask, uv, mapping, label, workspace, workspace_bytes, stream, flags, true);
}

// ---------------------------------------------------------- mid-slopes and ridges

namespace {
inline size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }
}  // namespace

// bytes of one bit-packed grid (64-pixel words along y)
static size_t bits_bytes(uint32_t nx, uint32_t ny) { return (size_t)nx * ((ny + 63u) / 64u) * 8u; }

extern "C" size_t slm_slopes_workspace_size(uint32_t nx, uint32_t ny) {
    const size_t n = (size_t)nx * ny;
    const size_t nb = std::max(al256(n), al256(bits_bytes(nx, ny)));
    return 4 * al256(n * 4) + 4 * al256(n) + 2 * nb + 512;
}

extern "C" int slm_gaussian_filter(const float* in, float* out, float* tmp, uint32_t nx, uint32_t ny,
                                   const double* weights, int radius, void* stream, uint32_t flags) {
    if (!in || !out || !tmp || !weights || !nx || !ny || radius < 0 || (uint64_t)nx * ny >= 0x80000000ull)
        return slm_fail(SLM_E_INVALID_ARG, "bad gaussian_filter arguments");
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((uint32_t)(((size_t)nx * ny + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_gauss<0>, grid, dim3(kBlock), 0, s, in, tmp, nx, ny, weights, radius);
    hipLaunchKernelGGL(k_gauss<1>, grid, dim3(kBlock), 0, s, (const float*)tmp, out, nx, ny, weights, radius);
    return slm_finish(stream, flags);
}

extern "C" int slm_map_midslopes_ridges(const float* sla, const uint8_t* mask, uint32_t nx, uint32_t ny,
                                        const double* w_mid, int r_mid, float mid_threshold, const double* w_ridge,
                                        int r_ridge, float ridge_level, uint32_t midslope_flag, uint32_t ridge_flag,
                                        uint32_t* mapping, void* workspace, size_t workspace_bytes, void* stream,
                                        uint32_t flags) {
    if (!sla || !mask || !mapping || !w_mid || !w_ridge || !workspace || !nx || !ny || r_mid < 0 || r_ridge < 0 ||
        (uint64_t)nx * ny >= 0x80000000ull)
        return slm_fail(SLM_E_INVALID_ARG, "bad map_midslopes_ridges arguments");
    if (workspace_bytes < slm_slopes_workspace_size(nx, ny))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes, need %zu", workspace_bytes,
                        slm_slopes_workspace_size(nx, ny));
    const size_t n = (size_t)nx * ny;
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    float* ang = (float*)w;
    float* tmp = (float*)(w + al256(n * 4));
    float* gm = (float*)(w + 2 * al256(n * 4));
    float* gr = (float*)(w + 3 * al256(n * 4));
    uint8_t* mid = w + 4 * al256(n * 4);
    uint8_t* ridge = mid + al256(n);
    uint8_t* fat = ridge + al256(n);
    uint8_t* scr = fat + al256(n);
    const size_t nb = std::max(al256(n), al256(bits_bytes(nx, ny)));
    uint8_t* ra = scr + al256(n);
    uint8_t* rb = ra + nb;
    uint32_t* ctr = (uint32_t*)(rb + nb);
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((uint32_t)((n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_sla_angle, grid, dim3(kBlock), 0, s, (const float2*)sla, (uint32_t)n, ang);
    hipLaunchKernelGGL(k_gauss<0>, grid, dim3(kBlock), 0, s, (const float*)ang, tmp, nx, ny, w_mid, r_mid);
    hipLaunchKernelGGL(k_gauss<1>, grid, dim3(kBlock), 0, s, (const float*)tmp, gm, nx, ny, w_mid, r_mid);
    hipLaunchKernelGGL(k_gauss<0>, grid, dim3(kBlock), 0, s, (const float*)ang, tmp, nx, ny, w_ridge, r_ridge);
    hipLaunchKernelGGL(k_gauss<1>, grid, dim3(kBlock), 0, s, (const float*)tmp, gr, nx, ny, w_ridge, r_ridge);
    hipLaunchKernelGGL(k_slope_marks, grid, dim3(kBlock), 0, s, (const float*)gm, (const float*)gr, mask,
                       (uint32_t)n, mid_threshold, ridge_level, mid, ridge);
    // ridges: dilate twice, fill holes
    hipLaunchKernelGGL(k_dilate8, grid, dim3(kBlock), 0, s, (const uint8_t*)ridge, scr, nx, ny);
    hipLaunchKernelGGL(k_dilate8, grid, dim3(kBlock), 0, s, (const uint8_t*)scr, fat, nx, ny);
    {
        const uint32_t nwy = (ny + 63u) / 64u;
        const uint32_t n_tiles = ((nx + 63u) / 64u) * nwy;
        const uint32_t n_words = nx * nwy;
        u64* open = (u64*)ra;
        u64* reach = (u64*)rb;
        hipLaunchKernelGGL(k_pack_open, dim3((n_words + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (const uint8_t*)fat,
                           nx, ny, nwy, open, reach);
        const uint32_t max_launches = 4u * n_tiles + 16u;
        for (uint32_t launches = 0;; launches++) {
            if (launches > max_launches)
                return slm_fail(SLM_E_HIP, "hole filling did not converge in %u launches", launches);
            zero_u64((u64*)ctr, 1, s);
            hipLaunchKernelGGL(k_fill_tiles, dim3((n_tiles + 3u) / 4u), dim3(256), 0, s, (const u64*)open, reach, nx,
                               ny, nwy, n_tiles, ctr);
            uint32_t grown = 0;
            HIP_TRY(hipMemcpyAsync(&grown, ctr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (!grown) break;
        }
        hipLaunchKernelGGL(k_fill_final, grid, dim3(kBlock), 0, s, fat, (const u64*)reach, nx, ny, nwy);
    }
    // skeletons
    int rc;
    if ((rc = zs_thin(mid, scr, nx, ny, ctr, s, nullptr))) return rc;
    if ((rc = zs_thin(fat, scr, nx, ny, ctr, s, nullptr))) return rc;
    hipLaunchKernelGGL(k_flag_marks, grid, dim3(kBlock), 0, s, (const uint8_t*)mid, (uint32_t)n, midslope_flag,
                       mapping);
    hipLaunchKernelGGL(k_flag_marks, grid, dim3(kBlock), 0, s, (const uint8_t*)fat, (uint32_t)n, ridge_flag, mapping);
    return slm_finish(stream, flags);
}
