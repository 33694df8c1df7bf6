// slm_hip.hip -- gfx950 kernels and the extern "C" entry points of libslm_hip.so.
//
// Entry points are declared in include/slm_hip.h.  Kernels:
//   k_rng_stream      Lehmer seeding / draws            (rng.cl, integratefields.cl:93-99)
//   k_tile_*          counting sort of seeds into Morton-ordered 8x8 pixel tiles
//   k_fields          persistent lane-refill tracer for integrate_fields, run-coalesced
//                     packed 64-bit accumulation checked on device (k_fields_sumlow/commit)
//                     (integratefields.cl:60-114, jittertrajectory.cl:42-88)
//   k_paths           one lane per seed for integrate_trajectory
//                     (integratetraj.cl:54-97, trajectory.cl:41-80)
//   k_fields_record   fields streamlines with positions recorded (test hook)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <stdarg.h>
#include <algorithm>
#include <mutex>
#include "../../include/slm_hip.h"
#include "trace_common.hpp"
#include "common.hpp"

using namespace slm;

// ------------------------------------------------------------------ errors

static thread_local char g_err[512];

int slm_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int slm_finish(void* stream, uint32_t flags) {
    HIP_TRY(hipGetLastError());
    if (flags & SLM_FLAG_SYNC) HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return SLM_OK;
}

extern "C" uint32_t slm_abi_version(void) { return SLM_ABI_VERSION; }
extern "C" const char* slm_last_error(void) { return g_err; }
extern "C" int slm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int slm_check_params(const SlmParams* p) {
    if (!p) return slm_fail(SLM_E_INVALID_ARG, "params is NULL");
    if (p->abi_version != SLM_ABI_VERSION)
        return slm_fail(SLM_E_ABI, "SlmParams.abi_version %u != %u", p->abi_version, SLM_ABI_VERSION);
    if (p->nx_padded == 0 || p->ny_padded == 0 ||
        (uint64_t)p->nx_padded * p->ny_padded != (uint64_t)p->nxy_padded)
        return slm_fail(SLM_E_INVALID_ARG, "grid %u x %u != nxy_padded %u", p->nx_padded, p->ny_padded,
                    p->nxy_padded);
    if (p->nx_padded > (1u << 20) || p->ny_padded > (1u << 20))
        return slm_fail(SLM_E_INVALID_ARG, "grid side above 2^20");
    if (!(p->trajectory_resolution > 0.0f))
        return slm_fail(SLM_E_INVALID_ARG, "trajectory_resolution must be > 0");
    if (p->seeds_chunk_offset > p->n_seed_points)
        return slm_fail(SLM_E_INVALID_ARG, "seeds_chunk_offset %u > n_seed_points %u", p->seeds_chunk_offset,
                    p->n_seed_points);
    return SLM_OK;
}
static inline int check_params(const SlmParams* p) { return slm_check_params(p); }


struct DevInfo {
    int cus = 0;
    int lds_per_block = 0;
};
static std::mutex g_dev_mu;
static DevInfo g_dev[64];

static int device_info(DevInfo& out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return slm_fail(SLM_E_HIP, "device id %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_dev[dev].cus == 0) {
        int cus = 0;
        HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        g_dev[dev].cus = cus > 0 ? cus : 1;
        int lds = 0;
        HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        g_dev[dev].lds_per_block = lds > 0 ? lds : 65536;
    }
    out = g_dev[dev];
    return SLM_OK;
}

// ------------------------------------------------------------------ utility kernels

__global__ void k_fill_u32(uint32_t* __restrict__ a, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = v;
}

static void fill_u32(uint32_t* a, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, a, n, v);
}

// ------------------------------------------------------------------ RNG stream

__global__ void k_rng_stream(uint32_t idx0, uint32_t n, uint32_t n_draws, uint32_t* __restrict__ rng0,
                             uint32_t* __restrict__ draws) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = rng_seed(idx0 + i);
    if (rng0) rng0[i] = s;
    if (draws)
        for (uint32_t k = 0; k < n_draws; k++) draws[(size_t)i * n_draws + k] = lehmer(s);
}

extern "C" int slm_rng_stream(uint32_t idx0, uint32_t n, uint32_t n_draws, uint32_t* rng0_out,
                              uint32_t* draws, void* stream, uint32_t flags) {
    if (n == 0) return slm_finish(stream, flags);
    if (!rng0_out && !draws) return slm_fail(SLM_E_INVALID_ARG, "no output buffer");
    if (draws && n_draws == 0) return slm_fail(SLM_E_INVALID_ARG, "n_draws is 0 with a draws buffer");
    hipLaunchKernelGGL(k_rng_stream, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, idx0, n,
                       n_draws, rng0_out, draws);
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ rounding check
//
// The short sqrt / divide forms of trace_common.hpp against the operators,
// bit for bit.  counts (uint64 [8], accumulated):
//   [0]/[1] sqrt_short(d) != sqrtf(d), of every float d in [2^-96, 2^100]
//   [2]/[3] recip_short(b) != 1/b, of every b in [2^-48, 2^50]
//   [4]/[5] div_short(a, b, recip_short(b)) != a/b: n_num numerators per b with
//           |a| in [2^-70, 8b], either sign (hashed exponent and mantissa), and
//           every significand of a in [1, 2) for 2048 significands of b in
//           [1, 2) (the 1024 largest, where a*RN(1/b) is furthest from a/b,
//           and 1024 hashed; scaling both by powers of two changes no bit)
//   [6]     sqrt_cr(d) / inv_sqrt_cr(d) differ from the operators, d = any bits
//   [7]     div_by_length(x, y) differs from x/len, y/len, x = any bits, y hashed

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void k_check_rounding(unsigned long long* __restrict__ counts, uint32_t n_num) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        const uint32_t bits = (uint32_t)i;
        const float d = __uint_as_float(bits);
        if (sqrt_short_ok(d)) {
            c[1]++;
            if (__float_as_uint(sqrt_short(d)) != __float_as_uint(fsqrt(d))) c[0]++;
        }
        if (d >= 0x1p-48f && d <= 0x1p50f) {
            c[3]++;
            const float r = recip_short(d);
            if (__float_as_uint(r) != __float_as_uint(fdiv(1.0f, d))) c[2]++;
            const uint32_t eb = (bits >> 23) & 0xffu;
            const uint32_t lo = 127u - 70u, span = eb + 3u - lo + 1u;
            for (uint32_t j = 0; j < n_num; j++) {
                const uint32_t h = mix32(bits * 0x9e3779b9u + j * 0x85ebca6bu + 1u);
                uint32_t nb = ((lo + h % span) << 23) | (mix32(h) & 0x7fffffu);
                if (__uint_as_float(nb) > 8.0f * d) nb = __float_as_uint(8.0f * d);
                nb |= h & 0x80000000u;
                const float a = __uint_as_float(nb);
                c[5]++;
                if (__float_as_uint(div_short(a, d, r)) != __float_as_uint(fdiv(a, d))) c[4]++;
            }
        }
        if (__float_as_uint(sqrt_cr(d)) != __float_as_uint(fsqrt(d)) ||
            __float_as_uint(inv_sqrt_cr(d)) != __float_as_uint(fdiv(1.0f, fsqrt(d))))
            c[6]++;
        const float y = __uint_as_float(mix32(bits ^ 0x5bd1e995u));
        const float2 q = div_by_length<false>(d, y);
        const float len = fsqrt(d * d + y * y);
        if (__float_as_uint(q.x) != __float_as_uint(fdiv(d, len)) || __float_as_uint(q.y) != __float_as_uint(fdiv(y, len)))
            c[7]++;
    }
    for (int k = 0; k < 8; k++)
        if (c[k]) atomicAdd(&counts[k], c[k]);
}

// Every dividend significand against one divisor significand per blockIdx.y.
__global__ void k_check_quotients(unsigned long long* __restrict__ counts) {
    const uint32_t j = blockIdx.y;
    const uint32_t bm = j < 1024u ? 0x7fffffu - j : (mix32(j * 0x9e3779b9u + 11u) & 0x7fffffu);
    const float b = __uint_as_float((127u << 23) | bm);
    const float y = recip_short(b);
    unsigned long long bad = 0, n = 0;
    for (uint32_t am = blockIdx.x * blockDim.x + threadIdx.x; am < (1u << 23); am += gridDim.x * blockDim.x) {
        const float a = __uint_as_float((127u << 23) | am);
        n++;
        if (__float_as_uint(div_short(a, b, y)) != __float_as_uint(fdiv(a, b))) bad++;
    }
    if (bad) atomicAdd(&counts[4], bad);
    atomicAdd(&counts[5], n);
}

extern "C" int slm_check_rounding(uint64_t* counts, uint32_t n_numerators, void* stream, uint32_t flags) {
    if (!counts) return slm_fail(SLM_E_INVALID_ARG, "counts is NULL");
    hipLaunchKernelGGL(k_check_rounding, dim3(8192), dim3(256), 0, (hipStream_t)stream,
                       (unsigned long long*)counts, n_numerators);
    hipLaunchKernelGGL(k_check_quotients, dim3(256, 2048), dim3(256), 0, (hipStream_t)stream,
                       (unsigned long long*)counts);
    return slm_finish(stream, flags);
}

// VALU issue-rate calibration: each lane runs n_iter rounds of 16 independent
// chains of one instruction kind, so each wave issues exactly 16*n_iter of it
// (plus a few dozen set-up and reduction instructions).  Timing it, and
// reading the VALU instruction counter over it, gives the chip's sustained
// wave64 issue rate of that instruction and what the counter counts per
// issued instruction.  Kinds: 0 v_fma_f32, 1 v_add_f32, 2 v_add_u32,
// 3 v_xor_b32, 4 v_mul_f32, 5 v_fma_f32 and v_add_u32 alternating,
// 6 v_cndmask_b32, 7 v_cmp_lt_f32 (to vcc), 8 v_mov_b32.
template <int KIND>
__global__ __launch_bounds__(256) void k_calib_valu(float* __restrict__ out, uint32_t n_iter, float a, float b) {
    float x[16];
    uint32_t u[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        x[i] = (float)(threadIdx.x + 16u * i);
        u[i] = threadIdx.x * 2654435761u + (uint32_t)i;
    }
    const uint32_t ua = __float_as_uint(a), ub = __float_as_uint(b);
    for (uint32_t k = 0; k < n_iter; ++k) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if constexpr (KIND == 0) x[i] = __builtin_fmaf(x[i], a, b);
            else if constexpr (KIND == 1) x[i] = x[i] + b;
            else if constexpr (KIND == 2) u[i] = u[i] + ua;
            else if constexpr (KIND == 3) u[i] = u[i] ^ (ub + (uint32_t)i);
            else if constexpr (KIND == 4) x[i] = x[i] * a;
            else if constexpr (KIND == 5) {
                // eight float and eight integer chains, interleaved
                if (i & 1) u[i] = u[i] + ua;
                else x[i] = __builtin_fmaf(x[i], a, b);
            } else if constexpr (KIND == 6) {
                __asm__ volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(u[i]) : "v"(ub));
            } else if constexpr (KIND == 7) {
                __asm__ volatile("v_cmp_lt_f32 vcc, %0, %1" : : "v"(x[i]), "v"(x[(i + 1) & 15]) : "vcc");
            } else {
                __asm__ volatile("v_mov_b32 %0, %1" : "=v"(u[i]) : "v"(u[(i + 1) & 15]));
            }
        }
        if constexpr (KIND == 2 || KIND == 3 || KIND == 5) {
            // integer chains with a loop-invariant operand have closed forms the
            // compiler could use instead of the loop; an empty asm statement that
            // claims to modify the registers keeps every round
            __asm__ volatile("" : "+v"(u[0]), "+v"(u[1]), "+v"(u[2]), "+v"(u[3]), "+v"(u[4]), "+v"(u[5]),
                             "+v"(u[6]), "+v"(u[7]), "+v"(u[8]), "+v"(u[9]), "+v"(u[10]), "+v"(u[11]),
                             "+v"(u[12]), "+v"(u[13]), "+v"(u[14]), "+v"(u[15]));
        }
    }
    float s = 0.0f;
    uint32_t t = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        s += x[i];
        t ^= u[i];
    }
    if (s == 1.25f || t == 0x9e3779b9u) out[blockIdx.x] = s;   // keeps the chains live; practically never true
}

extern "C" int slm_calibrate_valu(uint32_t kind, uint32_t n_blocks, uint32_t n_iter, float* out, void* stream,
                                  uint32_t flags) {
    if (!out || !n_blocks || kind > 8) return slm_fail(SLM_E_INVALID_ARG, "out is NULL, n_blocks is 0 or kind > 8");
    const hipStream_t s = (hipStream_t)stream;
    const dim3 g(n_blocks), blk(256);
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_calib_valu<0>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 1: hipLaunchKernelGGL(k_calib_valu<1>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 2: hipLaunchKernelGGL(k_calib_valu<2>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 3: hipLaunchKernelGGL(k_calib_valu<3>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 4: hipLaunchKernelGGL(k_calib_valu<4>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 5: hipLaunchKernelGGL(k_calib_valu<5>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 6: hipLaunchKernelGGL(k_calib_valu<6>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        case 7: hipLaunchKernelGGL(k_calib_valu<7>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
        default: hipLaunchKernelGGL(k_calib_valu<8>, g, blk, 0, s, out, n_iter, 0.999f, 0.001f); break;
    }
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ seed tiling
//
// Work order for the fields tracer: seeds are bucketed into 8x8-pixel tiles,
// tiles are visited in Morton (Z) order, so the 64 lanes of a wave start in
// one tile and neighbouring waves start in neighbouring tiles.  The RNG stays
// keyed on each seed's index in the caller's (shuffled) list.

static __host__ __device__ inline uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

struct TileGeom {
    uint32_t tx, ty;      // tiles along x and y
    uint32_t side;        // power-of-two side of the Morton square
    uint32_t n_keys;      // side*side
};

static TileGeom tile_geom(const SlmParams* p) {
    TileGeom g;
    g.tx = (p->nx_padded + 7u) / 8u;
    g.ty = (p->ny_padded + 7u) / 8u;
    uint32_t m = g.tx > g.ty ? g.tx : g.ty, side = 1;
    while (side < m) side <<= 1;
    g.side = side;
    g.n_keys = side * side;
    return g;
}

__device__ __forceinline__ uint32_t seed_tile_key(const SlmParams& P, float2 s) {
    const uint32_t px = pix(s.x, P.pad_width_pp5, P.nx_padded - 1u);
    const uint32_t py = pix(s.y, P.pad_width_pp5, P.ny_padded - 1u);
    return spread16(px >> 3) | (spread16(py >> 3) << 1);
}

__global__ void k_tile_count(SlmParams P, const float2* __restrict__ seeds, uint32_t n,
                             uint32_t* __restrict__ cnt) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    atomicAdd(&cnt[seed_tile_key(P, seeds[P.seeds_chunk_offset + k])], 1u);
}

// Exclusive scan of cnt[0..n) in place by one 1024-thread block; writes cursor copy.
__global__ __launch_bounds__(1024) void k_tile_scan(uint32_t* __restrict__ cnt, uint32_t* __restrict__ cur,
                                                    uint32_t n) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = t * per, e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += cnt[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = (t >= off) ? part[t - off] : 0u;
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

// Seed record of the work order: position and RNG key of one seed.
struct SeedRec {
    float x, y;
    uint32_t seed_idx, pad;
};

__global__ void k_tile_scatter(SlmParams P, const float2* __restrict__ seeds, uint32_t n,
                               uint32_t* __restrict__ cur, SeedRec* __restrict__ order) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float2 s = seeds[P.seeds_chunk_offset + k];
    const uint32_t key = seed_tile_key(P, s);
    SeedRec r;
    r.x = s.x;
    r.y = s.y;
    r.seed_idx = P.seeds_chunk_offset + k;
    r.pad = 0u;
    order[atomicAdd(&cur[key], 1u)] = r;
}

// ------------------------------------------------------------------ fields tracer
//
// Accumulation.  Every write of the reference adds (slc += 1, slt += (uint)(l+0.5f))
// at one pixel.  Here a lane keeps a run (pixel, count, length-sum) while its
// streamline stays in one pixel and flushes the run with ONE 64-bit atomic into
// acc[pixel] = (slt << 32) | slc.  Integer sums commute, so the result equals
// the reference's per-write uint32 atomics exactly, provided no pixel's slc
// carries out of the low 32 bits (slt's own carry falls off bit 63, which is
// exactly uint32 wrap-around).  That proviso is checked on the device, with
// no host round trip: sum over pixels of low32(acc) must equal the exact
// number of writes W counted by the kernel.  If it does not, k_fields_commit
// leaves the outputs untouched and a fallback launch re-traces with separate
// 32-bit atomics straight into slc/slt; otherwise the fallback launch exits
// at once.
//
// Workspace layout (bytes, 256-aligned blocks):
//   [0]    FieldsCtrl                      W, sum of low halves, queue heads
//   [4096] SeedRec order[n]                seeds in tile order with their RNG keys
//   [..]   uint32 cnt[n_keys], cur[n_keys] counting-sort buffers
//   [..]   uint64 acc[nxy_padded]          packed (slt << 32) | slc
//   [..]   Quad quad[nxy_padded]           interpolation corners per pixel
static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Work-queue heads: one per XCD slice of the work order, 256 bytes apart.
constexpr uint32_t kQueues = 8;
constexpr uint32_t kQueueStride = 64;  // uint32 words

struct FieldsCtrl {
    unsigned long long writes;
    unsigned long long sumlow;
    uint32_t pad[60];
    uint32_t queue[kQueues * kQueueStride];
};

struct FieldsWs {
    size_t off_order, off_cnt, off_cur, off_acc, off_quad, total;
};

static FieldsWs fields_ws(const SlmParams* p) {
    FieldsWs w;
    const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
    const TileGeom g = tile_geom(p);
    w.off_order = align256(sizeof(FieldsCtrl)) > 4096 ? align256(sizeof(FieldsCtrl)) : 4096;
    w.off_cnt = w.off_order + align256((size_t)n * sizeof(SeedRec));
    w.off_cur = w.off_cnt + align256((size_t)g.n_keys * 4);
    w.off_acc = w.off_cur + align256((size_t)g.n_keys * 4);
    w.off_quad = w.off_acc + align256((size_t)p->nxy_padded * 8);
    w.total = w.off_quad + align256((size_t)p->nxy_padded * sizeof(Quad));
    return w;
}

extern "C" size_t slm_trace_fields_workspace_size(const SlmParams* p) {
    if (check_params(p) != SLM_OK) return 0;
    return fields_ws(p).total;
}

__device__ __forceinline__ bool fields_packed_ok(const FieldsCtrl* ctrl, uint32_t force_fallback) {
    return !force_fallback && ctrl->sumlow == ctrl->writes;
}

// Wave-private accumulation table in LDS.  A flushed run (pixel, packed sum)
// is added into the wave's table instead of global memory; the table is
// direct-mapped by pixel position (x mod 16, y mod 32, so a 16x32 window of
// neighbouring pixels never collides).  A run whose slot holds another pixel
// evicts it: the evicted sum goes to acc with one global 64-bit atomic.  At
// the end the wave drains its table into acc.  Streamlines of a wave start
// in one 8x8 tile and converge downhill, so most flushes hit a pixel another
// lane flushed recently, and the table removes most of the global atomics
// and every same-address collision among the lanes of one atomic
// instruction.  All sums are the same integers added in another order.
//
// Protocol of one add() (every lane of the wave calls it together; lanes
// with has = false take no part).  The wave's LDS operations execute in
// program order, and wavefront-scope fences keep the compiler from moving
// them across phases:
//   1. read the slot's tag; a lane whose pixel is there adds its sum (hit);
//   2. a missing lane swaps the tag from what it read to its own pixel with a
//      compare-and-swap; the first such lane per slot wins;
//   3. the winner exchanges the slot's sum for its own (the old sum, which
//      already holds the hits of phase 1, is the evicted pixel's);
//   4. a loser whose pixel is the winner's adds into the slot; a loser with
//      another pixel, and the winner's evicted sum, go to global memory.
constexpr uint32_t kRunSlots = 512;  // per wave: 4 B tag + 8 B sum each
constexpr uint32_t kRunEmpty = 0xffffffffu;

__device__ __forceinline__ uint32_t slot_of(uint32_t ix, uint32_t iy) {
    return ((ix & 15u) << 5) | (iy & 31u);
}

// Table sizes of the batched tracer (a tuning option): 256 slots map a 16x16
// window of pixels, 512 a 16x32 window, 1024 a 32x32 window, collision-free.
template <uint32_t SLOTS>
__device__ __forceinline__ uint32_t slot_of_t(uint32_t ix, uint32_t iy) {
    static_assert(SLOTS == 256 || SLOTS == 512 || SLOTS == 1024, "table size");
    if constexpr (SLOTS == 256) return ((ix & 15u) << 4) | (iy & 15u);
    else if constexpr (SLOTS == 512) return ((ix & 15u) << 5) | (iy & 31u);
    else return ((ix & 31u) << 5) | (iy & 31u);
}

template <uint32_t SLOTS>
struct RunTableT {
    uint32_t* tag;
    unsigned long long* val;

    __device__ __forceinline__ void init(uint32_t lane) {
        for (uint32_t i = lane; i < SLOTS; i += 64u) {
            __hip_atomic_store(&tag[i], kRunEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            __hip_atomic_store(&val[i], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }

    __device__ __forceinline__ void add(bool has, uint32_t a, uint32_t s, unsigned long long v,
                                        unsigned long long* __restrict__ acc) {
        uint32_t r = kRunEmpty;
        bool miss = false;
        if (has) {
            r = __hip_atomic_load(&tag[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            if (r == a) __hip_atomic_fetch_add(&val[s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            else miss = true;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        bool won = false;
        uint32_t seen = r;
        if (miss)
            won = __hip_atomic_compare_exchange_strong(&tag[s], &seen, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        unsigned long long old = 0;
        if (won) old = __hip_atomic_exchange(&val[s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (miss && !won && seen == a)
            __hip_atomic_fetch_add(&val[s], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (won && r != kRunEmpty && old) atomicAdd(&acc[r], old);
        if (miss && !won && seen != a) atomicAdd(&acc[a], v);
    }

    __device__ __forceinline__ void drain(uint32_t lane, unsigned long long* __restrict__ acc) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (uint32_t i = lane; i < SLOTS; i += 64u) {
            const uint32_t t = __hip_atomic_load(&tag[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const unsigned long long v = __hip_atomic_load(&val[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            if (t != kRunEmpty && v) atomicAdd(&acc[t], v);
        }
    }
};
using RunTable = RunTableT<kRunSlots>;

// One persistent wave pulls streamlines from the queue; a lane that finishes
// takes the next item at once, so divergent lengths do not idle the wave.
// Work item w: seed ordinal order[w / d^2], sub-pixel start w % d^2 = j*d+i.
// PACKED: runs go to acc with one 64-bit atomic; W and the step count are
// recorded.  !PACKED: the fallback, which runs only when the packed result was
// rejected and adds runs to slc/slt with two 32-bit atomics.
// dt_fixed: the host has proven that the adaptive step can never leave DT_MAX
// (slm::dt_is_fixed: every dxy component is c/RES with c in [-128, 127], so
// step_error <= sqrt(8)*(128/RES)/GRID_SCALE, and a GRID_SCALE *
// ADJUSTED_MAX_ERROR above that bound by a rounding margin makes
// fmin(DT_MAX, ADJUSTED_MAX_ERROR*DT_MAX/step_error) == DT_MAX); step_error
// is then not computed.
template <bool C, bool PACKED, bool R2>
__global__ __launch_bounds__(256) void k_fields(SlmParams P, const float2* __restrict__ seeds,
                                                const SeedRec* __restrict__ order, uint32_t n_work,
                                                const uint8_t* __restrict__ mask,
                                                const Quad* __restrict__ quad,
                                                unsigned long long* __restrict__ acc,
                                                uint32_t* __restrict__ slc, uint32_t* __restrict__ slt,
                                                unsigned long long* __restrict__ steps,
                                                FieldsCtrl* __restrict__ ctrl, uint32_t force_fallback,
                                                uint32_t dt_fixed) {
    if constexpr (!PACKED) {
        if (fields_packed_ok(ctrl, force_fallback)) return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t d = P.subpixel_seed_point_density ? P.subpixel_seed_point_density : 1u;
    const uint32_t dd = d * d;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const float res = P.trajectory_resolution;
    const float combo = P.combo_factor;
    const float inv_res = 1.0f / res;
    const uint32_t nxm1 = P.nx_padded - 1u, nym1 = P.ny_padded - 1u;

    __shared__ uint32_t s_tag[PACKED ? 4u * kRunSlots : 1u];
    __shared__ unsigned long long s_val[PACKED ? 4u * kRunSlots : 1u];
    RunTable table{s_tag + (threadIdx.x >> 6) * kRunSlots, s_val + (threadIdx.x >> 6) * kRunSlots};
    if constexpr (PACKED) table.init(lane);

    bool active = false;
    bool exhausted = false;
    float x = 0.f, y = 0.f, l = 0.f, dt = 0.f;
    uint32_t n = 0, rng = 0;
    Quad Q;                                        // quad of the current pixel
    uint32_t cx = 0, cy = 0;                       // and its coordinates
    uint32_t run_idx = 0, run_slot = 0, run_n = 0, run_t = 0;   // current same-pixel run
    uint64_t my_steps = 0, my_writes = 0;

    // adds the runs of the lanes with has = true (all lanes call it together)
    auto emit = [&](bool has, uint32_t a, uint32_t s, uint32_t cn, uint32_t ct) {
        if constexpr (PACKED) {
            table.add(has, a, s, ((unsigned long long)ct << 32) | (unsigned long long)cn, acc);
        } else {
            if (has) {
                atomicAdd(&slt[a], ct);
                atomicAdd(&slc[a], cn);
            }
        }
    };

    // wave-private chunk of the work queue: one global atomic per kChunk items
    constexpr uint32_t kChunk = 64;
    uint32_t c_next = 0, c_end = 0;
    const uint32_t batch = fields_refill_batch(d);
    while (true) {
        if (!exhausted) {
            uint64_t need = __ballot(!active);
            if (__popcll(need) < batch && __any(active)) need = 0;   // batched refills
            while (need) {
                if (c_next >= c_end) {
                    // The work order is cut into kQueues slices; a wave starts on
                    // the slice of its XCD (its L2 then holds that part of the
                    // grid) and steals from the other slices when it runs dry.
                    uint32_t base = 0xffffffffu, lim = 0;
                    const uint32_t me = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11));  // XCC_ID
                    const uint32_t slice = (n_work + kQueues - 1u) / kQueues;
                    for (uint32_t t = 0; t < kQueues && base == 0xffffffffu; t++) {
                        const uint32_t q = (me + t) % kQueues;
                        const uint32_t lo = q * slice, hi = min(n_work, lo + slice);
                        uint32_t b = 0;
                        if (lane == 0) b = atomicAdd(&ctrl->queue[q * kQueueStride], kChunk);
                        b = __shfl(b, 0);
                        if (b < slice && lo + b < hi) {
                            base = lo + b;
                            lim = hi;
                        }
                    }
                    if (base == 0xffffffffu) {
                        exhausted = true;
                        break;
                    }
                    c_next = base;
                    c_end = min(base + kChunk, lim);
                }
                const uint32_t take = min((uint32_t)__popcll(need), c_end - c_next);
                const uint32_t rank = __popcll(need & lt_mask);
                const bool mine = ((need >> lane) & 1ull) && rank < take;
                if (mine) {
                    const uint32_t w = c_next + rank;
                    {
                        const uint32_t sub = w % dd;
                        float2 s;
                        uint32_t seed_idx;
                        if (order) {
                            const SeedRec r = order[w / dd];
                            s = make_float2(r.x, r.y);
                            seed_idx = r.seed_idx;
                        } else {
                            seed_idx = P.seeds_chunk_offset + w / dd;
                            s = seeds[seed_idx];
                        }
                        x = s.x + sub_offset<C>(P, sub % d);
                        y = s.y + sub_offset<C>(P, sub / d);
                        rng = rng_seed(seed_idx);
                        l = 0.f;
                        n = 0;
                        dt = P.dt_max;
                        const uint32_t ix = pix(x, P.pad_width_pp5, nxm1), iy = pix(y, P.pad_width_pp5, nym1);
                        const uint32_t idx = iy + P.ny_padded * ix;
                        if (!mask[idx]) {
                            // atomic_write_sl_data at the seed (jittertrajectory.cl:60-62)
                            Q = load_quad(quad, idx);
                            cx = ix;
                            cy = iy;
                            run_idx = idx;
                            run_slot = slot_of(ix, iy);
                            run_n = 1u;
                            run_t = slt_inc(0.f);
                            active = (0.f < P.max_length) && (0u < P.max_n_steps);
                        }
                    }
                }
                need &= ~__ballot(mine);
                c_next += take;
            }
        }
        // run that ends this step because the streamline entered another pixel
        bool f_has = false;
        uint32_t f_idx = 0, f_slot = 0, f_n = 0, f_t = 0;
        if (active) {
            // compute_step_vec_jittered (computestep.cl:86-106) on quads, with
            // the short sqrt / divide forms where their inputs allow
            const float2 u1 = jitter_unit_cr<C>(P, interp_quad_px_cr<C>(P, Q, x, y, cx, cy), rng);
            const float2 d1 = make_float2(approx_r<R2>(u1.x * combo * dt, res, inv_res), approx_r<R2>(u1.y * combo * dt, res, inv_res));
            const float mx = x + d1.x, my = y + d1.y;
            const uint32_t mix = pix(mx, P.pad_width_pp5, nxm1), miy = pix(my, P.pad_width_pp5, nym1);
            const Quad Q1 = load_quad(quad, miy + P.ny_padded * mix);
            const float2 u2 = jitter_unit_cr<C>(P, interp_quad_px_cr<C>(P, Q1, mx, my, mix, miy), rng);
            const float2 d2 = make_float2(approx_r<R2>(0.5f * mad<C>(u2.x * combo, dt, d1.x), res, inv_res),
                                          approx_r<R2>(0.5f * mad<C>(u2.y * combo, dt, d1.y), res, inv_res));
            const float nx = x + d2.x, ny = y + d2.y;
            const uint32_t nix = pix(nx, P.pad_width_pp5, nxm1), niy = pix(ny, P.pad_width_pp5, nym1);
            const uint32_t nidx = niy + P.ny_padded * nix;
            const uint8_t mk = mask[nidx];
            const Quad Qn = load_quad(quad, nidx);
            if (mk) {
                active = false;  // end pixel masked: step dropped, loop ends
            } else {
                // runge_kutta_step_write_sl_data (rungekutta.cl:163-186)
                const float dl = fast_length_cr<C>(d2.x, d2.y);
                x = nx;
                y = ny;
                Q = Qn;
                cx = nix;
                cy = niy;
                const bool stuck = dl < P.integration_halt_threshold;
                if (!stuck && !dt_fixed) {
                    StepOut o;
                    o.dxy1 = d1;
                    o.dxy2 = d2;
                    dt = next_dt(P, dt, step_error<C>(P, o));
                }
                l += dl;
                n += 1u;
                // check_atomic_write_sl_data (writearray.cl:70-82), run-coalesced
                if (nidx != run_idx) {
                    f_has = true;
                    f_idx = run_idx;
                    f_slot = run_slot;
                    f_n = run_n;
                    f_t = run_t;
                    run_idx = nidx;
                    run_slot = slot_of(nix, niy);
                    run_n = 0;
                    run_t = 0;
                }
                run_n += 1u;
                run_t += slt_inc(l);
                active = !stuck && (l < P.max_length) && (n < P.max_n_steps);
            }
        }
        emit(f_has, f_idx, f_slot, f_n, f_t);
        // runs of streamlines that ended (this step, or at their seed)
        const bool ended = !active && run_n;
        if (__any(ended)) {
            emit(ended, run_idx, run_slot, run_n, run_t);
            if (ended) {
                // the streamline's writes: its seed and n accepted steps
                my_steps += n;
                my_writes += (uint64_t)n + 1u;
                run_n = 0;
                run_t = 0;
            }
        }
        if (exhausted && !__any(active)) break;
    }
    if constexpr (PACKED) {
        table.drain(lane, acc);
        // wave reductions, one atomic per wave
        uint64_t v = my_steps, wv = my_writes;
        for (int off = 32; off > 0; off >>= 1) {
            v += __shfl_down(v, off);
            wv += __shfl_down(wv, off);
        }
        if (lane == 0) {
            if (steps && v) atomicAdd(steps, (unsigned long long)v);
            if (wv) atomicAdd(&ctrl->writes, (unsigned long long)wv);
        }
    }
}

// quad[NY_PADDED*x + y] from uv (layout described in trace_common.hpp)
__global__ void k_build_quads(SlmParams P, const float2* __restrict__ uv, Quad* __restrict__ quad) {
    const size_t n = P.nxy_padded;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i / P.ny_padded), y = (uint32_t)(i - (size_t)x * P.ny_padded);
        const uint32_t xr = min(P.nx_padded - 1u, x + 1u), yu = min(P.ny_padded - 1u, y + 1u);
        const float2 a = uv[(size_t)P.ny_padded * x + y], b = uv[(size_t)P.ny_padded * xr + y];
        const float2 e = uv[(size_t)P.ny_padded * x + yu], f = uv[(size_t)P.ny_padded * xr + yu];
        Quad q;
        q.lo = make_float4(a.x, a.y, b.x, b.y);
        q.hi = make_float4(e.x, e.y, f.x, f.y);
        quad[i] = q;
    }
}

// True when next_dt can only return DT_MAX (slm::dt_is_fixed).
static uint32_t dt_is_fixed(const SlmParams* p) { return slm::dt_is_fixed(*p) ? 1u : 0u; }

__global__ void k_fields_init(FieldsCtrl* ctrl) {
    if (threadIdx.x < kQueues) ctrl->queue[threadIdx.x * kQueueStride] = 0u;
    if (threadIdx.x == 0) {
        ctrl->writes = 0ull;
        ctrl->sumlow = 0ull;
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0ull;
}

// sum over pixels of low32(acc), one atomic per block
__global__ __launch_bounds__(256) void k_fields_sumlow(const unsigned long long* __restrict__ acc, size_t n,
                                                      FieldsCtrl* __restrict__ ctrl) {
    __shared__ unsigned long long part[4];
    unsigned long long v = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v += acc[i] & 0xffffffffull;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(&ctrl->sumlow, t);
    }
}

// Split acc into slc/slt (assign, or add when accumulating) if the packed
// result passed its check; otherwise clear the outputs (unless accumulating)
// for the fallback.  Resets the queue head for the fallback launch.
__global__ void k_fields_commit(const unsigned long long* __restrict__ acc, uint32_t* __restrict__ slc,
                                uint32_t* __restrict__ slt, size_t n, FieldsCtrl* __restrict__ ctrl,
                                uint32_t accumulate, uint32_t force_fallback) {
    const bool ok = fields_packed_ok(ctrl, force_fallback);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (ok) {
            const unsigned long long a = acc[i];
            const uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
            if (accumulate) {
                slc[i] += lo;
                slt[i] += hi;
            } else {
                slc[i] = lo;
                slt[i] = hi;
            }
        } else if (!accumulate) {
            slc[i] = 0u;
            slt[i] = 0u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < kQueues) ctrl->queue[threadIdx.x * kQueueStride] = 0u;
}

static uint32_t grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (uint32_t)(b > 4096 ? 4096 : (b ? b : 1));
}

static int trace_fields_ws(const SlmParams* p, const float* seeds_xy, const uint8_t* mask, const float* uv,
                           uint32_t* slc, uint32_t* slt, uint64_t* steps_out, void* workspace,
                           size_t workspace_bytes, void* stream, uint32_t flags) {
    int rc = check_params(p);
    if (rc) return rc;
    if (!mask || !uv || !slc || !slt) return slm_fail(SLM_E_INVALID_ARG, "NULL mask/uv/slc/slt");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
    const bool accumulate = (flags & SLM_FLAG_NO_ZERO) != 0;
    if (n == 0) {
        if (!accumulate) {
            fill_u32(slc, p->nxy_padded, 0u, s);
            fill_u32(slt, p->nxy_padded, 0u, s);
        }
        return slm_finish(stream, flags);
    }
    if (!seeds_xy) return slm_fail(SLM_E_INVALID_ARG, "NULL seeds");
    const uint32_t d = p->subpixel_seed_point_density ? p->subpixel_seed_point_density : 1u;
    if ((uint64_t)n * d * d >= (1ull << 32) - 64)
        return slm_fail(SLM_E_INVALID_ARG, "n_seeds * d^2 = %llu exceeds 2^32", (unsigned long long)n * d * d);
    const FieldsWs w = fields_ws(p);
    if (!workspace || workspace_bytes < w.total)
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, w.total);
    DevInfo di;
    rc = device_info(di);
    if (rc) return rc;

    char* ws = (char*)workspace;
    FieldsCtrl* ctrl = (FieldsCtrl*)ws;
    SeedRec* order = (SeedRec*)(ws + w.off_order);
    uint32_t* cnt = (uint32_t*)(ws + w.off_cnt);
    uint32_t* cur = (uint32_t*)(ws + w.off_cur);
    unsigned long long* acc = (unsigned long long*)(ws + w.off_acc);
    Quad* quad = (Quad*)(ws + w.off_quad);
    const TileGeom g = tile_geom(p);
    const float2* seeds = (const float2*)seeds_xy;
    const size_t nxy = p->nxy_padded;
    const uint32_t force = (flags & SLM_FLAG_FORCE_FALLBACK) ? 1u : 0u;

    if (!(flags & SLM_FLAG_KEEP_ORDER)) {
        fill_u32(cnt, g.n_keys, 0u, s);
        hipLaunchKernelGGL(k_tile_count, dim3((n + 255) / 256), dim3(256), 0, s, *p, seeds, n, cnt);
        hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, cnt, cur, g.n_keys);
        hipLaunchKernelGGL(k_tile_scatter, dim3((n + 255) / 256), dim3(256), 0, s, *p, seeds, n, cur, order);
    }
    if (!(flags & SLM_FLAG_KEEP_ORDER))
        hipLaunchKernelGGL(k_build_quads, dim3(grid_for(nxy)), dim3(256), 0, s, *p, (const float2*)uv, quad);
    hipLaunchKernelGGL(k_fields_init, dim3(1), dim3(64), 0, s, ctrl);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nxy)), dim3(256), 0, s, acc, nxy);

    const uint32_t n_work = n * d * d;
    // persistent grid: as many 256-thread blocks as can be resident at once
    int per_cu = 0;
    const bool r2 = slm_res_is_pow2(p);
    if (p->fp_contract)
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fields<true, true, true>, 256, 0));
    else
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fields<false, true, true>, 256, 0));
    if (per_cu < 1) per_cu = 1;
    uint32_t grid = (uint32_t)di.cus * (uint32_t)per_cu;
    const uint32_t max_useful = (n_work + 255) / 256;
    if (grid > max_useful) grid = max_useful ? max_useful : 1;
    unsigned long long* st = (unsigned long long*)steps_out;
    const uint32_t fixed = dt_is_fixed(p);
#define SLM_LAUNCH_FIELDS(CC, RR)                                                                            \
    hipLaunchKernelGGL((k_fields<CC, true, RR>), dim3(grid), dim3(256), 0, s, *p, seeds, order, n_work, mask, quad, \
                       acc, slc, slt, st, ctrl, force, fixed)
    if (p->fp_contract) {
        if (r2) SLM_LAUNCH_FIELDS(true, true);
        else SLM_LAUNCH_FIELDS(true, false);
    } else {
        if (r2) SLM_LAUNCH_FIELDS(false, true);
        else SLM_LAUNCH_FIELDS(false, false);
    }
#undef SLM_LAUNCH_FIELDS
    hipLaunchKernelGGL(k_fields_sumlow, dim3(grid_for(nxy)), dim3(256), 0, s, acc, nxy, ctrl);
    hipLaunchKernelGGL(k_fields_commit, dim3(grid_for(nxy)), dim3(256), 0, s, acc, slc, slt, nxy, ctrl,
                       accumulate ? 1u : 0u, force);
    if (p->fp_contract)
        hipLaunchKernelGGL((k_fields<true, false, false>), dim3(grid), dim3(256), 0, s, *p, seeds, order, n_work, mask,
                           quad, acc, slc, slt, nullptr, ctrl, force, fixed);
    else
        hipLaunchKernelGGL((k_fields<false, false, false>), dim3(grid), dim3(256), 0, s, *p, seeds, order, n_work, mask,
                           quad, acc, slc, slt, nullptr, ctrl, force, fixed);
    return slm_finish(stream, flags);
}

extern "C" int slm_trace_fields(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                const float* uv, uint32_t* slc, uint32_t* slt, uint64_t* steps_out,
                                void* workspace, size_t workspace_bytes, void* stream, uint32_t flags) {
    return trace_fields_ws(p, seeds_xy, mask, uv, slc, slt, steps_out, workspace, workspace_bytes, stream, flags);
}

// ------------------------------------------------------------------ fields without allocation
//
// slm_trace_fields_noalloc runs slm_trace_fields with its workspace in a
// fixed device array instead of a caller buffer, for callers that cannot
// allocate (graph capture, serving loops).  Each device holds its own copy of
// the array for the life of the process.  In host code the array's name is
// only a stand-in for the symbol: its device address on the current device is
// looked up once per device with hipGetSymbolAddress and used everywhere.
// Every kernel of the call clears what it needs before use (queue heads,
// counting-sort counts, the packed accumulator), so nothing left by an earlier
// call changes a result.  Calls that overlap on one device share the array and
// race.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static int scratch_ptr(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return slm_fail(SLM_E_HIP, "device id %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return slm_fail(SLM_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                            hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return SLM_OK;
}

extern "C" size_t slm_trace_fields_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int slm_trace_fields_noalloc(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                        const float* uv, uint32_t* slc, uint32_t* slt, uint64_t* steps_out,
                                        void* stream, uint32_t flags) {
    int rc = check_params(p);
    if (rc) return rc;
    const size_t need = fields_ws(p).total;
    if (need > kNoAllocScratchBytes)
        return slm_fail(SLM_E_WORKSPACE, "grid and seeds need %zu bytes of scratch, more than the %zu-byte maximum "
                        "of slm_trace_fields_noalloc (use slm_trace_fields with a workspace)",
                        need, kNoAllocScratchBytes);
    void* ws = nullptr;
    rc = scratch_ptr(&ws);
    if (rc) return rc;
    return trace_fields_ws(p, seeds_xy, mask, uv, slc, slt, steps_out, ws, kNoAllocScratchBytes, stream, flags);
}

// ------------------------------------------------------------------ fields record (test hook)

template <bool C>
__global__ void k_fields_record(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                                const float2* __restrict__ uv, const uint32_t* __restrict__ seed_sel,
                                const uint32_t* __restrict__ sub_sel, uint32_t n_sel, uint32_t max_rec,
                                float2* __restrict__ xy, uint32_t* __restrict__ nrec) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_sel) return;
    const uint32_t d = P.subpixel_seed_point_density ? P.subpixel_seed_point_density : 1u;
    const uint32_t seed_idx = P.seeds_chunk_offset + seed_sel[j];
    const uint32_t sub = sub_sel[j];
    const float2 s = seeds[seed_idx];
    float x = s.x + sub_offset<C>(P, sub % d);
    float y = s.y + sub_offset<C>(P, sub / d);
    uint32_t rng = rng_seed(seed_idx), n = 0, nr = 0;
    float l = 0.f, dt = P.dt_max;
    float2* rec = xy + (size_t)j * max_rec;
    if (nr < max_rec) rec[nr] = make_float2(x, y);
    nr++;
    uint32_t idx = array_idx(P, x, y);
    bool active = !mask[idx] && (0.f < P.max_length) && (0u < P.max_n_steps);
    while (active) {
        StepOut o;
        compute_step<C, true>(P, uv, dt, x, y, rng, o);
        if (mask[o.idx]) break;
        const float err = step_error<C>(P, o);
        const float dl = fast_length<C>(o.dxy2.x, o.dxy2.y);
        x = o.next.x;
        y = o.next.y;
        const bool stuck = dl < P.integration_halt_threshold;
        if (!stuck) dt = next_dt(P, dt, err);
        l += dl;
        n += 1u;
        if (nr < max_rec) rec[nr] = make_float2(x, y);
        nr++;
        active = !stuck && (l < P.max_length) && (n < P.max_n_steps);
    }
    nrec[j] = nr;
}

extern "C" int slm_trace_fields_record(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                       const float* uv, const uint32_t* seed_sel, const uint32_t* sub_sel,
                                       uint32_t n_sel, uint32_t max_rec, float* xy, uint32_t* nrec,
                                       void* stream, uint32_t flags) {
    int rc = check_params(p);
    if (rc) return rc;
    if (n_sel == 0) return slm_finish(stream, flags);
    if (!seeds_xy || !mask || !uv || !seed_sel || !sub_sel || !nrec || (max_rec && !xy))
        return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    if (p->fp_contract)
        hipLaunchKernelGGL(k_fields_record<true>, dim3((n_sel + 63) / 64), dim3(64), 0, s, *p,
                           (const float2*)seeds_xy, mask, (const float2*)uv, seed_sel, sub_sel, n_sel, max_rec,
                           (float2*)xy, nrec);
    else
        hipLaunchKernelGGL(k_fields_record<false>, dim3((n_sel + 63) / 64), dim3(64), 0, s, *p,
                           (const float2*)seeds_xy, mask, (const float2*)uv, seed_sel, sub_sel, n_sel, max_rec,
                           (float2*)xy, nrec);
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ trajectory record

// trajectory_record (trajectory.cl:41-80) with runge_kutta_step_record
// (rungekutta.cl:49-77) and euler_step_record (:104-120).
template <bool C>
__global__ void k_paths(SlmParams P, const float2* __restrict__ seeds, const uint8_t* __restrict__ mask,
                        const float2* __restrict__ uv, char2* __restrict__ traj, uint16_t* __restrict__ nsteps,
                        float* __restrict__ length, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float res = P.trajectory_resolution;
    const float2 s = seeds[P.seeds_chunk_offset + k];
    char2* rec = traj + (size_t)k * P.max_n_steps;
    float x = s.x, y = s.y, px = x, py = y;
    float l = 0.f, dl = 0.f, dt = P.dt_max;
    uint32_t nst = 0, rng = 0;
    uint32_t idx = array_idx(P, x, y);
    const bool fixed = dt_is_fixed(P);
    while (!mask[idx] && (l < P.max_length && nst < P.max_n_steps)) {
        StepOut o;
        compute_step<C, false, false, true>(P, uv, dt, x, y, rng, o);
        idx = o.idx;
        if (!mask[idx]) {
            dl = fast_length_cr<C>(o.dxy2.x, o.dxy2.y);
            float2 nx = o.next;
            if (dl < P.integration_halt_threshold) {
                nx = make_float2(x + o.dxy1.x, y + o.dxy1.y);
                dl = fast_length_cr<C>(o.dxy1.x, o.dxy1.y);
            }
            x = nx.x;
            y = nx.y;
            idx = array_idx(P, x, y);
            const bool stuck = dl < P.integration_halt_threshold;
            if (!stuck && !fixed) dt = next_dt(P, dt, step_error<C>(P, o));
            l += dl;
            rec[nst] = make_char2(compress1(x - px, res), compress1(y - py, res));
            nst += 1u;
            if (stuck) break;
            px = x;
            py = y;
        } else {
            // Euler hop to the nearest pixel edge along DOWNUP_SIGN*uv1
            const float su = o.uv1.x * P.downup_sign, sv = o.uv1.y * P.downup_sign;
            float dtx = 0.f, dty = 0.f;
            if (su > 0.f) dtx = fdiv((float)cvt_i32(x + 1.5f) - (x + 0.5f), su);
            if (su < 0.f) dtx = fdiv(-((x + 0.5f) - (float)cvt_i32(x + 0.5f)), su);
            if (sv > 0.f) dty = fdiv((float)cvt_i32(y + 1.5f) - (y + 0.5f), sv);
            if (sv < 0.f) dty = fdiv(-((y + 0.5f) - (float)cvt_i32(y + 0.5f)), sv);
            const float ax = fabsf(dtx), ay = fabsf(dty);
            dt = (ax < ay) ? dtx : ((ay < ax) ? dty : fminf(dtx, dty));
            x = mad<C>(su, dt, x);
            y = mad<C>(sv, dt, y);
            x = approx_abs(x, res, P.approx_wrap);
            y = approx_abs(y, res, P.approx_wrap);
            x = fminf(fmaxf(x, -0.5f), P.nxf_mp5);
            y = fminf(fmaxf(y, -0.5f), P.nyf_mp5);
            dl = fast_length<C>(x - px, y - py);
            l += dl;
            rec[nst] = make_char2(compress1(x - px, res), compress1(y - py, res));
            nst += 1u;
            break;
        }
    }
    nsteps[k] = (uint16_t)nst;
    length[k] = l;
}

extern "C" int slm_trace_paths(const SlmParams* p, const float* seeds_xy, const uint8_t* mask, const float* uv,
                               int8_t* traj, uint16_t* nsteps, float* length, void* stream, uint32_t flags) {
    int rc = check_params(p);
    if (rc) return rc;
    const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
    if (n == 0) return slm_finish(stream, flags);
    if (!seeds_xy || !mask || !uv || !traj || !nsteps || !length) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (p->max_n_steps > 65535u)
        return slm_fail(SLM_E_INVALID_ARG, "max_n_steps %u does not fit the uint16 step count", p->max_n_steps);
    hipStream_t s = (hipStream_t)stream;
    if (p->fp_contract)
        hipLaunchKernelGGL(k_paths<true>, dim3((n + 63) / 64), dim3(64), 0, s, *p, (const float2*)seeds_xy, mask,
                           (const float2*)uv, (char2*)traj, nsteps, length, n);
    else
        hipLaunchKernelGGL(k_paths<false>, dim3((n + 63) / 64), dim3(64), 0, s, *p, (const float2*)seeds_xy, mask,
                           (const float2*)uv, (char2*)traj, nsteps, length, n);
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ trajectory packing
//
// The recorded trajectories of slm_trace_paths sit in rows of max_n_steps
// steps, mostly unused.  Packing them end to end takes an exclusive prefix
// sum of the step counts (offsets, n + 1 entries) and one copy of the used
// steps, so that only those leave HBM.  The sum runs in place over offsets,
// in blocks of kPackBlock counts: pass 1 leaves block b's total in its first
// slot offsets[b * kPackBlock]; pass 2 (one block) turns those slots into
// exclusive block bases and writes the grand total to offsets[n]; pass 3
// scans each block from its base.  Counts above max_n_steps are clamped in
// every pass, so the packed rows never read past their row.

constexpr uint32_t kPackThreads = 256;
constexpr uint32_t kPackPer = 4;
constexpr uint32_t kPackBlock = kPackThreads * kPackPer;

// exclusive scan of one value per thread of a kPackThreads block; returns the
// thread's prefix and leaves the block total in *total (sh: kPackThreads slots)
__device__ inline uint64_t pack_block_scan(uint64_t v, uint64_t* sh, uint64_t* total) {
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < kPackThreads; d <<= 1) {
        const uint64_t add = t >= d ? sh[t - d] : 0ull;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const uint64_t incl = sh[t];
    *total = sh[kPackThreads - 1];
    __syncthreads();
    return incl - v;
}

__device__ inline uint32_t pack_count(const uint16_t* nsteps, uint32_t k, uint32_t n, uint32_t max_n_steps) {
    return k < n ? min((uint32_t)nsteps[k], max_n_steps) : 0u;
}

__global__ __launch_bounds__(kPackThreads) void k_pack_sums(const uint16_t* __restrict__ nsteps, uint32_t n,
                                                            uint32_t max_n_steps, uint64_t* __restrict__ offsets) {
    __shared__ uint64_t sh[kPackThreads];
    const size_t k0 = (size_t)blockIdx.x * kPackBlock + threadIdx.x * kPackPer;
    uint64_t s = 0;
    for (uint32_t j = 0; j < kPackPer; ++j)
        if (k0 + j < n) s += pack_count(nsteps, (uint32_t)(k0 + j), n, max_n_steps);
    uint64_t total;
    pack_block_scan(s, sh, &total);
    if (threadIdx.x == 0) offsets[(size_t)blockIdx.x * kPackBlock] = total;
}

__global__ __launch_bounds__(kPackThreads) void k_pack_bases(uint32_t n_blocks, uint32_t n,
                                                             uint64_t* __restrict__ offsets) {
    __shared__ uint64_t sh[kPackThreads];
    uint64_t carry = 0;
    for (uint32_t b0 = 0; b0 < n_blocks; b0 += kPackThreads) {
        const uint32_t b = b0 + threadIdx.x;
        const uint64_t v = b < n_blocks ? offsets[(size_t)b * kPackBlock] : 0ull;
        uint64_t total;
        const uint64_t pre = pack_block_scan(v, sh, &total);
        if (b < n_blocks) offsets[(size_t)b * kPackBlock] = carry + pre;
        carry += total;
    }
    if (threadIdx.x == 0) offsets[n] = carry;
}

__global__ __launch_bounds__(kPackThreads) void k_pack_scan(const uint16_t* __restrict__ nsteps, uint32_t n,
                                                            uint32_t max_n_steps, uint64_t* __restrict__ offsets) {
    __shared__ uint64_t sh[kPackThreads];
    const size_t first = (size_t)blockIdx.x * kPackBlock;
    const uint64_t base = offsets[first];      // read by every thread before any write below
    const size_t k0 = first + threadIdx.x * kPackPer;
    uint32_t c[kPackPer];
    uint64_t s = 0;
    for (uint32_t j = 0; j < kPackPer; ++j) {
        c[j] = k0 + j < n ? pack_count(nsteps, (uint32_t)(k0 + j), n, max_n_steps) : 0u;
        s += c[j];
    }
    uint64_t total;
    uint64_t o = base + pack_block_scan(s, sh, &total);    // its barriers order the reads of base first
    for (uint32_t j = 0; j < kPackPer; ++j) {
        if (k0 + j < n) offsets[k0 + j] = o;
        o += c[j];
    }
}

// one wave per record: its used steps (2 bytes each) copied to their offset
__global__ __launch_bounds__(kPackThreads) void k_pack_copy(const uint16_t* __restrict__ traj,
                                                            const uint16_t* __restrict__ nsteps,
                                                            const uint64_t* __restrict__ offsets, uint32_t n,
                                                            uint32_t max_n_steps, uint16_t* __restrict__ packed) {
    const uint32_t lane = threadIdx.x & 63u;
    const size_t waves = (size_t)gridDim.x * (kPackThreads / 64);
    for (size_t k = (size_t)blockIdx.x * (kPackThreads / 64) + (threadIdx.x >> 6); k < n; k += waves) {
        const uint32_t cnt = min((uint32_t)nsteps[k], max_n_steps);
        const uint16_t* src = traj + k * max_n_steps;
        uint16_t* dst = packed + offsets[k];
        for (uint32_t i = lane; i < cnt; i += 64) dst[i] = src[i];
    }
}

extern "C" int slm_paths_offsets(const uint16_t* nsteps, uint32_t n, uint32_t max_n_steps, uint64_t* offsets,
                                 void* stream, uint32_t flags) {
    if (!offsets || (n && !nsteps)) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (unsigned long long*)offsets, (size_t)1);
        return slm_finish(stream, flags);
    }
    const uint32_t nb = (uint32_t)(((size_t)n + kPackBlock - 1) / kPackBlock);
    hipLaunchKernelGGL(k_pack_sums, dim3(nb), dim3(kPackThreads), 0, s, nsteps, n, max_n_steps, offsets);
    hipLaunchKernelGGL(k_pack_bases, dim3(1), dim3(kPackThreads), 0, s, nb, n, offsets);
    hipLaunchKernelGGL(k_pack_scan, dim3(nb), dim3(kPackThreads), 0, s, nsteps, n, max_n_steps, offsets);
    return slm_finish(stream, flags);
}

extern "C" int slm_paths_pack(const int8_t* traj, const uint16_t* nsteps, const uint64_t* offsets, uint32_t n,
                              uint32_t max_n_steps, int8_t* packed, void* stream, uint32_t flags) {
    if (n == 0) return slm_finish(stream, flags);
    if (!traj || !nsteps || !offsets || !packed) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (max_n_steps == 0) return slm_fail(SLM_E_INVALID_ARG, "max_n_steps is 0");
    hipStream_t s = (hipStream_t)stream;
    const size_t blocks = std::min<size_t>(((size_t)n + 3) / 4, 65536);
    hipLaunchKernelGGL(k_pack_copy, dim3((uint32_t)blocks), dim3(kPackThreads), 0, s, (const uint16_t*)traj, nsteps,
                       offsets, n, max_n_steps, (uint16_t*)packed);
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ hillslope lengths

// hillslope_lengths (lengths.cl:37-116) with lengths_runge_kutta_step
// (rungekutta.cl:492-516): one lane per mid-slope/ridge seed.
template <bool C>
__global__ __launch_bounds__(256) void k_hsl(SlmParams P, const float2* __restrict__ seeds,
                                             const uint8_t* __restrict__ mask, const float2* __restrict__ uv,
                                             const uint32_t* __restrict__ mapping, uint32_t thin_flag,
                                             uint32_t from_flag, float* __restrict__ length) {
    float x = 0.f, y = 0.f, l = 0.f, dt = 0.f;
    uint32_t n = 0, rng = 0, idx = 0, k = 0;
    bool moved_off = false;
    const bool fixed = dt_is_fixed(P);
    CycleSkip cyc;
    auto start = [&](uint32_t kk) -> bool {
        const float2 s = seeds[kk];
        k = kk;
        x = s.x;
        y = s.y;
        l = 0.f;
        dt = P.dt_max;
        n = 0;
        moved_off = false;
        idx = array_idx(P, x, y);
        cyc.reset();
        return true;
    };
    auto step = [&]() -> bool {
        if (!(((~mapping[idx]) & thin_flag) && n < P.max_n_steps)) {
            if (mapping[idx] & thin_flag) length[k] = l * P.pixel_size;
            return false;
        }
        StepOut o;
        compute_step<C, false, false, true>(P, uv, dt, x, y, rng, o);
        idx = o.idx;
        if (mask[idx]) return false;
        if (!moved_off && ((~mapping[idx]) & from_flag)) moved_off = true;
        else if (moved_off && (mapping[idx] & from_flag)) return false;
        float dl = fast_length_cr<C>(o.dxy2.x, o.dxy2.y);
        float2 nx = o.next;
        if (dl < P.integration_halt_threshold) {
            nx = make_float2(x + o.dxy1.x, y + o.dxy1.y);
            dl = fast_length_cr<C>(o.dxy1.x, o.dxy1.y);
        }
        x = nx.x;
        y = nx.y;
        idx = array_idx(P, x, y);
        if (dl < P.integration_halt_threshold) return false;
        if (!fixed) dt = next_dt(P, dt, step_error<C>(P, o));
        l += dl;
        n += 1u;
        // A repeated state (position, dt, moved_off) comes back around a cycle
        // of non-thin pixels: the trace would run to MAX_N_STEPS and end off
        // the thin pixels, writing nothing.
        if (cyc.check(x, y, dt, moved_off ? 1u : 0u, n, P.max_n_steps)) return false;
        return true;
    };
    persist_seeds(P.seeds_chunk_offset, P.n_seed_points - P.seeds_chunk_offset, start, step);
}

extern "C" int slm_hillslope_lengths(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                     const float* uv, const uint32_t* mapping, uint32_t thin_flag,
                                     uint32_t from_flag, float* length, void* stream, uint32_t flags) {
    int rc = check_params(p);
    if (rc) return rc;
    const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
    if (n == 0) return slm_finish(stream, flags);
    if (!seeds_xy || !mask || !uv || !mapping || !length) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    if (p->fp_contract)
        hipLaunchKernelGGL(k_hsl<true>, persist_grid(k_hsl<true>, n), dim3(256), 0, s, *p, (const float2*)seeds_xy,
                           mask, (const float2*)uv, mapping, thin_flag, from_flag, length);
    else
        hipLaunchKernelGGL(k_hsl<false>, persist_grid(k_hsl<false>, n), dim3(256), 0, s, *p, (const float2*)seeds_xy,
                           mask, (const float2*)uv, mapping, thin_flag, from_flag, length);
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ tiled fields tracer
//
// LDS-owned tiles.  The padded grid is cut into T x T pixel tiles.  A unit of
// work is (tile, up to U streamline states queued in that tile); one 256-thread
// block processes a unit with the tile's region in LDS:
//   region   = tile + H-pixel hysteresis band + M-pixel access margin, side R
//   uv_l     = uv of the region plus one row/column (interpolation corners)
//   mask_l   = mask of the region          acc_l = packed (slt<<32)|slc sums
// A streamline steps while its pixel stays inside tile+band (a step moves a
// position by less than one pixel, M = 2 covers the mid and end pixels) and
// at most cap_steps times per visit; then its state is emitted to the queue
// of the tile that owns its pixel.  Rounds repeat: bin emitted states by tile
// (LDS-histogram counting sort), cut bins into units, process, until no state
// is left.  Each unit ends by adding its non-zero acc_l entries to the global
// acc with one 64-bit atomic per pixel.  All sums are integers, so the
// result is the reference's per-write atomics exactly; the packed-sum check
// and 32-bit fallback of the queue tracer apply unchanged.

struct TState {  // 32 bytes
    float x, y, l, dt;
    uint32_t n, rng, tile, fresh;
};

struct TileCfg {
    uint32_t T, H, M, R, RU;   // R = T + 2H + 2M; RU = R + 1 (uv side)
    uint32_t ntx, nty, n_tiles;
    uint32_t unit, cap_steps;
};

struct TUnit {
    uint32_t tile, begin, end, pad;
};

struct TileCtrl {
    FieldsCtrl f;               // queue (unused here), W, sum of low halves
    uint32_t n_out;             // states emitted by the current round
    uint32_t unit_head;         // next unit to take
    uint32_t n_units;           // units of the current round
    uint32_t pad;
};

static void tile_opts_defaults(SlmTileOpts* o) {
    o->tile = o->tile ? o->tile : 32u;
    o->block = o->block ? o->block : 256u;
    o->unit = o->unit ? o->unit : 8u * o->block;
    o->cap_steps = o->cap_steps ? o->cap_steps : 4u * o->tile;
    o->halo = o->halo ? o->halo : 4u;
}

static int check_tile_opts(const SlmTileOpts* o) {
    if (o->tile < 8 || o->tile > 128) return slm_fail(SLM_E_INVALID_ARG, "tile side %u outside [8, 128]", o->tile);
    if (o->block != 256 && o->block != 512 && o->block != 1024)
        return slm_fail(SLM_E_INVALID_ARG, "block %u is not 256, 512 or 1024", o->block);
    if (o->halo > 32) return slm_fail(SLM_E_INVALID_ARG, "halo %u above 32", o->halo);
    if (o->unit == 0 || o->cap_steps == 0) return slm_fail(SLM_E_INVALID_ARG, "unit and cap_steps must be > 0");
    return SLM_OK;
}

static TileCfg tile_cfg(const SlmParams* p, const SlmTileOpts& o) {
    TileCfg c;
    c.T = o.tile;
    c.H = o.halo;
    c.M = 2;
    c.R = c.T + 2 * c.H + 2 * c.M;
    c.RU = c.R + 1;
    c.ntx = (p->nx_padded + c.T - 1) / c.T;
    c.nty = (p->ny_padded + c.T - 1) / c.T;
    c.n_tiles = c.ntx * c.nty;
    c.unit = o.unit;
    c.cap_steps = o.cap_steps;
    return c;
}

static size_t tile_lds_bytes(const TileCfg& c) {
    return (size_t)c.RU * c.RU * 8 + (size_t)c.R * c.R * 8 + ((size_t)c.R * c.R + 15) / 16 * 16;
}

__device__ __forceinline__ uint32_t tile_of(const TileCfg& c, uint32_t px, uint32_t py) {
    return (px / c.T) * c.nty + (py / c.T);
}

// fresh states: work item w = seed ordinal w/d^2, sub-pixel start w%d^2
template <bool C>
__global__ void k_tl_fresh(SlmParams P, TileCfg c, const float2* __restrict__ seeds, uint32_t n_work,
                           TState* __restrict__ out) {
    const uint32_t d = P.subpixel_seed_point_density ? P.subpixel_seed_point_density : 1u;
    const uint32_t dd = d * d;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n_work; w += gridDim.x * blockDim.x) {
        const uint32_t k = w / dd, sub = w - k * dd;
        const uint32_t seed_idx = P.seeds_chunk_offset + k;
        const float2 s = seeds[seed_idx];
        TState t;
        t.x = s.x + sub_offset<C>(P, sub % d);
        t.y = s.y + sub_offset<C>(P, sub / d);
        t.l = 0.f;
        t.dt = P.dt_max;
        t.n = 0u;
        t.rng = rng_seed(seed_idx);
        t.tile = tile_of(c, pix(t.x, P.pad_width_pp5, P.nx_padded - 1u), pix(t.y, P.pad_width_pp5, P.ny_padded - 1u));
        t.fresh = 1u;
        out[w] = t;
    }
}

// Per-tile counts of n states; LDS histogram when the tiles fit (lds_bins).
__global__ __launch_bounds__(256) void k_tl_count(const TState* __restrict__ st, uint32_t n, uint32_t n_tiles,
                                                  uint32_t* __restrict__ cnt, uint32_t use_lds) {
    extern __shared__ uint32_t h[];
    if (use_lds) {
        for (uint32_t i = threadIdx.x; i < n_tiles; i += blockDim.x) h[i] = 0u;
        __syncthreads();
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t t = st[i].tile;
        if (use_lds) atomicAdd(&h[t], 1u);
        else atomicAdd(&cnt[t], 1u);
    }
    if (use_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n_tiles; i += blockDim.x)
            if (h[i]) atomicAdd(&cnt[i], h[i]);
    }
}

// Exclusive scan of tile counts -> off (and cursor copy), units per tile ->
// unit offsets; total units into ctrl.  One 1024-thread block.
__global__ __launch_bounds__(1024) void k_tl_scan(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ off,
                                                  uint32_t* __restrict__ cur, uint32_t* __restrict__ uoff,
                                                  uint32_t n_tiles, uint32_t unit, TileCtrl* __restrict__ ctrl) {
    __shared__ uint32_t ps[1024], pu[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n_tiles + 1023u) / 1024u;
    const uint32_t b = t * per, e = min(n_tiles, b + per);
    uint32_t s = 0, u = 0;
    for (uint32_t i = b; i < e; i++) {
        s += cnt[i];
        u += (cnt[i] + unit - 1u) / unit;
    }
    ps[t] = s;
    pu[t] = u;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        const uint32_t vs = (t >= o) ? ps[t - o] : 0u, vu = (t >= o) ? pu[t - o] : 0u;
        __syncthreads();
        ps[t] += vs;
        pu[t] += vu;
        __syncthreads();
    }
    uint32_t rs = t ? ps[t - 1] : 0u, ru = t ? pu[t - 1] : 0u;
    for (uint32_t i = b; i < e; i++) {
        off[i] = rs;
        cur[i] = rs;
        uoff[i] = ru;
        rs += cnt[i];
        ru += (cnt[i] + unit - 1u) / unit;
    }
    if (t == 1023) {
        ctrl->n_units = pu[1023];
        ctrl->unit_head = 0u;
        ctrl->n_out = 0u;
    }
}

// Scatter states into tile order: per block, an LDS histogram reserves one
// global range per tile, then each state takes a slot in its tile's range.
__global__ __launch_bounds__(256) void k_tl_scatter(const TState* __restrict__ st, uint32_t n, uint32_t n_tiles,
                                                    uint32_t* __restrict__ cur, TState* __restrict__ out,
                                                    uint32_t use_lds) {
    extern __shared__ uint32_t h[];  // [0, n_tiles): counts then bases; [n_tiles, 2n_tiles): ranks
    uint32_t* rank = h + n_tiles;
    const uint32_t per_block = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t b = blockIdx.x * per_block, e = min(n, b + per_block);
    if (use_lds) {
        for (uint32_t i = threadIdx.x; i < n_tiles; i += blockDim.x) {
            h[i] = 0u;
            rank[i] = 0u;
        }
        __syncthreads();
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) atomicAdd(&h[st[i].tile], 1u);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n_tiles; i += blockDim.x)
            if (h[i]) h[i] = atomicAdd(&cur[i], h[i]);
        __syncthreads();
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
            const TState s = st[i];
            out[h[s.tile] + atomicAdd(&rank[s.tile], 1u)] = s;
        }
    } else {
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
            const TState s = st[i];
            out[atomicAdd(&cur[s.tile], 1u)] = s;
        }
    }
}

__global__ void k_tl_units(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                           const uint32_t* __restrict__ uoff, uint32_t n_tiles, uint32_t unit,
                           TUnit* __restrict__ units) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles; t += gridDim.x * blockDim.x) {
        const uint32_t c = cnt[t];
        uint32_t u = uoff[t];
        for (uint32_t b = 0; b < c; b += unit, u++) {
            TUnit x;
            x.tile = t;
            x.begin = off[t] + b;
            x.end = off[t] + min(c, b + unit);
            x.pad = 0u;
            units[u] = x;
        }
    }
}

// Process units: see the section comment.  in: binned states; out: emitted.
template <bool C, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_tl_process(SlmParams P, TileCfg c, const TState* __restrict__ in,
                                                    TState* __restrict__ out, const TUnit* __restrict__ units,
                                                    const uint8_t* __restrict__ mask, const float2* __restrict__ uv,
                                                    unsigned long long* __restrict__ acc,
                                                    unsigned long long* __restrict__ steps,
                                                    TileCtrl* __restrict__ ctrl, uint32_t dt_fixed,
                                                    uint32_t out_cap) {
    extern __shared__ __align__(16) unsigned char lds[];
    float2* uv_l = (float2*)lds;                                          // RU*RU
    unsigned long long* acc_l = (unsigned long long*)(lds + (size_t)c.RU * c.RU * 8);   // R*R
    uint8_t* mask_l = lds + (size_t)c.RU * c.RU * 8 + (size_t)c.R * c.R * 8;            // R*R
    __shared__ uint32_t s_unit, s_next;

    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const float res = P.trajectory_resolution;
    const float combo = P.combo_factor;
    const uint32_t R = c.R, RU = c.RU;
    uint64_t my_steps = 0, my_writes = 0;

    while (true) {
        if (threadIdx.x == 0) s_unit = atomicAdd(&ctrl->unit_head, 1u);
        __syncthreads();
        const uint32_t u = s_unit;
        if (u >= ctrl->n_units) break;
        const TUnit U = units[u];
        const int32_t ox = (int32_t)((U.tile / c.nty) * c.T) - (int32_t)(c.H + c.M);
        const int32_t oy = (int32_t)((U.tile % c.nty) * c.T) - (int32_t)(c.H + c.M);
        const int32_t sx0 = ox + (int32_t)c.M, sy0 = oy + (int32_t)c.M;   // stay band origin
        const int32_t sw = (int32_t)(c.T + 2 * c.H);                         // stay band side
        for (uint32_t i = threadIdx.x; i < RU * RU; i += blockDim.x) {
            const int32_t gx = min(max(ox + (int32_t)(i / RU), 0), (int32_t)P.nx_padded - 1);
            const int32_t gy = min(max(oy + (int32_t)(i % RU), 0), (int32_t)P.ny_padded - 1);
            uv_l[i] = uv[(size_t)gx * P.ny_padded + gy];
        }
        for (uint32_t i = threadIdx.x; i < R * R; i += blockDim.x) {
            const int32_t gx = min(max(ox + (int32_t)(i / R), 0), (int32_t)P.nx_padded - 1);
            const int32_t gy = min(max(oy + (int32_t)(i % R), 0), (int32_t)P.ny_padded - 1);
            mask_l[i] = mask[(size_t)gx * P.ny_padded + gy];
            acc_l[i] = 0ull;
        }
        if (threadIdx.x == 0) s_next = U.begin;
        __syncthreads();

        // local index of a global pixel, or 0xffffffff outside the mask/acc region
        auto loc = [&](uint32_t px, uint32_t py) -> uint32_t {
            const uint32_t li = (uint32_t)((int32_t)px - ox), lj = (uint32_t)((int32_t)py - oy);
            return (li < R && lj < R) ? li * R + lj : 0xffffffffu;
        };
        // speed_interpolator from the region (corners read at +1 row/column of uv_l)
        auto interp = [&](float x, float y) -> float2 {
            const uint32_t xl = pix(x, P.pad_width_pp5, P.nx_padded - 1u);
            const uint32_t yd = pix(y, P.pad_width_pp5, P.ny_padded - 1u);
            const uint32_t li = (uint32_t)((int32_t)xl - ox), lj = (uint32_t)((int32_t)yd - oy);
            Quad Q;
            if (li < R && lj < R) {
                const float2 a = uv_l[li * RU + lj], b = uv_l[(li + 1) * RU + lj];
                const float2 e = uv_l[li * RU + lj + 1], f = uv_l[(li + 1) * RU + lj + 1];
                Q.lo = make_float4(a.x, a.y, b.x, b.y);
                Q.hi = make_float4(e.x, e.y, f.x, f.y);
            } else {  // outside the region: read HBM (never taken: steps move < 1 pixel)
                const uint32_t xr = min(P.nx_padded - 1u, xl + 1u), yu = min(P.ny_padded - 1u, yd + 1u);
                const float2 a = uv[(size_t)P.ny_padded * xl + yd], b = uv[(size_t)P.ny_padded * xr + yd];
                const float2 e = uv[(size_t)P.ny_padded * xl + yu], f = uv[(size_t)P.ny_padded * xr + yu];
                Q.lo = make_float4(a.x, a.y, b.x, b.y);
                Q.hi = make_float4(e.x, e.y, f.x, f.y);
            }
            return interp_quad<C>(P, Q, x, y);
        };

        bool active = false, done = false;
        TState S;
        uint32_t run_li = 0, run_n = 0, run_t = 0, visit = 0;
        auto flush = [&]() {
            if (run_n) {
                atomicAdd(&acc_l[run_li], ((unsigned long long)run_t << 32) | (unsigned long long)run_n);
                run_n = 0;
                run_t = 0;
            }
        };
        while (true) {
            if (!done) {
                const uint64_t need = __ballot(!active);
                if (need) {
                    const uint32_t cnt = __popcll(need);
                    const uint32_t leader = __ffsll((unsigned long long)need) - 1u;
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(&s_next, cnt);
                    base = __shfl(base, leader);
                    if (base + cnt >= U.end) done = true;
                    if (!active) {
                        const uint32_t w = base + __popcll(need & lt_mask);
                        if (w < U.end) {
                            S = in[w];
                            visit = 0;
                            active = true;
                            if (S.fresh) {
                                // atomic_write_sl_data at the seed (jittertrajectory.cl:60-62)
                                S.fresh = 0u;
                                const uint32_t px = pix(S.x, P.pad_width_pp5, P.nx_padded - 1u);
                                const uint32_t py = pix(S.y, P.pad_width_pp5, P.ny_padded - 1u);
                                const uint32_t li = loc(px, py);
                                const uint8_t mk = (li != 0xffffffffu) ? mask_l[li] : mask[(size_t)px * P.ny_padded + py];
                                if (mk) {
                                    active = false;
                                } else {
                                    my_writes += 1u;
                                    if (li != 0xffffffffu) {
                                        run_li = li;
                                        run_n = 1u;
                                        run_t = slt_inc(0.f);
                                    } else {
                                        atomicAdd(&acc[(size_t)px * P.ny_padded + py], 1ull);
                                    }
                                    active = (0.f < P.max_length) && (0u < P.max_n_steps);
                                    if (!active) flush();
                                }
                            }
                        }
                    }
                }
            }
            if (!__any(active)) {
                if (done) break;
                continue;
            }
            if (active) {
                // compute_step_vec_jittered (computestep.cl:86-106) from LDS
                const float2 u1 = jitter_unit<C>(P, interp(S.x, S.y), S.rng);
                const float2 d1 = make_float2(approx(u1.x * combo * S.dt, res), approx(u1.y * combo * S.dt, res));
                const float2 u2 = jitter_unit<C>(P, interp(S.x + d1.x, S.y + d1.y), S.rng);
                const float2 d2 = make_float2(approx(0.5f * mad<C>(u2.x * combo, S.dt, d1.x), res),
                                              approx(0.5f * mad<C>(u2.y * combo, S.dt, d1.y), res));
                const float nx = S.x + d2.x, ny = S.y + d2.y;
                const uint32_t npx = pix(nx, P.pad_width_pp5, P.nx_padded - 1u);
                const uint32_t npy = pix(ny, P.pad_width_pp5, P.ny_padded - 1u);
                const uint32_t nli = loc(npx, npy);
                const size_t ngi = (size_t)npx * P.ny_padded + npy;
                const uint8_t mk = (nli != 0xffffffffu) ? mask_l[nli] : mask[ngi];
                if (mk) {
                    active = false;  // end pixel masked: step dropped, loop ends
                    flush();
                } else {
                    // runge_kutta_step_write_sl_data (rungekutta.cl:163-186)
                    const float dl = fast_length<C>(d2.x, d2.y);
                    S.x = nx;
                    S.y = ny;
                    const bool stuck = dl < P.integration_halt_threshold;
                    if (!stuck && !dt_fixed) {
                        StepOut o;
                        o.dxy1 = d1;
                        o.dxy2 = d2;
                        S.dt = next_dt(P, S.dt, step_error<C>(P, o));
                    }
                    S.l += dl;
                    S.n += 1u;
                    my_steps += 1u;
                    my_writes += 1u;
                    if (nli != 0xffffffffu) {
                        if (nli != run_li || run_n == 0u) {
                            flush();
                            run_li = nli;
                        }
                        run_n += 1u;
                        run_t += slt_inc(S.l);
                    } else {
                        flush();
                        atomicAdd(&acc[ngi], ((unsigned long long)slt_inc(S.l) << 32) | 1ull);
                    }
                    active = !stuck && (S.l < P.max_length) && (S.n < P.max_n_steps);
                    visit += 1u;
                    if (!active) {
                        flush();
                    } else {
                        const bool inside = (uint32_t)((int32_t)npx - sx0) < (uint32_t)sw &&
                                            (uint32_t)((int32_t)npy - sy0) < (uint32_t)sw;
                        if (!inside || visit >= c.cap_steps) {
                            flush();
                            S.tile = tile_of(c, npx, npy);
                            // wave-aggregated append to the emitted list
                            const uint64_t em = __ballot(true);
                            const uint32_t lead = __ffsll((unsigned long long)em) - 1u;
                            uint32_t ob = 0;
                            if (lane == lead) ob = atomicAdd(&ctrl->n_out, (uint32_t)__popcll(em));
                            ob = __shfl(ob, lead);
                            const uint32_t slot = ob + __popcll(em & lt_mask);
                            if (slot < out_cap) out[slot] = S;   // overflow is reported by the host
                            active = false;
                        }
                    }
                }
            }
        }
        __syncthreads();
        // add this unit's sums to the grid (one 64-bit atomic per touched pixel)
        for (uint32_t i = threadIdx.x; i < R * R; i += blockDim.x) {
            const unsigned long long v = acc_l[i];
            const int32_t gx = ox + (int32_t)(i / R), gy = oy + (int32_t)(i % R);
            if (v && gx >= 0 && gy >= 0 && gx < (int32_t)P.nx_padded && gy < (int32_t)P.ny_padded)
                atomicAdd(&acc[(size_t)gx * P.ny_padded + gy], v);
        }
        __syncthreads();
    }
    uint64_t v = my_steps, wv = my_writes;
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off);
        wv += __shfl_down(wv, off);
    }
    if (lane == 0) {
        if (steps && v) atomicAdd(steps, (unsigned long long)v);
        if (wv) atomicAdd(&ctrl->f.writes, (unsigned long long)wv);
    }
}

struct TiledWs {
    size_t off_acc, off_quad, off_a, off_b, off_cnt, off_off, off_cur, off_uoff, off_units, total;
    uint32_t max_units;
};

static TiledWs tiled_ws(const SlmParams* p, const TileCfg& c) {
    TiledWs w;
    const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
    const uint32_t d = p->subpixel_seed_point_density ? p->subpixel_seed_point_density : 1u;
    const size_t n_work = (size_t)n * d * d;
    w.max_units = (uint32_t)((n_work + c.unit - 1) / c.unit + c.n_tiles);
    w.off_acc = align256(sizeof(TileCtrl));
    w.off_quad = w.off_acc + align256((size_t)p->nxy_padded * 8);
    w.off_a = w.off_quad + align256((size_t)p->nxy_padded * sizeof(Quad));
    w.off_b = w.off_a + align256(n_work * sizeof(TState));
    w.off_cnt = w.off_b + align256(n_work * sizeof(TState));
    w.off_off = w.off_cnt + align256((size_t)c.n_tiles * 4);
    w.off_cur = w.off_off + align256((size_t)c.n_tiles * 4);
    w.off_uoff = w.off_cur + align256((size_t)c.n_tiles * 4);
    w.off_units = w.off_uoff + align256((size_t)c.n_tiles * 4);
    w.total = w.off_units + align256((size_t)w.max_units * sizeof(TUnit));
    return w;
}

extern "C" size_t slm_trace_fields_tiled_workspace_size(const SlmParams* p, const SlmTileOpts* opts) {
    if (check_params(p) != SLM_OK) return 0;
    SlmTileOpts o = {};
    if (opts) o = *opts;
    tile_opts_defaults(&o);
    if (check_tile_opts(&o) != SLM_OK) return 0;
    return tiled_ws(p, tile_cfg(p, o)).total;
}

template <bool C>
static hipError_t tl_occupancy(int* per_cu, uint32_t block, size_t lds) {
    if (block == 1024) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_tl_process<C, 1024>, 1024, lds);
    if (block == 512) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_tl_process<C, 512>, 512, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_tl_process<C, 256>, 256, lds);
}

template <bool C>
static void tl_launch(uint32_t block, uint32_t grid, size_t lds, hipStream_t s, const SlmParams& P, const TileCfg& c,
                      const TState* in, TState* out, const TUnit* units, const uint8_t* mask, const float2* uv,
                      unsigned long long* acc, unsigned long long* steps, TileCtrl* ctrl, uint32_t fixed,
                      uint32_t cap) {
    if (block == 1024)
        hipLaunchKernelGGL((k_tl_process<C, 1024>), dim3(grid), dim3(1024), lds, s, P, c, in, out, units, mask, uv,
                           acc, steps, ctrl, fixed, cap);
    else if (block == 512)
        hipLaunchKernelGGL((k_tl_process<C, 512>), dim3(grid), dim3(512), lds, s, P, c, in, out, units, mask, uv,
                           acc, steps, ctrl, fixed, cap);
    else
        hipLaunchKernelGGL((k_tl_process<C, 256>), dim3(grid), dim3(256), lds, s, P, c, in, out, units, mask, uv,
                           acc, steps, ctrl, fixed, cap);
}

extern "C" int slm_trace_fields_tiled(const SlmParams* p, const SlmTileOpts* opts, const float* seeds_xy,
                                      const uint8_t* mask, const float* uv, uint32_t* slc, uint32_t* slt,
                                      uint64_t* steps_out, void* workspace, size_t workspace_bytes,
                                      void* stream, uint32_t flags) {
    int rc = check_params(p);
    if (rc) return rc;
    SlmTileOpts o = {};
    if (opts) o = *opts;
    tile_opts_defaults(&o);
    rc = check_tile_opts(&o);
    if (rc) return rc;
    if (!mask || !uv || !slc || !slt) return slm_fail(SLM_E_INVALID_ARG, "NULL mask/uv/slc/slt");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
    const bool accumulate = (flags & SLM_FLAG_NO_ZERO) != 0;
    if (n == 0) {
        if (!accumulate) {
            fill_u32(slc, p->nxy_padded, 0u, s);
            fill_u32(slt, p->nxy_padded, 0u, s);
        }
        return slm_finish(stream, flags);
    }
    if (!seeds_xy) return slm_fail(SLM_E_INVALID_ARG, "NULL seeds");
    const uint32_t d = p->subpixel_seed_point_density ? p->subpixel_seed_point_density : 1u;
    if ((uint64_t)n * d * d >= (1ull << 31))
        return slm_fail(SLM_E_INVALID_ARG, "n_seeds * d^2 = %llu exceeds 2^31", (unsigned long long)n * d * d);
    const TileCfg c = tile_cfg(p, o);
    const TiledWs w = tiled_ws(p, c);
    if (!workspace || workspace_bytes < w.total)
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, w.total);
    DevInfo di;
    rc = device_info(di);
    if (rc) return rc;
    const size_t lds = tile_lds_bytes(c);
    if (lds > (size_t)di.lds_per_block)
        return slm_fail(SLM_E_INVALID_ARG, "tile %u needs %zu bytes of LDS, the device allows %d per block", o.tile,
                        lds, di.lds_per_block);

    char* ws = (char*)workspace;
    TileCtrl* ctrl = (TileCtrl*)ws;
    unsigned long long* acc = (unsigned long long*)(ws + w.off_acc);
    Quad* quad = (Quad*)(ws + w.off_quad);
    TState* A = (TState*)(ws + w.off_a);
    TState* B = (TState*)(ws + w.off_b);
    uint32_t* cnt = (uint32_t*)(ws + w.off_cnt);
    uint32_t* off = (uint32_t*)(ws + w.off_off);
    uint32_t* cur = (uint32_t*)(ws + w.off_cur);
    uint32_t* uoff = (uint32_t*)(ws + w.off_uoff);
    TUnit* units = (TUnit*)(ws + w.off_units);
    const size_t nxy = p->nxy_padded;
    const uint32_t n_work = n * d * d;
    const uint32_t force = (flags & SLM_FLAG_FORCE_FALLBACK) ? 1u : 0u;
    const uint32_t fixed = dt_is_fixed(p);
    const uint32_t use_lds_hist = (c.n_tiles <= 8192u) ? 1u : 0u;
    const size_t hist_lds = use_lds_hist ? (size_t)c.n_tiles * 4 : 0;

    hipLaunchKernelGGL(k_fields_init, dim3(1), dim3(64), 0, s, &ctrl->f);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nxy)), dim3(256), 0, s, acc, nxy);
    hipLaunchKernelGGL(k_build_quads, dim3(grid_for(nxy)), dim3(256), 0, s, *p, (const float2*)uv, quad);
    if (p->fp_contract)
        hipLaunchKernelGGL(k_tl_fresh<true>, dim3(grid_for(n_work)), dim3(256), 0, s, *p, c,
                           (const float2*)seeds_xy, n_work, A);
    else
        hipLaunchKernelGGL(k_tl_fresh<false>, dim3(grid_for(n_work)), dim3(256), 0, s, *p, c,
                           (const float2*)seeds_xy, n_work, A);

    int per_cu = 0;
    if (p->fp_contract)
        HIP_TRY(tl_occupancy<true>(&per_cu, o.block, lds));
    else
        HIP_TRY(tl_occupancy<false>(&per_cu, o.block, lds));
    if (per_cu < 1) per_cu = 1;
    const uint32_t pgrid = (uint32_t)di.cus * (uint32_t)per_cu;

    uint32_t* h_nout = nullptr;
    HIP_TRY(hipHostMalloc((void**)&h_nout, sizeof(uint32_t), hipHostMallocDefault));
    uint32_t n_in = n_work;
    int rounds = 0;
    hipError_t herr = hipSuccess;
    // every visit accepts at least one step or ends the streamline, so the
    // number of rounds is bounded by max_n_steps + 1
    const int max_rounds = (int)p->max_n_steps + 2;
    while (n_in > 0 && rounds < max_rounds) {
        const uint32_t g = grid_for(n_in) < 1024u ? grid_for(n_in) : 1024u;
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(c.n_tiles)), dim3(256), 0, s, cnt, (size_t)c.n_tiles, 0u);
        hipLaunchKernelGGL(k_tl_count, dim3(g), dim3(256), hist_lds, s, A, n_in, c.n_tiles, cnt, use_lds_hist);
        hipLaunchKernelGGL(k_tl_scan, dim3(1), dim3(1024), 0, s, cnt, off, cur, uoff, c.n_tiles, c.unit, ctrl);
        hipLaunchKernelGGL(k_tl_scatter, dim3(g), dim3(256), 2 * hist_lds, s, A, n_in, c.n_tiles, cur, B,
                           use_lds_hist);
        hipLaunchKernelGGL(k_tl_units, dim3(grid_for(c.n_tiles)), dim3(256), 0, s, cnt, off, uoff, c.n_tiles,
                           c.unit, units);
        if (p->fp_contract)
            tl_launch<true>(o.block, pgrid, lds, s, *p, c, B, A, units, mask, (const float2*)uv, acc,
                            (unsigned long long*)steps_out, ctrl, fixed, n_work);
        else
            tl_launch<false>(o.block, pgrid, lds, s, *p, c, B, A, units, mask, (const float2*)uv, acc,
                             (unsigned long long*)steps_out, ctrl, fixed, n_work);
        herr = hipMemcpyAsync(h_nout, &ctrl->n_out, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        if (herr == hipSuccess) herr = hipStreamSynchronize(s);
        if (herr != hipSuccess) break;
        n_in = *h_nout;
        if (n_in > n_work) {
            (void)hipHostFree(h_nout);
            return slm_fail(SLM_E_HIP, "tiled round %d emitted %u states, more than the %u streamlines", rounds,
                            n_in, n_work);
        }
        rounds++;
    }
    (void)hipHostFree(h_nout);
    if (herr != hipSuccess) return slm_fail(SLM_E_HIP, "tiled round %d failed: %s", rounds, hipGetErrorString(herr));
    if (n_in > 0) return slm_fail(SLM_E_HIP, "tiled tracer did not finish in %d rounds", rounds);

    hipLaunchKernelGGL(k_fields_sumlow, dim3(grid_for(nxy)), dim3(256), 0, s, acc, nxy, &ctrl->f);
    hipLaunchKernelGGL(k_fields_commit, dim3(grid_for(nxy)), dim3(256), 0, s, acc, slc, slt, nxy, &ctrl->f,
                       accumulate ? 1u : 0u, force);
    // fallback: the queue tracer with 32-bit atomics, identity order; exits at once when committed
    int fb_cu = 0;
    if (p->fp_contract)
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&fb_cu, k_fields<true, false, false>, 256, 0));
    else
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&fb_cu, k_fields<false, false, false>, 256, 0));
    if (fb_cu < 1) fb_cu = 1;
    uint32_t fgrid = (uint32_t)di.cus * (uint32_t)fb_cu;
    const uint32_t max_useful = (n_work + 255) / 256;
    if (fgrid > max_useful) fgrid = max_useful ? max_useful : 1;
    const float2* seeds = (const float2*)seeds_xy;
    if (p->fp_contract)
        hipLaunchKernelGGL((k_fields<true, false, false>), dim3(fgrid), dim3(256), 0, s, *p, seeds, nullptr, n_work, mask,
                           quad, acc, slc, slt, nullptr, &ctrl->f, force, fixed);
    else
        hipLaunchKernelGGL((k_fields<false, false, false>), dim3(fgrid), dim3(256), 0, s, *p, seeds, nullptr, n_work,
                           mask, quad, acc, slc, slt, nullptr, &ctrl->f, force, fixed);
    return slm_finish(stream, flags);
}

// ------------------------------------------------------------------ batched fields tracer
//
// Many tiles (grids of any size, each with its own macro set: DT_MAX,
// COMBO_FACTOR and direction differ per tile) traced by ONE persistent
// launch.  The work list is the concatenation of every job's Morton-ordered
// work items; waves take wave-private chunks from XCD-sliced queues exactly
// as k_fields does, so a lane can hold a streamline of any job.  At each
// refill a lane finds its job by binary search over the jobs' work offsets
// and loads that job's constants and pointers into registers.  All jobs
// accumulate into one concatenated acc with the same packed-sum check and
// fallback as a single tile.  No host synchronisation and no memset nodes:
// the whole call can be captured in a hipGraph.

struct BatchDesc {
    SlmParams P;
    const float2* seeds;
    const uint8_t* mask;
    const Quad* quad;
    unsigned long long* acc;
    uint32_t* slc;
    uint32_t* slt;
    const SeedRec* order;
    uint32_t work_begin, n_work;
    uint32_t pix_begin, dt_fixed;
    uint32_t seed_begin, n_seeds;   // this job's slice of the concatenated seed records
    uint32_t key_begin, n_keys;     // and of the concatenated Morton-tile counters
};

constexpr uint32_t kDescPerLaunch = 8;

// Binary search: the job whose [begin, next begin) holds i, for the member
// `field` (jobs are stored in increasing order of every begin).
#define BATCH_FIND(desc, n_jobs, field, i, out)                    \
    do {                                                           \
        uint32_t lo_ = 0, hi_ = (n_jobs);                          \
        while (hi_ - lo_ > 1u) {                                   \
            const uint32_t mid_ = (lo_ + hi_) >> 1;                \
            if ((desc)[mid_].field <= (i)) lo_ = mid_; else hi_ = mid_; \
        }                                                          \
        (out) = lo_;                                               \
    } while (0)

// BATCH_FIND for item i of a wave whose items run on from i0 (the wave's
// first item, the same in every lane): the job of i0 by one search with
// scalar loads, and a lane's own search only when its item lies at or past
// the next job's begin (a wave across a job boundary).
#define BATCH_FIND_WAVE(desc, n_jobs, field, i0, i, out)                                 \
    do {                                                                                 \
        const uint32_t w0_ = (uint32_t)__builtin_amdgcn_readfirstlane((int)(i0));        \
        uint32_t j0_;                                                                    \
        BATCH_FIND(desc, n_jobs, field, w0_, j0_);                                       \
        const uint32_t e_ = j0_ + 1u < (n_jobs) ? (desc)[j0_ + 1u].field : 0xFFFFFFFFu;  \
        (out) = j0_;                                                                     \
        if ((i) >= e_) BATCH_FIND(desc, n_jobs, field, i, out);                          \
    } while (0)

// The seed sort and quad build of every job, one launch each for all jobs.
__global__ void k_batch_count(const BatchDesc* __restrict__ desc, uint32_t n_jobs, uint32_t n_seeds,
                              uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_seeds) return;
    uint32_t j;
    BATCH_FIND_WAVE(desc, n_jobs, seed_begin, i - (threadIdx.x & 63u), i, j);
    const BatchDesc& D = desc[j];
    const uint32_t k = i - D.seed_begin;
    atomicAdd(&cnt[D.key_begin + seed_tile_key(D.P, D.seeds[D.P.seeds_chunk_offset + k])], 1u);
}

// Exclusive scan of each job's counters: one 1024-thread block per job.
__global__ __launch_bounds__(1024) void k_batch_scan(const BatchDesc* __restrict__ desc, uint32_t* __restrict__ cnt,
                                                     uint32_t* __restrict__ cur) {
    const BatchDesc& D = desc[blockIdx.x];
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x, n = D.n_keys;
    uint32_t* c = cnt + D.key_begin;
    uint32_t* u = cur + D.key_begin;
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t b = t * per, e = min(n, b + per);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; i++) s += c[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = (t == 0) ? 0u : part[t - 1];
    for (uint32_t i = b; i < e; i++) {
        const uint32_t x = c[i];
        c[i] = run;
        u[i] = run;
        run += x;
    }
}

__global__ void k_batch_scatter(const BatchDesc* __restrict__ desc, uint32_t n_jobs, uint32_t n_seeds,
                                uint32_t* __restrict__ cur, SeedRec* __restrict__ order_all) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_seeds) return;
    uint32_t j;
    BATCH_FIND_WAVE(desc, n_jobs, seed_begin, i - (threadIdx.x & 63u), i, j);
    const BatchDesc& D = desc[j];
    const uint32_t k = i - D.seed_begin;
    const float2 sd = D.seeds[D.P.seeds_chunk_offset + k];
    SeedRec r;
    r.x = sd.x;
    r.y = sd.y;
    r.seed_idx = D.P.seeds_chunk_offset + k;
    r.pad = 0u;
    order_all[D.seed_begin + atomicAdd(&cur[D.key_begin + seed_tile_key(D.P, sd)], 1u)] = r;
}

__global__ void k_batch_quads(const BatchDesc* __restrict__ desc, uint32_t n_jobs, size_t n_pix,
                              const float2* const* __restrict__ uvs, Quad* __restrict__ quad_all) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_pix; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t j;
        BATCH_FIND_WAVE(desc, n_jobs, pix_begin, (uint32_t)i - (threadIdx.x & 63u), (uint32_t)i, j);
        const SlmParams& P = desc[j].P;
        const float2* uv = uvs[j];
        const size_t li = i - desc[j].pix_begin;
        const uint32_t x = (uint32_t)(li / P.ny_padded), y = (uint32_t)(li - (size_t)x * P.ny_padded);
        const uint32_t xr = min(P.nx_padded - 1u, x + 1u), yu = min(P.ny_padded - 1u, y + 1u);
        const float2 a = uv[(size_t)P.ny_padded * x + y], b = uv[(size_t)P.ny_padded * xr + y];
        const float2 e = uv[(size_t)P.ny_padded * x + yu], f = uv[(size_t)P.ny_padded * xr + yu];
        Quad q;
        q.lo = make_float4(a.x, a.y, b.x, b.y);
        q.hi = make_float4(e.x, e.y, f.x, f.y);
        quad_all[i] = q;
    }
}

struct UvPack {
    const float2* uv[kDescPerLaunch];
};
__global__ void k_put_uv(UvPack pack, uint32_t first, uint32_t count, const float2** __restrict__ out) {
    const uint32_t i = threadIdx.x;
    if (i < count) out[first + i] = pack.uv[i];
}
struct DescPack {
    BatchDesc d[kDescPerLaunch];
};

__global__ void k_put_desc(DescPack pack, uint32_t first, uint32_t count, BatchDesc* __restrict__ out) {
    const uint32_t i = threadIdx.x;
    if (i < count) out[first + i] = pack.d[i];
}

template <bool C, bool PACKED, bool R2, uint32_t SLOTS = kRunSlots>
__global__ __launch_bounds__(256) void k_fields_batch(const BatchDesc* __restrict__ desc, uint32_t n_jobs,
                                                      uint32_t n_work, unsigned long long* __restrict__ steps,
                                                      FieldsCtrl* __restrict__ ctrl, uint32_t force_fallback,
                                                      uint32_t chunk, uint32_t n_slices, uint32_t batch,
                                                      unsigned long long* __restrict__ acc_all) {
    if constexpr (!PACKED) {
        if (fields_packed_ok(ctrl, force_fallback)) return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64u - lane));

    // per-wave accumulation table; its tags are pixel indices into the
    // concatenated accumulator of all jobs (fewer than 2^32 - 1 pixels)
    __shared__ uint32_t s_tag[PACKED ? 4u * SLOTS : 1u];
    __shared__ unsigned long long s_val[PACKED ? 4u * SLOTS : 1u];
    RunTableT<SLOTS> table{s_tag + (threadIdx.x >> 6) * SLOTS, s_val + (threadIdx.x >> 6) * SLOTS};
    if constexpr (PACKED) table.init(lane);

    // per-lane job context
    SlmParams P;
    GBytePtr mask = nullptr;
    GQuadPtr quad = nullptr;
    uint32_t* slc = nullptr;
    uint32_t* slt = nullptr;
    uint32_t dt_fixed = 0, pix_begin = 0;
    float inv_res = 0.f;   // 1/TRAJECTORY_RESOLUTION of the lane's job (exact when R2)

    bool active = false, exhausted = false;
    float x = 0.f, y = 0.f, l = 0.f, dt = 0.f;
    uint32_t n = 0, rng = 0;
    Quad Q;
    uint32_t cx = 0, cy = 0;   // pixel of the current position
    uint32_t run_idx = 0, run_slot = 0, run_n = 0, run_t = 0;
    uint64_t my_steps = 0, my_writes = 0;

    // adds the runs of the lanes with has = true (all lanes call it together);
    // a is the job-local pixel index
    auto emit = [&](bool has, uint32_t a, uint32_t s, uint32_t cn, uint32_t ct) {
        if constexpr (PACKED) {
            table.add(has, pix_begin + a, s, ((unsigned long long)ct << 32) | (unsigned long long)cn, acc_all);
        } else {
            if (has) {
                atomicAdd(&slt[a], ct);
                atomicAdd(&slc[a], cn);
            }
        }
    };

    uint32_t c_next = 0, c_end = 0;
    const uint32_t slice = (n_work + n_slices - 1u) / n_slices;
    while (true) {
        if (!exhausted) {
            uint64_t need = __ballot(!active);
            if (__popcll(need) < batch && __any(active)) need = 0;   // batched refills
            while (need) {
                if (c_next >= c_end) {
                    uint32_t base = 0xffffffffu, lim = 0;
                    const uint32_t me = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) % n_slices;
                    for (uint32_t t = 0; t < n_slices && base == 0xffffffffu; t++) {
                        const uint32_t q = (me + t) % n_slices;
                        const uint32_t lo = q * slice, hi = min(n_work, lo + slice);
                        uint32_t b = 0;
                        if (lane == 0) b = atomicAdd(&ctrl->queue[q * kQueueStride], chunk);
                        b = __builtin_amdgcn_readfirstlane(b);   // lane 0's, wave-uniform
                        if (b < slice && lo + b < hi) {
                            base = lo + b;
                            lim = hi;
                        }
                    }
                    if (base == 0xffffffffu) {
                        exhausted = true;
                        break;
                    }
                    c_next = base;
                    c_end = min(base + chunk, lim);
                }
                const uint32_t take = min((uint32_t)__popcll(need), c_end - c_next);
                const uint32_t rank = __popcll(need & lt_mask);
                const bool mine = ((need >> lane) & 1ull) && rank < take;
                // the job holding work item c_next, found once for the wave
                // (c_next is uniform: scalar loads); items before the next
                // job's begin belong to it too, and only a lane whose item
                // lies beyond (a chunk across a job boundary) searches again
                uint32_t j0;
                BATCH_FIND(desc, n_jobs, work_begin, c_next, j0);
                const uint32_t j0_end = j0 + 1u < n_jobs ? desc[j0 + 1u].work_begin : 0xffffffffu;
                if (mine) {
                    const uint32_t w = c_next + rank;
                    uint32_t lo = j0;
                    if (w >= j0_end) BATCH_FIND(desc, n_jobs, work_begin, w, lo);
                    const BatchDesc* D = desc + lo;
                    P = D->P;
                    mask = as_global(D->mask);
                    quad = as_global(D->quad);
                    slc = D->slc;
                    slt = D->slt;
                    dt_fixed = D->dt_fixed;
                    pix_begin = D->pix_begin;
                    inv_res = 1.0f / P.trajectory_resolution;
                    const uint32_t d = P.subpixel_seed_point_density ? P.subpixel_seed_point_density : 1u;
                    const uint32_t wl = w - D->work_begin;
                    // work item wl = seed * d^2 + sub, sub = j * d + i (i along x);
                    // d = 1 and 2 without the integer divisions
                    uint32_t seed, si, sj;
                    if (d == 1u) {
                        seed = wl;
                        si = sj = 0;
                    } else if (d == 2u) {
                        seed = wl >> 2;
                        si = wl & 1u;
                        sj = (wl >> 1) & 1u;
                    } else {
                        const uint32_t dd = d * d, sub = wl % dd;
                        seed = wl / dd;
                        si = sub % d;
                        sj = sub / d;
                    }
                    const SeedRec r = D->order[seed];
                    x = r.x + sub_offset<C>(P, si);
                    y = r.y + sub_offset<C>(P, sj);
                    rng = rng_seed(r.seed_idx);
                    l = 0.f;
                    n = 0;
                    dt = P.dt_max;
                    const uint32_t ix = pix(x, P.pad_width_pp5, P.nx_padded - 1u);
                    const uint32_t iy = pix(y, P.pad_width_pp5, P.ny_padded - 1u);
                    const uint32_t idx = iy + P.ny_padded * ix;
                    // the seed pixel's quad is loaded with its mask byte (a
                    // masked seed leaves the lane idle, and Q is set again at
                    // its next refill)
                    const uint8_t mk = mask[idx];
                    Q = load_quad(quad, idx);
                    if (!mk) {
                        // atomic_write_sl_data at the seed (jittertrajectory.cl:60-62)
                        cx = ix;
                        cy = iy;
                        run_idx = idx;
                        run_slot = slot_of_t<SLOTS>(ix, iy);
                        run_n = 1u;
                        run_t = slt_inc(0.f);
                        active = (0.f < P.max_length) && (0u < P.max_n_steps);
                    }
                }
                need &= ~__ballot(mine);
                c_next += take;
            }
        }
        bool f_has = false, stepped = false;
        uint32_t nidx = 0, nix = 0, niy = 0;
        if (active) {
            // compute_step_vec_jittered (computestep.cl:86-106) on quads, with
            // the short sqrt / divide forms where their inputs allow
            const float res = P.trajectory_resolution;
            const float combo = P.combo_factor;
            const float2 u1 = jitter_unit_cr<C>(P, interp_quad_px_cr<C>(P, Q, x, y, cx, cy), rng);
            const float2 d1 = make_float2(approx_r<R2>(u1.x * combo * dt, res, inv_res), approx_r<R2>(u1.y * combo * dt, res, inv_res));
            const float mx = x + d1.x, my = y + d1.y;
            const uint32_t mix = pix(mx, P.pad_width_pp5, P.nx_padded - 1u);
            const uint32_t miy = pix(my, P.pad_width_pp5, P.ny_padded - 1u);
            const Quad Q1 = load_quad(quad, miy + P.ny_padded * mix);
            const float2 u2 = jitter_unit_cr<C>(P, interp_quad_px_cr<C>(P, Q1, mx, my, mix, miy), rng);
            const float2 d2 = make_float2(approx_r<R2>(0.5f * mad<C>(u2.x * combo, dt, d1.x), res, inv_res),
                                          approx_r<R2>(0.5f * mad<C>(u2.y * combo, dt, d1.y), res, inv_res));
            const float nx = x + d2.x, ny = y + d2.y;
            nix = pix(nx, P.pad_width_pp5, P.nx_padded - 1u);
            niy = pix(ny, P.pad_width_pp5, P.ny_padded - 1u);
            nidx = niy + P.ny_padded * nix;
            // the end pixel's mask byte and quad, issued together: Q is taken
            // whatever the mask says (a masked end pixel ends the streamline,
            // and Q is not read again before a refill sets it), so the quad
            // load is not deferred behind the mask test
            const uint8_t mk = mask[nidx];
            Q = load_quad(quad, nidx);
            // the position and its pixel advance whatever the mask says, for
            // the same reason
            x = nx;
            y = ny;
            cx = nix;
            cy = niy;
            if (mk) {
                active = false;
            } else {
                const float dl = fast_length_cr<C>(d2.x, d2.y);
                const bool stuck = dl < P.integration_halt_threshold;
                if (!stuck && !dt_fixed) {
                    StepOut o;
                    o.dxy1 = d1;
                    o.dxy2 = d2;
                    dt = next_dt(P, dt, step_error<C>(P, o));
                }
                l += dl;
                n += 1u;
                stepped = true;
                f_has = nidx != run_idx;
                active = !stuck && (l < P.max_length) && (n < P.max_n_steps);
            }
        }
        // a step into another pixel closes the run of the one it left; the
        // run state is passed as it stands, and the new run starts after
        emit(f_has, run_idx, run_slot, run_n, run_t);
        if (f_has) {
            run_idx = nidx;
            run_slot = slot_of_t<SLOTS>(nix, niy);
            run_n = 0;
            run_t = 0;
        }
        if (stepped) {
            run_n += 1u;
            run_t += slt_inc(l);
        }
        const bool ended = !active && run_n;
        if (__any(ended)) {
            emit(ended, run_idx, run_slot, run_n, run_t);
            if (ended) {
                // the streamline's writes: its seed and n accepted steps
                my_steps += n;
                my_writes += (uint64_t)n + 1u;
                run_n = 0;
                run_t = 0;
            }
        }
        if (exhausted && !__any(active)) break;
    }
    if constexpr (PACKED) {
        table.drain(lane, acc_all);
        uint64_t v = my_steps, wv = my_writes;
        for (int off = 32; off > 0; off >>= 1) {
            v += __shfl_down(v, off);
            wv += __shfl_down(wv, off);
        }
        if (lane == 0) {
            if (steps && v) atomicAdd(steps, (unsigned long long)v);
            if (wv) atomicAdd(&ctrl->writes, (unsigned long long)wv);
        }
    }
}

// Split the concatenated acc into every job's slc/slt (or clear them for the
// fallback), as k_fields_commit does for one grid.
__global__ void k_batch_commit(const BatchDesc* __restrict__ desc, uint32_t n_jobs, size_t n_pix,
                               FieldsCtrl* __restrict__ ctrl, uint32_t accumulate, uint32_t force_fallback,
                               const unsigned long long* __restrict__ acc_all) {
    const bool ok = fields_packed_ok(ctrl, force_fallback);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_pix; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t lo;
        BATCH_FIND_WAVE(desc, n_jobs, pix_begin, (uint32_t)i - (threadIdx.x & 63u), (uint32_t)i, lo);
        const BatchDesc* D = desc + lo;
        const size_t j = i - D->pix_begin;
        if (ok) {
            const unsigned long long a = acc_all[i];
            const uint32_t l32 = (uint32_t)a, h32 = (uint32_t)(a >> 32);
            if (accumulate) {
                D->slc[j] += l32;
                D->slt[j] += h32;
            } else {
                D->slc[j] = l32;
                D->slt[j] = h32;
            }
        } else if (!accumulate) {
            D->slc[j] = 0u;
            D->slt[j] = 0u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < kQueues) ctrl->queue[threadIdx.x * kQueueStride] = 0u;
}

struct BatchWs {
    size_t off_desc, off_uvs, off_cnt, off_cur, off_order, off_acc, off_quad, total;
    uint64_t n_keys, n_seeds, n_pix, n_work;
};

static int batch_ws(const SlmFieldsJob* jobs, uint32_t n_jobs, BatchWs& w) {
    w.n_keys = w.n_seeds = w.n_pix = w.n_work = 0;
    for (uint32_t j = 0; j < n_jobs; j++) {
        const SlmParams* p = &jobs[j].params;
        int rc = check_params(p);
        if (rc) return rc;
        const TileGeom g = tile_geom(p);
        w.n_keys += g.n_keys;
        const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
        const uint32_t d = p->subpixel_seed_point_density ? p->subpixel_seed_point_density : 1u;
        w.n_seeds += n;
        w.n_pix += p->nxy_padded;
        w.n_work += (uint64_t)n * d * d;
    }
    if (w.n_work >= (1ull << 32) - 4096 || w.n_pix >= (1ull << 32) || w.n_keys >= (1ull << 32) ||
        w.n_seeds >= (1ull << 32))
        return slm_fail(SLM_E_INVALID_ARG, "batch too large: %llu work items, %llu pixels",
                        (unsigned long long)w.n_work, (unsigned long long)w.n_pix);
    w.off_desc = align256(sizeof(FieldsCtrl)) > 4096 ? align256(sizeof(FieldsCtrl)) : 4096;
    w.off_uvs = w.off_desc + align256((size_t)n_jobs * sizeof(BatchDesc));
    w.off_cnt = w.off_uvs + align256((size_t)n_jobs * sizeof(void*));
    w.off_cur = w.off_cnt + align256((size_t)w.n_keys * 4);
    w.off_order = w.off_cur + align256((size_t)w.n_keys * 4);
    w.off_acc = w.off_order + align256((size_t)w.n_seeds * sizeof(SeedRec));
    w.off_quad = w.off_acc + align256((size_t)w.n_pix * 8);
    w.total = w.off_quad + align256((size_t)w.n_pix * sizeof(Quad));
    return SLM_OK;
}

extern "C" size_t slm_trace_fields_batch_workspace_size(const SlmFieldsJob* jobs, uint32_t n_jobs) {
    if (!jobs || n_jobs == 0) return 0;
    BatchWs w;
    if (batch_ws(jobs, n_jobs, w) != SLM_OK) return 0;
    return w.total;
}

extern "C" int slm_trace_fields_batch(const SlmFieldsJob* jobs, uint32_t n_jobs, const SlmBatchOpts* opts,
                                      uint64_t* steps_out, void* workspace, size_t workspace_bytes, void* stream,
                                      uint32_t flags) {
    if (!jobs || n_jobs == 0) return slm_fail(SLM_E_INVALID_ARG, "empty batch");
    BatchWs w;
    int rc = batch_ws(jobs, n_jobs, w);
    if (rc) return rc;
    if (!workspace || workspace_bytes < w.total)
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, w.total);
    const uint32_t C0 = jobs[0].params.fp_contract ? 1u : 0u;
    for (uint32_t j = 0; j < n_jobs; j++) {
        const SlmFieldsJob& J = jobs[j];
        if ((J.params.fp_contract ? 1u : 0u) != C0)
            return slm_fail(SLM_E_INVALID_ARG, "job %u: fp_contract differs within the batch", j);
        if (!J.mask || !J.uv || !J.slc || !J.slt) return slm_fail(SLM_E_INVALID_ARG, "job %u: NULL buffer", j);
        if (J.params.n_seed_points > J.params.seeds_chunk_offset && !J.seeds_xy)
            return slm_fail(SLM_E_INVALID_ARG, "job %u: NULL seeds", j);
    }
    uint32_t chunk = 64, blocks_per_cu = 0, slices = kQueues, slots = kRunSlots;
    // the refill batch of the densest job (fields_refill_batch)
    uint32_t max_d = 1;
    for (uint32_t j = 0; j < n_jobs; j++) max_d = std::max(max_d, jobs[j].params.subpixel_seed_point_density);
    uint32_t batch = fields_refill_batch(max_d);
    if (opts) {
        if (opts->chunk) chunk = opts->chunk;
        if (opts->blocks_per_cu) blocks_per_cu = opts->blocks_per_cu;
        if (opts->queue_slices) slices = opts->queue_slices;
        if (opts->lds_slots) slots = opts->lds_slots;
        if (opts->refill_batch) batch = opts->refill_batch;
    }
    if (batch > 64) return slm_fail(SLM_E_INVALID_ARG, "refill_batch %u outside [1, 64]", batch);
    if (chunk < 1 || chunk > 65536) return slm_fail(SLM_E_INVALID_ARG, "chunk %u outside [1, 65536]", chunk);
    if (slices < 1 || slices > kQueues) return slm_fail(SLM_E_INVALID_ARG, "queue_slices %u outside [1, 8]", slices);
    if (slots != 256 && slots != 512 && slots != 1024)
        return slm_fail(SLM_E_INVALID_ARG, "lds_slots %u is not 256, 512 or 1024", slots);
    DevInfo di;
    rc = device_info(di);
    if (rc) return rc;

    hipStream_t s = (hipStream_t)stream;
    const bool accumulate = (flags & SLM_FLAG_NO_ZERO) != 0;
    const uint32_t force = (flags & SLM_FLAG_FORCE_FALLBACK) ? 1u : 0u;
    char* ws = (char*)workspace;
    FieldsCtrl* ctrl = (FieldsCtrl*)ws;
    BatchDesc* desc = (BatchDesc*)(ws + w.off_desc);
    uint32_t* cnt = (uint32_t*)(ws + w.off_cnt);
    uint32_t* cur = (uint32_t*)(ws + w.off_cur);
    SeedRec* order_all = (SeedRec*)(ws + w.off_order);
    unsigned long long* acc_all = (unsigned long long*)(ws + w.off_acc);
    Quad* quad_all = (Quad*)(ws + w.off_quad);

    const float2** uvs = (const float2**)(ws + w.off_uvs);
    // descriptors (and the uv pointer table) through kernel arguments, so the
    // call stays capturable; then the seed sort and quads of all jobs at once
    DescPack pack;
    UvPack upack;
    uint32_t in_pack = 0, pack_first = 0;
    uint64_t work_off = 0, seed_off = 0, pix_off = 0, key_off = 0;
    for (uint32_t j = 0; j < n_jobs; j++) {
        const SlmFieldsJob& J = jobs[j];
        const SlmParams* p = &J.params;
        const uint32_t n = p->n_seed_points - p->seeds_chunk_offset;
        const uint32_t d = p->subpixel_seed_point_density ? p->subpixel_seed_point_density : 1u;
        const TileGeom g = tile_geom(p);
        BatchDesc& D = pack.d[in_pack];
        D.P = *p;
        D.seeds = (const float2*)J.seeds_xy;
        D.mask = J.mask;
        D.quad = quad_all + pix_off;
        D.acc = acc_all + pix_off;
        D.slc = J.slc;
        D.slt = J.slt;
        D.order = order_all + seed_off;
        D.work_begin = (uint32_t)work_off;
        D.n_work = n * d * d;
        D.pix_begin = (uint32_t)pix_off;
        D.dt_fixed = dt_is_fixed(p);
        D.seed_begin = (uint32_t)seed_off;
        D.n_seeds = n;
        D.key_begin = (uint32_t)key_off;
        D.n_keys = g.n_keys;
        upack.uv[in_pack] = (const float2*)J.uv;
        work_off += (uint64_t)n * d * d;
        seed_off += n;
        pix_off += p->nxy_padded;
        key_off += g.n_keys;
        if (++in_pack == kDescPerLaunch || j + 1 == n_jobs) {
            hipLaunchKernelGGL(k_put_desc, dim3(1), dim3(64), 0, s, pack, pack_first, in_pack, desc);
            hipLaunchKernelGGL(k_put_uv, dim3(1), dim3(64), 0, s, upack, pack_first, in_pack, uvs);
            pack_first += in_pack;
            in_pack = 0;
        }
    }
    if (!(flags & SLM_FLAG_KEEP_ORDER)) {
        const uint32_t ns = (uint32_t)w.n_seeds;
        fill_u32(cnt, (size_t)w.n_keys, 0u, s);
        if (ns) {
            hipLaunchKernelGGL(k_batch_count, dim3((ns + 255) / 256), dim3(256), 0, s, (const BatchDesc*)desc, n_jobs,
                               ns, cnt);
            hipLaunchKernelGGL(k_batch_scan, dim3(n_jobs), dim3(1024), 0, s, (const BatchDesc*)desc, cnt, cur);
            hipLaunchKernelGGL(k_batch_scatter, dim3((ns + 255) / 256), dim3(256), 0, s, (const BatchDesc*)desc,
                               n_jobs, ns, cur, order_all);
        }
        hipLaunchKernelGGL(k_batch_quads, dim3(grid_for(w.n_pix)), dim3(256), 0, s, (const BatchDesc*)desc, n_jobs,
                           (size_t)w.n_pix, (const float2* const*)uvs, quad_all);
    }
    hipLaunchKernelGGL(k_fields_init, dim3(1), dim3(64), 0, s, ctrl);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(w.n_pix)), dim3(256), 0, s, acc_all, (size_t)w.n_pix);

    const uint32_t n_work = (uint32_t)w.n_work;
    int per_cu = 0;
    bool r2 = true;
    for (uint32_t j = 0; j < n_jobs; j++) r2 = r2 && slm_res_is_pow2(&jobs[j].params);
#define SLM_OCC(CC, SS) HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fields_batch<CC, true, true, SS>, 256, 0))
    if (C0) {
        if (slots == 256) SLM_OCC(true, 256);
        else if (slots == 1024) SLM_OCC(true, 1024);
        else SLM_OCC(true, 512);
    } else {
        if (slots == 256) SLM_OCC(false, 256);
        else if (slots == 1024) SLM_OCC(false, 1024);
        else SLM_OCC(false, 512);
    }
#undef SLM_OCC
    if (per_cu < 1) per_cu = 1;
    if (blocks_per_cu && (int)blocks_per_cu < per_cu) per_cu = (int)blocks_per_cu;
    uint32_t grid = (uint32_t)di.cus * (uint32_t)per_cu;
    const uint32_t max_useful = (n_work + 255) / 256;
    if (grid > max_useful) grid = max_useful ? max_useful : 1;
    unsigned long long* st = (unsigned long long*)steps_out;
    if (n_work) {
#define SLM_LAUNCH_BATCH(CC, RR, SS)                                                                     \
    hipLaunchKernelGGL((k_fields_batch<CC, true, RR, SS>), dim3(grid), dim3(256), 0, s, (const BatchDesc*)desc, \
                       n_jobs, n_work, st, ctrl, force, chunk, slices, batch, acc_all)
#define SLM_LAUNCH_SLOTS(CC, RR)                          \
    do {                                                  \
        if (slots == 256) SLM_LAUNCH_BATCH(CC, RR, 256);   \
        else if (slots == 1024) SLM_LAUNCH_BATCH(CC, RR, 1024); \
        else SLM_LAUNCH_BATCH(CC, RR, 512);                \
    } while (0)
        if (C0) {
            if (r2) SLM_LAUNCH_SLOTS(true, true);
            else SLM_LAUNCH_SLOTS(true, false);
        } else {
            if (r2) SLM_LAUNCH_SLOTS(false, true);
            else SLM_LAUNCH_SLOTS(false, false);
        }
#undef SLM_LAUNCH_SLOTS
#undef SLM_LAUNCH_BATCH
    }
    hipLaunchKernelGGL(k_fields_sumlow, dim3(grid_for(w.n_pix)), dim3(256), 0, s, acc_all, (size_t)w.n_pix, ctrl);
    hipLaunchKernelGGL(k_batch_commit, dim3(grid_for(w.n_pix)), dim3(256), 0, s, (const BatchDesc*)desc, n_jobs,
                       (size_t)w.n_pix, ctrl, accumulate ? 1u : 0u, force, (const unsigned long long*)acc_all);
    if (n_work) {
        if (C0)
            hipLaunchKernelGGL((k_fields_batch<true, false, false>), dim3(grid), dim3(256), 0, s, (const BatchDesc*)desc,
                               n_jobs, n_work, nullptr, ctrl, force, chunk, slices, batch, acc_all);
        else
            hipLaunchKernelGGL((k_fields_batch<false, false, false>), dim3(grid), dim3(256), 0, s, (const BatchDesc*)desc,
                               n_jobs, n_work, nullptr, ctrl, force, chunk, slices, batch, acc_all);
    }
    return slm_finish(stream, flags);
}
