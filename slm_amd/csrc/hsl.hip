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
#include "../../include/slm_hip.h"
#include "common.hpp"

namespace {

typedef unsigned long long u64;
constexpr int kBlock = 256;
constexpr double kFix = 16777216.0;   // 2^24

__device__ __forceinline__ uint32_t abs_label(int32_t l) {
    // |label|; the most negative int32 (a flank bit with no label) is no label
    return (l == INT32_MIN) ? 0u : (uint32_t)(l < 0 ? -l : l);
}

// Label pairs whose dilated domains can share a pixel: pixels at Manhattan
// distance 1 or 2 with different non-zero |label|.  Half of the offsets, so
// each unordered pixel pair is seen once; duplicates of a label pair remain.
__global__ void k_fmm_pairs(const int32_t* __restrict__ label, uint32_t nx, uint32_t ny,
                            const uint8_t* __restrict__ colour, uint32_t n_labels, u64* __restrict__ pairs,
                            uint32_t cap, uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const uint32_t a = abs_label(label[i]);
    if (!a || a >= n_labels || colour[a] == 0xffu) return;
    const int x = (int)(i / ny), y = (int)(i % ny);
    const int off[6][2] = {{0, 1}, {0, 2}, {1, -1}, {1, 0}, {1, 1}, {2, 0}};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int xx = x + off[k][0], yy = y + off[k][1];
        if (xx < 0 || yy < 0 || xx >= (int)nx || yy >= (int)ny) continue;
        const uint32_t b = abs_label(label[(size_t)xx * ny + yy]);
        if (!b || b == a || b >= n_labels || colour[b] == 0xffu) continue;
        const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        const uint32_t slot = atomicAdd(count, 1u);
        if (slot < cap) pairs[slot] = ((u64)lo << 32) | hi;
    }
}

// problem key (|label| + 0 = none) and phi sign (1 = negative) of colour c
__global__ void k_fmm_setup(const int32_t* __restrict__ label, const uint32_t* __restrict__ mapping,
                            uint32_t thin_flag, const uint8_t* __restrict__ colour, uint32_t n_labels, uint32_t c,
                            uint32_t nx, uint32_t ny, uint32_t* __restrict__ dom, uint8_t* __restrict__ neg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
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

__global__ void k_fmm_init(const uint32_t* __restrict__ dom, const uint8_t* __restrict__ neg, uint32_t nx,
                           uint32_t ny, double dx, double* __restrict__ u, uint8_t* __restrict__ fixed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
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

__global__ __launch_bounds__(256) void k_fmm_relax_tiles(const uint32_t* __restrict__ dom,
                                                         const uint8_t* __restrict__ fixed, uint32_t nx, uint32_t ny,
                                                         double dx, double* __restrict__ u,
                                                         uint32_t* __restrict__ changed) {
    __shared__ double su[kTH * kTH];
    __shared__ uint32_t sk[kTH * kTH];
    __shared__ uint8_t sfix[kTH * kTH];
    const uint32_t nty = (ny + kT - 1) / kT;
    const int x0 = (int)(blockIdx.x / nty) * kT - 1, y0 = (int)(blockIdx.x % nty) * kT - 1;
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
    if (threadIdx.x == 0) atomicAdd(changed, 1u);
}

// Per problem: fixed-point sum of the signed distances of the points a
// channel reaches, their count, the channel pixels and the own pixels.  A
// wave whose active lanes share one key adds once.
__global__ void k_fmm_reduce(const uint32_t* __restrict__ dom, const uint8_t* __restrict__ neg,
                             const int32_t* __restrict__ label, const double* __restrict__ u, uint32_t nx,
                             uint32_t ny, u64* __restrict__ sums, uint32_t* __restrict__ cnt,
                             uint32_t* __restrict__ nch, uint32_t* __restrict__ nphi) {
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
    const uint32_t first = __shfl(key, 0);
    const bool uniform = __all(key == first);
    if (uniform) {
        if (!first) return;
        for (int o = 32; o > 0; o >>= 1) {
            q += __shfl_down(q, o);
            c += __shfl_down(c, o);
            ch += __shfl_down(ch, o);
            own += __shfl_down(own, o);
        }
        if ((threadIdx.x & 63u) == 0u) {
            if (q) atomicAdd(&sums[first], (u64)q);
            if (c) atomicAdd(&cnt[first], c);
            if (ch) atomicAdd(&nch[first], ch);
            if (own) atomicAdd(&nphi[first], own);
        }
    } else if (key) {
        if (q) atomicAdd(&sums[key], (u64)q);
        if (c) atomicAdd(&cnt[key], c);
        if (ch) atomicAdd(&nch[key], ch);
        if (own) atomicAdd(&nphi[key], own);
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

extern "C" size_t slm_fmm_workspace_size(uint32_t nx, uint32_t ny) {
    const size_t n = (size_t)nx * ny;
    return al256(n * 4) + al256(n) + al256(n * 8) + al256(n) + 256 + 256;
}

extern "C" int slm_fmm_hsl(const int32_t* label, const uint32_t* mapping, uint32_t thin_flag, const uint8_t* colour,
                           uint32_t n_labels, uint32_t n_colours, uint32_t nx, uint32_t ny, double dx,
                           uint64_t* sums, uint32_t* cnt, uint32_t* nch, uint32_t* nphi, void* workspace,
                           size_t workspace_bytes, void* stream, uint32_t flags) {
    if (!label || !mapping || !colour || !sums || !cnt || !nch || !nphi || !workspace || !nx || !ny ||
        (uint64_t)nx * ny >= 0x80000000ull || !(dx > 0.0) || n_colours > 255)
        return slm_fail(SLM_E_INVALID_ARG, "bad fmm_hsl arguments");
    if (workspace_bytes < slm_fmm_workspace_size(nx, ny))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes, need %zu", workspace_bytes,
                        slm_fmm_workspace_size(nx, ny));
    const size_t n = (size_t)nx * ny;
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    uint32_t* dom = (uint32_t*)w;
    uint8_t* neg = w + al256(n * 4);
    double* u = (double*)(neg + al256(n));
    uint8_t* fixed = (uint8_t*)u + al256(n * 8);
    uint32_t* ctr = (uint32_t*)(fixed + al256(n));
    hipStream_t s = (hipStream_t)stream;
    zero(sums, (size_t)n_labels * 8, s);
    zero(cnt, (size_t)n_labels * 4, s);
    zero(nch, (size_t)n_labels * 4, s);
    zero(nphi, (size_t)n_labels * 4, s);
    const dim3 g = grid_for(n);
    // a bound on the launches keeps a bad input from looping
    const uint32_t n_tiles = ((nx + kT - 1) / kT) * ((ny + kT - 1) / kT);
    const uint32_t max_launches = 2u * (nx + ny) + 64u;
    for (uint32_t c = 0; c < n_colours; c++) {
        hipLaunchKernelGGL(k_fmm_setup, g, dim3(kBlock), 0, s, label, mapping, thin_flag, colour, n_labels, c, nx, ny,
                           dom, neg);
        hipLaunchKernelGGL(k_fmm_init, g, dim3(kBlock), 0, s, (const uint32_t*)dom, (const uint8_t*)neg, nx, ny, dx,
                           u, fixed);
        for (uint32_t launches = 0;; launches++) {
            if (launches > max_launches) return slm_fail(SLM_E_HIP, "eikonal relaxation did not converge");
            zero(ctr, sizeof(uint32_t), s);
            hipLaunchKernelGGL(k_fmm_relax_tiles, dim3(n_tiles), dim3(kBlock), 0, s, (const uint32_t*)dom,
                               (const uint8_t*)fixed, nx, ny, dx, u, ctr);
            uint32_t dec = 0;
            HIP_TRY(hipMemcpyAsync(&dec, ctr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            if (!dec) break;
        }
        hipLaunchKernelGGL(k_fmm_reduce, g, dim3(kBlock), 0, s, (const uint32_t*)dom, (const uint8_t*)neg, label,
                           (const double*)u, nx, ny, (u64*)sums, cnt, nch, nphi);
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
