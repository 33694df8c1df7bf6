// prep.hip -- the grids either side of the tracer: the uv field it consumes
// and the finished slc / slt / sla fields it produces.
//
//   slm_condition_uv       preprocess.py:531-600 with :196-473: the same uv
//                          after the loop and blockage fixes
//   slm_gradient_uv        preprocess.py:51-62 (np.gradient / pixel_size),
//                          :100-193 (u,v = -grad, speed, normalise),
//                          :584-600 (re-normalise, 'edge' pad), :602-621
//                          (NaN uv -> masked, uv 0), with the DTM mask of
//                          geodata (NaN -> masked, pad masked)
//   slm_fields_finish_raw  fields.py:296-309 (sla, slt / d^2)
//   slm_fields_finish_scaled fields.py:141-164 (slt * ds/pixel_size, sqrt;
//                          slc -> sqrt(2) (slc/d^2)^(2/3) in float64)
//
// Float32 arithmetic follows the dtypes numpy gave the reference (1.x
// promotion rules): every op rounded to float32 unless stated.  HBM-bound
// elementwise kernels, one thread per pixel.
#include <hip/hip_runtime.h>

#include <math.h>

#include "common.hpp"

namespace {

constexpr int kBlock = 256;

// np.hypot on float32 equals the float64 sqrt of the float64 sum of squares
// rounded once to float32 (exact squares; glibc's hypotf computes it so).
__device__ __forceinline__ float hypot32(float u, float v) {
    if (isinf(u) || isinf(v)) return INFINITY;
    const double d = (double)u * (double)u + (double)v * (double)v;
    return (float)__builtin_sqrt(d);
}

// set_speed_field + normalize_velocity_field (preprocess.py:124-160)
__device__ __forceinline__ void normalise(float& u, float& v) {
    float s = hypot32(u, v);
    if (!isfinite(s) || s == 0.0f) s = 1.0f;
    u = u / s;
    v = v / s;
}

// u, v = -grad h / pixel_size at ROI pixel (x, y) (np.gradient, edge_order 1,
// unit spacing: one-sided at the ends), preprocess.py:51-62, :100-117
__device__ __forceinline__ void neg_gradient(const float* __restrict__ r, uint32_t nx, uint32_t ny, uint32_t x,
                                             uint32_t y, float ps, float& u, float& v) {
    auto at = [&](uint32_t a, uint32_t b) { return r[(size_t)a * ny + b]; };
    float gx, gy;
    if (x == 0)
        gx = (at(1, y) - at(0, y)) / 1.0f;
    else if (x == nx - 1)
        gx = (at(nx - 1, y) - at(nx - 2, y)) / 1.0f;
    else
        gx = (at(x + 1, y) - at(x - 1, y)) / 2.0f;
    if (y == 0)
        gy = (at(x, 1) - at(x, 0)) / 1.0f;
    else if (y == ny - 1)
        gy = (at(x, ny - 1) - at(x, ny - 2)) / 1.0f;
    else
        gy = (at(x, y + 1) - at(x, y - 1)) / 2.0f;
    u = -(gx / ps);
    v = -(gy / ps);
}

__global__ void k_gradient_uv(const float* __restrict__ roi, uint32_t nx, uint32_t ny, uint32_t pad, float ps,
                              int renorm, float2* __restrict__ uv, uint8_t* __restrict__ mask) {
    const uint32_t nxp = nx + 2 * pad, nyp = ny + 2 * pad;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxp * nyp) return;
    const uint32_t xp = i / nyp, yp = i % nyp;
    // 'edge' padding of u and v: the nearest ROI pixel
    const uint32_t x = (uint32_t)min(max((int)xp - (int)pad, 0), (int)nx - 1);
    const uint32_t y = (uint32_t)min(max((int)yp - (int)pad, 0), (int)ny - 1);
    const float* r = roi;
    auto at = [&](uint32_t a, uint32_t b) { return r[(size_t)a * ny + b]; };
    // np.gradient, edge_order 1, unit spacing: one-sided at the ends
    float gx, gy;
    if (x == 0)
        gx = (at(1, y) - at(0, y)) / 1.0f;
    else if (x == nx - 1)
        gx = (at(nx - 1, y) - at(nx - 2, y)) / 1.0f;
    else
        gx = (at(x + 1, y) - at(x - 1, y)) / 2.0f;
    if (y == 0)
        gy = (at(x, 1) - at(x, 0)) / 1.0f;
    else if (y == ny - 1)
        gy = (at(x, ny - 1) - at(x, ny - 2)) / 1.0f;
    else
        gy = (at(x, y + 1) - at(x, y - 1)) / 2.0f;
    gx = gx / ps;
    gy = gy / ps;
    float u = -gx, v = -gy;
    normalise(u, v);
    if (renorm) normalise(u, v);
    const bool uv_bad = isnan(u) || isnan(v);
    const bool in_roi = xp >= pad && xp < nx + pad && yp >= pad && yp < ny + pad;
    const bool dtm_bad = !in_roi || isnan(at(x, y));
    uv[i] = uv_bad ? make_float2(0.0f, 0.0f) : make_float2(u, v);
    mask[i] = (uv_bad || dtm_bad) ? 1 : 0;
}

// fields.py:296-309 per pixel and direction k: raw counts c (uint32) and
// raw length sums t (uint32, cast to float32 as fields.py:288 does).
__global__ void k_finish_raw(const uint32_t* __restrict__ slc_raw, const uint32_t* __restrict__ slt_raw,
                             uint32_t nxy, float d2, float* __restrict__ slt, float* __restrict__ sla) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const uint32_t c = slc_raw[(size_t)k * nxy + i];
        float t = (float)slt_raw[(size_t)k * nxy + i];
        float a = 0.0f;
        if (c > 0u) {
            // np.sqrt(2.0)*slt (float32) / slc (float64 division), stored float32
            const float prod = 1.41421353816986083984375f * t;
            a = (float)((double)prod / (double)c);
            t = t / d2;
        }
        slt[2 * (size_t)i + k] = t;
        sla[2 * (size_t)i + k] = a;
    }
}

// fields.py:141-164: slt[...,k] *= ds_k/pixel_size (float32); slt = sqrt;
// slc = sqrt(2.0) * (slc / d^2)^(2/3) in float64.
__global__ void k_finish_scaled(const uint32_t* __restrict__ slc_raw, uint32_t nxy, double d2, float sd, float su,
                                float* __restrict__ slt, double* __restrict__ slc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const size_t o = 2 * (size_t)i + k;
        slt[o] = __builtin_sqrtf(slt[o] * (k == 0 ? sd : su));
        const double c = (double)slc_raw[(size_t)k * nxy + i];
        slc[o] = 1.4142135623730951 * pow(c / d2, 2.0 / 3.0);
    }
}

// ------------------------------------------------------------------ DTM conditioning
// Preprocess.conditioned_gradient_vector_field (preprocess.py:531-600): find
// blockages (single diagonal outflows) and the pixels upstream of them, find
// loops (2x2 cells of the normalised uv field whose mean vector is short,
// which converge and curl), point loop corners at their lowest neighbour and
// blockages along their outflow, then normalise, pad and mask as the plain
// path.  Every fix depends on the DTM alone, so the order of the reference's
// fix loops does not matter and each pixel is computed independently.

// scipy.ndimage 'reflect' border for a 3x3 window (index -1 -> 0, n -> n-1)
__device__ __forceinline__ uint32_t reflect1(int c, uint32_t n) {
    return (uint32_t)(c < 0 ? -c - 1 : (c >= (int)n ? 2 * (int)n - c - 1 : c));
}

// get_flow_vector (preprocess.py:196-227): SW, SE, NW, NE, W, E, S, N
__device__ __forceinline__ void flow_vec(uint32_t nn, int& dx, int& dy) {
    dx = 0;
    dy = 0;
    if (nn & 16) { dx = -1; dy = -1; }
    else if (nn & 32) { dx = 1; dy = -1; }
    else if (nn & 64) { dx = -1; dy = 1; }
    else if (nn & 128) { dx = 1; dy = 1; }
    else if (nn & 1) { dx = -1; }
    else if (nn & 2) { dx = 1; }
    else if (nn & 4) { dy = -1; }
    else if (nn & 8) { dy = 1; }
}

// Pass A, per ROI pixel: the blockage code (has_one_diagonal_outflow,
// :390-432, on the 3x3 window with 'reflect' borders; NaN compares false),
// the normalised u, v and the raw speed (compute_gradient_velocity_field,
// :177-193).
__global__ void k_cond_a(const float* __restrict__ roi, uint32_t nx, uint32_t ny, float ps,
                         uint8_t* __restrict__ code_b, float2* __restrict__ uvn, float* __restrict__ raw) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const uint32_t x = i / ny, y = i % ny;
    const float h = roi[i];
    // (dx, dy, bit) in the callback's window order; bits: SW 16, S 4, SE 32,
    // W 1, E 2, NW 64, N 8, NE 128
    const int dxs[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
    const int dys[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    const uint32_t bits[8] = {16, 4, 32, 1, 2, 64, 8, 128};
    uint32_t nn = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t xx = reflect1((int)x + dxs[k], nx), yy = reflect1((int)y + dys[k], ny);
        if (roi[(size_t)xx * ny + yy] < h) nn += bits[k];
    }
    code_b[i] = ((nn & 240u) == nn) ? (uint8_t)nn : 0;
    float u, v;
    neg_gradient(roi, nx, ny, x, y, ps, u, v);
    float sp = hypot32(u, v);
    if (!isfinite(sp) || sp == 0.0f) sp = 1.0f;
    uvn[i] = make_float2(u / sp, v / sp);
    raw[i] = sp;
}

// Pass B, per ROI pixel: the blocked-neighbour code (upstream_of_diagonal_
// outflow, :434-473, over the blockage codes with 'reflect' borders) and the
// loop flag of the 2x2 cell at (x, y) (check_has_loop, :229-261: the mean
// vector in float64 as numba types it, the +-1 dot products and their sums
// in float32; the cells of the last row and column, where the reference
// reads past its arrays, are never loops).
__global__ void k_cond_b(const uint8_t* __restrict__ code_b, const float2* __restrict__ uvn, uint32_t nx,
                         uint32_t ny, double vecsum_thr, double div_thr, double curl_thr,
                         uint8_t* __restrict__ code_n, uint8_t* __restrict__ loop) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    const uint32_t x = i / ny, y = i % ny;
    auto c = [&](int dx, int dy) {
        return (uint32_t)code_b[(size_t)reflect1((int)x + dx, nx) * ny + reflect1((int)y + dy, ny)];
    };
    const uint32_t w0 = c(-1, -1), w1 = c(0, -1), w2 = c(1, -1), w3 = c(-1, 0);
    const uint32_t w5 = c(1, 0), w6 = c(-1, 1), w7 = c(0, 1), w8 = c(1, 1);
    uint32_t nn = 0;
    nn += (w0 == 16) ? 16 : 0;
    nn += (w1 == 16) ? 4 : 0;
    nn += (w1 == 32) ? 4 : 0;
    nn += (w2 == 32) ? 32 : 0;
    nn += (w3 == 16) ? 1 : 0;
    nn += (w3 == 64) ? 1 : 0;
    nn += (w5 == 128) ? 2 : 0;
    nn += (w5 == 32) ? 2 : 0;
    nn += (w6 == 64) ? 64 : 0;
    nn += (w7 == 64) ? 8 : 0;
    nn += (w7 == 128) ? 8 : 0;
    nn += (w8 == 128) ? 128 : 0;
    code_n[i] = (uint8_t)nn;
    uint8_t looped = 0;
    if (x + 1 < nx && y + 1 < ny) {
        const float2 a = uvn[i], b = uvn[i + ny], d = uvn[i + ny + 1], e = uvn[i + 1];   // 00, 10, 11, 01
        const double mu = (double)(((a.x + b.x) + d.x) + e.x) / 4.0;
        const double mv = (double)(((a.y + b.y) + d.y) + e.y) / 4.0;
        const float speed = (float)__builtin_sqrt(mu * mu + mv * mv);
        const float div = (((-a.x - a.y) + (b.x - b.y)) + (d.x + d.y)) + (-e.x + e.y);
        const float curl = (((a.x - a.y) + (b.x + b.y)) + (-d.x + d.y)) + (-e.x - e.y);
        looped = ((double)speed < vecsum_thr && (double)div <= div_thr && (double)fabsf(curl) >= curl_thr) ? 1 : 0;
    }
    loop[i] = looped;
}

// Pass C, per padded output pixel (edge padding: the nearest ROI pixel):
// find_and_fix_loops (:303-350) + break_out_of_loop (:263-301) for a pixel
// at a corner of a looped cell and well inside the grid, then fix_blockages
// (:352-388: blockages, then blocked neighbours, override), then the
// (un)normalisation, padding and NaN mask of :584-621.
__global__ void k_cond_c(const float* __restrict__ roi, uint32_t nx, uint32_t ny, uint32_t pad,
                         const uint8_t* __restrict__ code_b, const uint8_t* __restrict__ code_n,
                         const uint8_t* __restrict__ loop, const float2* __restrict__ uvn,
                         const float* __restrict__ raw, int renorm, float2* __restrict__ uv,
                         uint8_t* __restrict__ mask) {
#pragma clang fp contract(off)
    const uint32_t nxp = nx + 2 * pad, nyp = ny + 2 * pad;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxp * nyp) return;
    const uint32_t xp = i / nyp, yp = i % nyp;
    const uint32_t x = (uint32_t)min(max((int)xp - (int)pad, 0), (int)nx - 1);
    const uint32_t y = (uint32_t)min(max((int)yp - (int)pad, 0), (int)ny - 1);
    const size_t j = (size_t)x * ny + y;
    float2 f = uvn[j];
    if (x >= 1 && x + 2 <= nx && y >= 1 && y + 2 <= ny) {
        const bool looped = loop[j - ny - 1] | loop[j - ny] | loop[j - 1] | loop[j];
        if (looped) {
            // W, E, S, N, SW, SE, NW, NE; strict <, so the first lowest wins
            const int dxs[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
            const int dys[8] = {0, 0, -1, 1, -1, -1, 1, 1};
            float h_min = 3.40282346638528859812e+38f;
            int best = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float hk = roi[(size_t)(x + dxs[k]) * ny + (y + dys[k])];
                if (h_min > hk) {
                    h_min = hk;
                    best = k;
                }
            }
            const double len = (dxs[best] && dys[best]) ? 1.4142135623730951 : 1.0;
            f = make_float2((float)((double)dxs[best] / len), (float)((double)dys[best] / len));
        }
    }
    int dx, dy;
    if (code_b[j]) {
        flow_vec(code_b[j], dx, dy);
        f = make_float2((float)((double)(float)dx / 1.4142135623730951), (float)((double)(float)dy / 1.4142135623730951));
    }
    if (code_n[j]) {
        flow_vec(code_n[j], dx, dy);
        f = make_float2((float)((double)(float)dx / 1.4142135623730951), (float)((double)(float)dy / 1.4142135623730951));
    }
    float u = f.x, v = f.y;
    if (renorm) {
        normalise(u, v);
    } else {
        u = u * raw[j];
        v = v * raw[j];
    }
    const bool uv_bad = isnan(u) || isnan(v);
    const bool in_roi = xp >= pad && xp < nx + pad && yp >= pad && yp < ny + pad;
    const bool dtm_bad = !in_roi || isnan(roi[j]);
    uv[i] = uv_bad ? make_float2(0.0f, 0.0f) : make_float2(u, v);
    mask[i] = (uv_bad || dtm_bad) ? 1 : 0;
}

inline dim3 grid_of(size_t n) { return dim3((uint32_t)((n + kBlock - 1) / kBlock)); }

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Preprocess.slope_array (preprocess.py:597): degrees(arctan(raw speed)),
// 'edge'-padded, float32; raw speed = |grad h| / pixel_size (np.hypot on
// float32), 1 where it is 0 or not finite (set_speed_field).  The arctan is
// the float64 one rounded to float32, then the float32 product by
// float32(180/pi) (numpy's float32 rad2deg); numpy's float32 arctan is a
// vectorised approximation that can differ by an ulp.
__global__ void k_slope_deg(const float* __restrict__ roi, uint32_t nx, uint32_t ny, uint32_t pad, float ps,
                            float* __restrict__ slope) {
    const uint32_t nxp = nx + 2 * pad, nyp = ny + 2 * pad;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxp * nyp) return;
    const uint32_t x = (uint32_t)min(max((int)(i / nyp) - (int)pad, 0), (int)nx - 1);
    const uint32_t y = (uint32_t)min(max((int)(i % nyp) - (int)pad, 0), (int)ny - 1);
    float u, v;
    neg_gradient(roi, nx, ny, x, y, ps, u, v);
    float sp = hypot32(u, v);
    if (!isfinite(sp) || sp == 0.0f) sp = 1.0f;
    const float a = (float)atan((double)sp);
    slope[i] = a * (float)(180.0 / 3.14159265358979323846);
}

}  // namespace

extern "C" int slm_slope_deg(const float* roi, uint32_t nx, uint32_t ny, uint32_t pad, float pixel_size,
                             float* slope, void* stream, uint32_t flags) {
    if (nx < 2 || ny < 2) return slm_fail(SLM_E_INVALID_ARG, "grid %u x %u: np.gradient needs 2 points", nx, ny);
    if ((uint64_t)(nx + 2ull * pad) * (ny + 2ull * pad) >= 0x80000000ull)
        return slm_fail(SLM_E_INVALID_ARG, "padded grid above 2^31 pixels");
    if (!roi || !slope) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (!(pixel_size > 0.0f)) return slm_fail(SLM_E_INVALID_ARG, "pixel_size must be > 0");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)(nx + 2 * pad) * (ny + 2 * pad);
    hipLaunchKernelGGL(k_slope_deg, grid_of(n), dim3(kBlock), 0, s, roi, nx, ny, pad, pixel_size, slope);
    return slm_finish(stream, flags);
}

extern "C" int slm_gradient_uv(const float* roi, uint32_t nx, uint32_t ny, uint32_t pad, float pixel_size,
                               int do_normalize_speed, float* uv, uint8_t* mask, void* stream, uint32_t flags) {
    if (nx < 2 || ny < 2) return slm_fail(SLM_E_INVALID_ARG, "grid %u x %u: np.gradient needs 2 points", nx, ny);
    if ((uint64_t)(nx + 2ull * pad) * (ny + 2ull * pad) >= 0x80000000ull)
        return slm_fail(SLM_E_INVALID_ARG, "padded grid above 2^31 pixels");
    if (!roi || !uv || !mask) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (!(pixel_size > 0.0f)) return slm_fail(SLM_E_INVALID_ARG, "pixel_size must be > 0");
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)(nx + 2 * pad) * (ny + 2 * pad);
    hipLaunchKernelGGL(k_gradient_uv, grid_of(n), dim3(kBlock), 0, s, roi, nx, ny, pad, pixel_size,
                       do_normalize_speed ? 1 : 0, (float2*)uv, mask);
    return slm_finish(stream, flags);
}

extern "C" int slm_fields_finish_raw(const uint32_t* slc_raw, const uint32_t* slt_raw, uint32_t nxy,
                                     uint32_t density, float* slt, float* sla, void* stream, uint32_t flags) {
    if (!nxy) return slm_finish(stream, flags);
    if (!slc_raw || !slt_raw || !slt || !sla) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (density == 0) return slm_fail(SLM_E_INVALID_ARG, "density must be > 0");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_finish_raw, grid_of(nxy), dim3(kBlock), 0, s, slc_raw, slt_raw, nxy,
                       (float)(density * density), slt, sla);
    return slm_finish(stream, flags);
}

extern "C" int slm_fields_finish_scaled(const uint32_t* slc_raw, uint32_t nxy, uint32_t density,
                                        float ds_scale_down, float ds_scale_up, float* slt, double* slc,
                                        void* stream, uint32_t flags) {
    if (!nxy) return slm_finish(stream, flags);
    if (!slc_raw || !slt || !slc) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (density == 0) return slm_fail(SLM_E_INVALID_ARG, "density must be > 0");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_finish_scaled, grid_of(nxy), dim3(kBlock), 0, s, slc_raw, nxy,
                       (double)density * (double)density, ds_scale_down, ds_scale_up, slt, slc);
    return slm_finish(stream, flags);
}

extern "C" size_t slm_condition_workspace_size(uint32_t nx, uint32_t ny) {
    const size_t n = (size_t)nx * ny;
    return 3 * align256(n) + align256(n * 8) + align256(n * 4) + 256;
}

extern "C" int slm_condition_uv(const float* roi, uint32_t nx, uint32_t ny, uint32_t pad, float pixel_size,
                                int do_normalize_speed, double vecsum_threshold, double divergence_threshold,
                                double curl_threshold, float* uv, uint8_t* mask, uint8_t* maps, void* workspace,
                                size_t workspace_bytes, void* stream, uint32_t flags) {
    if (nx < 2 || ny < 2) return slm_fail(SLM_E_INVALID_ARG, "grid %u x %u: np.gradient needs 2 points", nx, ny);
    if ((uint64_t)(nx + 2ull * pad) * (ny + 2ull * pad) >= 0x80000000ull)
        return slm_fail(SLM_E_INVALID_ARG, "padded grid above 2^31 pixels");
    if (!roi || !uv || !mask || !workspace) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (!(pixel_size > 0.0f)) return slm_fail(SLM_E_INVALID_ARG, "pixel_size must be > 0");
    if (workspace_bytes < slm_condition_workspace_size(nx, ny))
        return slm_fail(SLM_E_WORKSPACE, "workspace %zu bytes, need %zu", workspace_bytes,
                        slm_condition_workspace_size(nx, ny));
    const size_t n = (size_t)nx * ny;
    uint8_t* base = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    uint8_t* code_b = maps ? maps : base;
    uint8_t* code_n = maps ? maps + n : base + align256(n);
    uint8_t* loop = maps ? maps + 2 * n : base + 2 * align256(n);
    float2* uvn = (float2*)(base + 3 * align256(n));
    float* raw = (float*)(base + 3 * align256(n) + align256(n * 8));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cond_a, grid_of(n), dim3(kBlock), 0, s, roi, nx, ny, pixel_size, code_b, uvn, raw);
    hipLaunchKernelGGL(k_cond_b, grid_of(n), dim3(kBlock), 0, s, (const uint8_t*)code_b, (const float2*)uvn, nx, ny,
                       vecsum_threshold, divergence_threshold, curl_threshold, code_n, loop);
    const size_t np_ = (size_t)(nx + 2 * pad) * (ny + 2 * pad);
    hipLaunchKernelGGL(k_cond_c, grid_of(np_), dim3(kBlock), 0, s, roi, nx, ny, pad, (const uint8_t*)code_b,
                       (const uint8_t*)code_n, (const uint8_t*)loop, (const float2*)uvn, (const float*)raw,
                       do_normalize_speed ? 1 : 0, (float2*)uv, mask);
    return slm_finish(stream, flags);
}
