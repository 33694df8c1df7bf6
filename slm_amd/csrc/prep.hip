// prep.hip -- the grids either side of the tracer: the uv field it consumes
// and the finished slc / slt / sla fields it produces.
//
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

inline dim3 grid_of(size_t n) { return dim3((uint32_t)((n + kBlock - 1) / kBlock)); }

}  // namespace

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
