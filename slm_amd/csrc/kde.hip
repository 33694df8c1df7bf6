// kde.hip -- univariate kernel-density estimation (histogram + FP64 filter).
//
//   k_kde_hist   histogram_univariate (opencl/kde.cl:383-394): one lane per
//                datum, bins privatised in LDS per block when they fit, merged
//                with one global atomic per non-empty bin.
//   k_kde_pdf    pdf_univariate + filter (opencl/kde.cl:426-526): one lane per
//                pdf point, the reference's loop order and FP64 accumulation
//                kept so that the sums are identical.
// Rounding: built with -ffp-contract=off; the multiply-adds clang's OpenCL
// default would fuse (x = edge + BIN_DX*(h+0.5); 1 - q*q) are fused only when
// fp_contract is set.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "../../include/slm_hip.h"
#include "common.hpp"

namespace {

constexpr uint32_t kLdsBins = 8192;  // 32 KB of LDS per block

__global__ __launch_bounds__(256) void k_kde_hist(SlmKdeParams K, const float* __restrict__ logx,
                                                  uint32_t* __restrict__ hist, uint32_t use_lds) {
    __shared__ uint32_t bins[kLdsBins];
    const uint32_t nb = K.n_hist_bins;
    if (use_lds) {
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) bins[i] = 0u;
        __syncthreads();
    }
    const float fnb = (float)nb;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < K.n_data; g += gridDim.x * blockDim.x) {
        const float bin_x = (logx[g] - K.logx_min) / K.logx_range;
        const float t = bin_x * fnb;
        // (uint)(t) then max(0u, .) and min(., N-1): saturating, NaN -> 0
        const uint32_t u = (t > 0.0f) ? ((t >= 4294967296.0f) ? 0xffffffffu : (uint32_t)t) : 0u;
        const uint32_t idx = min(u, nb - 1u);
        if (use_lds) atomicAdd(&bins[idx], 1u);
        else atomicAdd(&hist[idx], 1u);
    }
    if (use_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x)
            if (bins[i]) atomicAdd(&hist[i], bins[i]);
    }
}

// k_sample (opencl/kde.cl:36-92); w arrives as the double of a float literal.
template <bool C>
__device__ __forceinline__ double k_sample(uint32_t kind, double x, double w) {
    switch (kind) {
        case SLM_KDF_TOPHAT:
            return (fabs(x) < w) ? 1.0 : 0.0;
        case SLM_KDF_TRIANGLE:
            return fmax(1.0 - fabs(x / w), 0.0);
        case SLM_KDF_EPANECHNIKOV: {
            const double q = x / w;
            return fmax(C ? __builtin_fma(-q, q, 1.0) : 1.0 - q * q, 0.0);
        }
        case SLM_KDF_COSINE:
            return fmax(cos(1.57079632679489661923 * x), 0.0);
        default: {  // gaussian
            const double q = x / (w * (double)0.4f);
            return exp(-0.5 * q * q);
        }
    }
}

template <bool C>
__global__ void k_kde_pdf(SlmKdeParams K, const uint32_t* __restrict__ hist, float* __restrict__ pdf) {
    const uint32_t pdf_idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (pdf_idx >= K.n_pdf_points) return;
    const uint32_t nbp = K.n_hist_bins / K.n_pdf_points;
    const double pdf_dx = (double)K.pdf_dx;
    const double bin_dx = (double)K.bin_dx;
    const double half = (double)(K.pdf_dx / 2.0f);  // PDF_DX/2.0f in float
    const double w = (double)K.kdf_width_x;
    const int npdf = (int)K.n_pdf_points;
    double acc = 0.0, kw = 0.0;
    for (uint32_t k_idx = 0; k_idx <= K.n_kdf_part_points_x; k_idx++) {
        const double k_bin_edge = (double)k_idx * pdf_dx - half;
        const int col_r = (int)(pdf_idx + k_idx);
        const int col_l = (int)(pdf_idx - k_idx + 1u);
        const bool ok_r = col_r < npdf;
        const bool ok_l = (k_idx != 0u) && (col_l > 0);
        for (uint32_t h = 0; h < nbp; h++) {
            const double hh = (double)h + 0.5;
            const double x = C ? __builtin_fma(bin_dx, hh, k_bin_edge) : k_bin_edge + bin_dx * hh;
            const double kv = k_sample<C>(K.kernel, x, w);
            const uint32_t hr = ok_r ? (uint32_t)(col_r * (int)nbp + (int)h) : 0u;
            const uint32_t hl = ok_l ? (uint32_t)(col_l * (int)nbp - (int)h - 1) : 0u;
            acc += ok_r ? (double)hist[hr] * kv : 0.0;
            acc += ok_l ? (double)hist[hl] * kv : 0.0;
            kw += ok_r ? kv : 0.0;
            kw += ok_l ? kv : 0.0;
        }
    }
    pdf[pdf_idx] = (float)((double)(float)acc / kw);
}

int check_kde(const SlmKdeParams* k) {
    if (!k) return slm_fail(SLM_E_INVALID_ARG, "kde params is NULL");
    if (k->abi_version != SLM_ABI_VERSION)
        return slm_fail(SLM_E_ABI, "SlmKdeParams.abi_version %u != %u", k->abi_version, SLM_ABI_VERSION);
    if (k->n_hist_bins == 0 || k->n_pdf_points == 0 || k->n_hist_bins < k->n_pdf_points)
        return slm_fail(SLM_E_INVALID_ARG, "need n_hist_bins >= n_pdf_points > 0 (%u, %u)", k->n_hist_bins,
                        k->n_pdf_points);
    if (k->kernel > SLM_KDF_GAUSSIAN) return slm_fail(SLM_E_INVALID_ARG, "unknown kernel %u", k->kernel);
    return SLM_OK;
}

__global__ void k_zero_u32(uint32_t* a, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0u;
}

}  // namespace

extern "C" int slm_kde_histogram(const SlmKdeParams* k, const float* logx, uint32_t* hist, void* stream,
                                 uint32_t flags) {
    int rc = check_kde(k);
    if (rc) return rc;
    if (!hist || (k->n_data && !logx)) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    if (!(flags & SLM_FLAG_NO_ZERO))
        hipLaunchKernelGGL(k_zero_u32, dim3((k->n_hist_bins + 255) / 256), dim3(256), 0, s, hist, k->n_hist_bins);
    if (k->n_data) {
        uint32_t blocks = (k->n_data + 255) / 256;
        if (blocks > 2048) blocks = 2048;
        const uint32_t use_lds = (k->n_hist_bins <= kLdsBins) ? 1u : 0u;
        hipLaunchKernelGGL(k_kde_hist, dim3(blocks), dim3(256), 0, s, *k, logx, hist, use_lds);
    }
    return slm_finish(stream, flags);
}

extern "C" int slm_kde_pdf(const SlmKdeParams* k, const uint32_t* hist, float* pdf, void* stream,
                           uint32_t flags) {
    int rc = check_kde(k);
    if (rc) return rc;
    if (!hist || !pdf) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t n = k->n_pdf_points;
    if (k->fp_contract)
        hipLaunchKernelGGL(k_kde_pdf<true>, dim3((n + 63) / 64), dim3(64), 0, s, *k, hist, pdf);
    else
        hipLaunchKernelGGL(k_kde_pdf<false>, dim3((n + 63) / 64), dim3(64), 0, s, *k, hist, pdf);
    return slm_finish(stream, flags);
}
