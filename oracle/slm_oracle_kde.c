/*
 * slm_oracle_kde.c -- CPU restatement of slm's univariate KDE kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see slm_oracle.c).  Restates opencl/kde.cl:
 *   histogram_univariate  :383-394
 *   pdf_univariate        :500-526 with filter :426-484 and k_sample :36-92
 * with the constants from SlmKdeParams.  Compiled with -ffp-contract=off;
 * the multiply-adds clang's OpenCL default fuses are explicit fma() calls,
 * taken when fp_contract is set.
 */
#include <math.h>
#include <stdint.h>
#include "../include/slm_hip.h"

static uint32_t cvt_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

void slmo_kde_histogram(const SlmKdeParams* k, const float* logx, uint32_t* hist) {
    const uint32_t nb = k->n_hist_bins;
    for (uint32_t g = 0; g < k->n_data; g++) {
        const float bin_x = (logx[g] - k->logx_min) / k->logx_range;
        uint32_t idx = cvt_u32(bin_x * (float)nb);
        if (idx > nb - 1u) idx = nb - 1u;
        hist[idx] += 1u;
    }
}

static double k_sample(const SlmKdeParams* k, double x, double w) {
    switch (k->kernel) {
        case SLM_KDF_TOPHAT:
            return (fabs(x) < w) ? 1.0 : 0.0;
        case SLM_KDF_TRIANGLE:
            return fmax(1.0 - fabs(x / w), 0.0);
        case SLM_KDF_EPANECHNIKOV: {
            const double q = x / w;
            return fmax(k->fp_contract ? fma(-q, q, 1.0) : 1.0 - q * q, 0.0);
        }
        case SLM_KDF_COSINE:
            return fmax(cos(1.57079632679489661923 * x), 0.0);
        default: {
            const double q = x / (w * (double)0.4f);
            return exp(-0.5 * q * q);
        }
    }
}

void slmo_kde_pdf(const SlmKdeParams* k, const uint32_t* hist, float* pdf) {
    const uint32_t nbp = k->n_hist_bins / k->n_pdf_points;
    const double pdf_dx = (double)k->pdf_dx, bin_dx = (double)k->bin_dx;
    const double half = (double)(k->pdf_dx / 2.0f);
    const double w = (double)k->kdf_width_x;
    for (uint32_t pdf_idx = 0; pdf_idx < k->n_pdf_points; pdf_idx++) {
        double acc = 0.0, kw = 0.0;
        for (uint32_t k_idx = 0; k_idx <= k->n_kdf_part_points_x; k_idx++) {
            /* filter (kde.cl:426-484) */
            const double k_bin_point = (double)k_idx * pdf_dx;
            const double k_bin_edge = k_bin_point - half;
            const int col_r = (int)(pdf_idx + k_idx);
            const int col_l = (int)(pdf_idx - k_idx + 1u);
            for (uint32_t h = 0; h < nbp; h++) {
                const double hh = (double)h + 0.5;
                const double x = k->fp_contract ? fma(bin_dx, hh, k_bin_edge) : k_bin_edge + bin_dx * hh;
                const double kv = k_sample(k, x, w);
                const int ok_r = col_r < (int)k->n_pdf_points;
                const int ok_l = (k_idx != 0u) && (col_l > 0);
                const uint32_t hr = ok_r ? (uint32_t)col_r * nbp + h : 0u;
                const uint32_t hl = ok_l ? (uint32_t)col_l * nbp - h - 1u : 0u;
                acc += ok_r ? (double)hist[hr] * kv : 0.0;
                acc += ok_l ? (double)hist[hl] * kv : 0.0;
                kw += ok_r ? kv : 0.0;
                kw += ok_l ? kv : 0.0;
            }
        }
        pdf[pdf_idx] = (float)((double)(float)acc / kw);
    }
}
