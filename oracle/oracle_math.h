/*
 * oracle_math.h -- shared arithmetic of the CPU oracle (test infrastructure
 * only): the float semantics, RNG, essentials.cl and computestep.cl helpers
 * restated once for slm_oracle.c (FP32 and FP64 builds) and
 * slm_oracle_map.c (FP32).  Define SLM_F64 before including for the FP64 twin.
 */
#ifndef SLM_ORACLE_MATH_H
#define SLM_ORACLE_MATH_H

#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include "../include/slm_hip.h"

#ifdef SLM_F64
#define SLM_REAL double
#define SLM_PREFIX(name) slmo64_##name
#else
#define SLM_REAL float
#define SLM_PREFIX(name) slmo_##name
#endif

typedef SLM_REAL real;

#define IS_F32 (sizeof(real) == sizeof(float))

static inline real r_sqrt(real x) { return IS_F32 ? (real)sqrtf((float)x) : (real)sqrt((double)x); }
static inline real r_fma(real a, real b, real c) {
    return IS_F32 ? (real)fmaf((float)a, (float)b, (float)c) : (real)fma((double)a, (double)b, (double)c);
}
static inline real r_fmin(real a, real b) { return IS_F32 ? (real)fminf((float)a, (float)b) : (real)fmin(a, b); }
static inline real r_fmax(real a, real b) { return IS_F32 ? (real)fmaxf((float)a, (float)b) : (real)fmax(a, b); }

/* a*b+c, fused only in FP_CONTRACT-ON mode */
static inline real MAD(int contract, real a, real b, real c) {
    return contract ? r_fma(a, b, c) : a * b + c;
}

/* (uint)f for f >= 0 after max(0,.): saturating, NaN -> 0 (v_cvt_u32_f32 semantics) */
static inline uint32_t cvt_u32(real f) {
    if (!(f > (real)0)) return 0u;
    if (f >= (real)4294967296.0) return 0xffffffffu;
    return (uint32_t)f;
}
/* (int)f: saturating truncation, NaN -> 0 (v_cvt_i32_f32 semantics) */
static inline int32_t cvt_i32(real f) {
    if (f != f) return 0;
    if (f >= (real)2147483648.0) return INT32_MAX;
    if (f <= (real)-2147483648.0) return INT32_MIN;
    return (int32_t)f;
}
/* (char)f: low byte of the int32 conversion, sign-extended */
static inline int8_t cvt_i8(real f) {
    return (int8_t)(uint8_t)((uint32_t)cvt_i32(f) & 0xffu);
}

/* ---------------------------------------------------------------- RNG (rng.cl) */

/* lehmer_rand_uint (rng.cl:37-45): s <- ((u64)(uint32)(s+1) * 279470273) mod (2^32-5).
 * The +1 wraps in 32 bits before widening. */
static inline uint32_t lehmer_u32(uint32_t* s) {
    uint64_t x = (uint64_t)(uint32_t)(*s + 1u) * 279470273ull;
    *s = (uint32_t)(x % 0xfffffffbull);
    return *s;
}

/* lehmer_rand_vec (rng.cl:62-67): x drawn first, then y.
 * (float)0xfffffffb rounds to 2^32; the FP64 build divides by the same 2^32
 * so that it differs from the FP32 build only in arithmetic rounding. */
static inline void rand_vec(uint32_t* s, real out[2]) {
    const real m = (real)(float)0xfffffffbu;
    uint32_t a = lehmer_u32(s);
    uint32_t b = lehmer_u32(s);
    if (IS_F32) {
        out[0] = (real)((float)a / (float)m - 0.5f);
        out[1] = (real)((float)b / (float)m - 0.5f);
    } else {
        out[0] = (real)a / m - (real)0.5;
        out[1] = (real)b / m - (real)0.5;
    }
}

static inline uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

/* integratefields.cl:93-99: byte-reverse, one Lehmer draw, byte-reverse */
static inline uint32_t rng_seed_u32(uint32_t seed_idx) {
    uint32_t s = bswap32(seed_idx);
    lehmer_u32(&s);
    return bswap32(s);
}

/* ------------------------------------------------------- essentials.cl */

/* get_array_idx (essentials.cl:56-59) */
static inline uint32_t array_idx(const SlmParams* p, real x, real y) {
    uint32_t xi = cvt_u32(r_fmax((real)0, x + (real)p->pad_width_pp5));
    uint32_t yi = cvt_u32(r_fmax((real)0, y + (real)p->pad_width_pp5));
    if (xi > p->nx_padded - 1u) xi = p->nx_padded - 1u;
    if (yi > p->ny_padded - 1u) yi = p->ny_padded - 1u;
    return yi + p->ny_padded * xi;
}

static inline real dot2(int c, real x, real y) { return MAD(c, x, x, y * y); }
static inline real fast_length(int c, real x, real y) { return r_sqrt(dot2(c, x, y)); }

/* speed_interpolator (essentials.cl:24-45): extrapolating "bilinear" blend of
 * pixels round(x), round(x)+1 (same for y), then fast_normalize. */
static inline void speed_interp(const SlmParams* p, const float* uv, real x, real y, real out[2]) {
    const int c = (int)p->fp_contract;
    uint32_t xl = cvt_u32(r_fmax((real)0, x + (real)p->pad_width_pp5));
    uint32_t yd = cvt_u32(r_fmax((real)0, y + (real)p->pad_width_pp5));
    if (xl > p->nx_padded - 1u) xl = p->nx_padded - 1u;
    if (yd > p->ny_padded - 1u) yd = p->ny_padded - 1u;
    uint32_t xr = xl + 1u; if (xr > p->nx_padded - 1u) xr = p->nx_padded - 1u;
    uint32_t yu = yd + 1u; if (yu > p->ny_padded - 1u) yu = p->ny_padded - 1u;
    const real rx = x - (real)xl + (real)p->pad_width;
    const real lx = (real)1 - rx;
    const real uy = y - (real)yd + (real)p->pad_width;
    const real dy = (real)1 - uy;
    const size_t nyp = p->ny_padded;
    const float* a = uv + 2 * (nyp * xl + yd);   /* (xl, yd) */
    const float* b = uv + 2 * (nyp * xr + yd);   /* (xr, yd) */
    const float* e = uv + 2 * (nyp * xl + yu);   /* (xl, yu) */
    const float* f = uv + 2 * (nyp * xr + yu);   /* (xr, yu) */
    real dwn[2], upp[2], s[2];
    for (int k = 0; k < 2; k++) {
        dwn[k] = MAD(c, (real)a[k], lx, (real)b[k] * rx);
        upp[k] = MAD(c, (real)e[k], lx, (real)f[k] * rx);
    }
    for (int k = 0; k < 2; k++) s[k] = MAD(c, dwn[k], dy, upp[k] * uy);
    /* fast_normalize */
    const real inv = (real)1 / r_sqrt(dot2(c, s[0], s[1]));
    out[0] = s[0] * inv;
    out[1] = s[1] * inv;
}

/* approximate (essentials.cl:104-109): truncate onto the 1/res lattice via char */
static inline real approx(const SlmParams* p, real v) {
    const real res = (real)p->trajectory_resolution;
    return (real)cvt_i8(v * res) / res;
}

/* dt_to_nearest_edge (essentials.cl:121-129) */
static inline real dt_edge(real x, real u) {
    real dt = 0;
    if (u > (real)0) dt = ((real)cvt_i32(x + (real)1.5) - (x + (real)0.5)) / u;
    if (u < (real)0) dt = -((x + (real)0.5) - (real)cvt_i32(x + (real)0.5)) / u;
    return dt;
}

static inline real minmag(real x, real y) {
    real ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
    if (ax < ay) return x;
    if (ay < ax) return y;
    return r_fmin(x, y);
}

/* ------------------------------------------------------- computestep.cl */

typedef struct Step {
    real uv1[2], uv2[2], dxy1[2], dxy2[2], next[2];
    uint32_t idx;
} Step;

/* compute_step_vec (computestep.cl:42-57) and compute_step_vec_jittered (:86-106) */
static inline void compute_step(const SlmParams* p, const float* uv, real dt, const real vec[2],
                                uint32_t* rng, int jitter, Step* st) {
    const int c = (int)p->fp_contract;
    const real combo = (real)p->combo_factor;
    const real J = (real)p->jitter_magnitude;
    real r[2];
    speed_interp(p, uv, vec[0], vec[1], st->uv1);
    if (jitter) {
        rand_vec(rng, r);
        for (int k = 0; k < 2; k++) st->uv1[k] = MAD(c, r[k], J, st->uv1[k]);
        const real len = fast_length(c, st->uv1[0], st->uv1[1]);
        st->uv1[0] = st->uv1[0] / len;
        st->uv1[1] = st->uv1[1] / len;
    }
    for (int k = 0; k < 2; k++) st->dxy1[k] = approx(p, st->uv1[k] * combo * dt);
    speed_interp(p, uv, vec[0] + st->dxy1[0], vec[1] + st->dxy1[1], st->uv2);
    if (jitter) {
        rand_vec(rng, r);
        for (int k = 0; k < 2; k++) st->uv2[k] = MAD(c, r[k], J, st->uv2[k]);
        const real len = fast_length(c, st->uv2[0], st->uv2[1]);
        st->uv2[0] = st->uv2[0] / len;
        st->uv2[1] = st->uv2[1] / len;
    }
    for (int k = 0; k < 2; k++)
        st->dxy2[k] = approx(p, (real)0.5 * MAD(c, st->uv2[k] * combo, dt, st->dxy1[k]));
    st->next[0] = vec[0] + st->dxy2[0];
    st->next[1] = vec[1] + st->dxy2[1];
    st->idx = array_idx(p, st->next[0], st->next[1]);
}

/* step_error and adaptive dt shared by runge_kutta_step_* (rungekutta.cl:56, :69-70) */
static inline real step_error(const SlmParams* p, const Step* st) {
    const int c = (int)p->fp_contract;
    const real gs = (real)p->grid_scale;
    return fast_length(c, (st->dxy2[0] - st->dxy1[0]) / gs, (st->dxy2[1] - st->dxy1[1]) / gs);
}
static inline real next_dt(const SlmParams* p, real dt, real err) {
    if (err == (real)0) return (real)p->dt_max;
    return r_fmin((real)p->dt_max, ((real)p->adjusted_max_error * dt) / err);
}

/* (uint)(l+0.5f) of writearray.cl:37, :78 */
static inline uint32_t slt_inc(real l) { return cvt_u32(l + (real)0.5); }

#endif /* SLM_ORACLE_MATH_H */
