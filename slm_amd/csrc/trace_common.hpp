// trace_common.hpp -- device-side streamline arithmetic shared by every tracing
// kernel of libslm_hip.so.
//
// Each function restates one helper of the reference's OpenCL C (file:line
// cited per function) in gfx950 HIP.  Rounding is pinned so that results are
// bit-identical to the CPU oracle:
//   * built with -ffp-contract=off; fused multiply-adds appear only through
//     mad<C>() and only where clang's OpenCL default (FP_CONTRACT ON) fuses
//     (a*b+c / a+=b*c in one expression, left product first) when C is true;
//   * division and sqrt are IEEE correctly rounded: plain '/' and
//     __builtin_sqrtf lower to the full-precision gfx950 sequences (HIP's
//     __fsqrt_rn is the native approximate sqrt unless
//     OCML_BASIC_ROUNDED_OPERATIONS is defined, so it is not used), or the
//     short forms below, which give the same bits in their ranges;
//   * fast_length(v) = sqrt(dot(v,v)); fast_normalize(v) = v*(1/sqrt(dot(v,v)));
//   * float->uint / float->char conversions saturate and map NaN to 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/slm_hip.h"

namespace slm {

template <bool C>
__device__ __forceinline__ float mad(float a, float b, float c) {
    if constexpr (C) return __builtin_fmaf(a, b, c);
    else return a * b + c;
}

__device__ __forceinline__ float fdiv(float a, float b) { return a / b; }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_sqrtf(a); }

// Correctly rounded sqrt and divide, short forms.  gfx950 has no correctly
// rounded f32 sqrt or divide instruction; the compiler expands '/' and
// __builtin_sqrtf into IEEE sequences (operand scaling, a hardware estimate
// refined by fma residuals, special-value fix-ups).  The short forms below are
// shorter refinements that return the same bits in their ranges:
//   sqrt_short(d), d in [2^-96, 2^100]: y = v_rsq_f32(d), s = d*y, then one
//       correction s + (d - s*s)*(y/2);
//   recip_short(b), b in [2^-48, 2^50]: one fma Newton step on v_rcp_f32,
//       which is already the correctly rounded 1/b;
//   div_short(a, b, y) with y = recip_short(b) = RN(1/b), |a| in [2^-70, 8b]:
//       q = a*y corrected once, q + (a - b*q)*y (Markstein's theorem: one
//       correction of a quotient from the correctly rounded reciprocal is
//       the correctly rounded quotient).  One reciprocal serves two quotients.
// slm_check_rounding compares them with the operators on the device: every
// float in the sqrt and reciprocal ranges, every dividend significand for
// 2048 divisor significands (half of them the largest, where the first
// quotient estimate is worst), and 6.6e9 hashed quotients.  The guarded
// wrappers below (sqrt_cr, inv_sqrt_cr, div_by_length) use the operators
// outside the ranges; sqrt_short itself is only called inside them.
// Each form trades the compares and selects of the IEEE sequences for
// multiplies, which gfx950 can issue two at a time.
// d in [2^-96, 2^100] as one unsigned compare of the bits: the bit patterns
// of non-negative floats are ordered as the values, and negative values, NaN
// and infinity fall outside (the same result as the two float compares for
// every bit pattern)
__device__ __forceinline__ bool sqrt_short_ok(float d) {
    return __float_as_uint(d) - 0x0f800000u <= 0x71800000u - 0x0f800000u;
}
__device__ __forceinline__ float sqrt_short(float d) {
    const float y = __builtin_amdgcn_rsqf(d);
    const float s = d * y;
    const float h = 0.5f * y;
    const float r = __builtin_fmaf(-s, s, d);
    return __builtin_fmaf(r, h, s);
}
// RN(1/b) for b in [2^-48, 2^50]
__device__ __forceinline__ float recip_short(float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}
// RN(a/b) from y = recip_short(b)
__device__ __forceinline__ float div_short(float a, float b, float y) {
    const float q = a * y;
    const float e = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(e, y, q);
}
// sqrt(d) and 1/sqrt(d), each rounding as the operators do
__device__ __forceinline__ float sqrt_cr(float d) {
    if (__builtin_expect(sqrt_short_ok(d), 1)) return sqrt_short(d);
    return fsqrt(d);
}
__device__ __forceinline__ float inv_sqrt_cr(float d) {
    if (__builtin_expect(sqrt_short_ok(d), 1)) {
        return recip_short(sqrt_short(d));   // the root is in [2^-48, 2^50]
    }
    return fdiv(1.0f, fsqrt(d));
}

template <bool C>
__device__ __forceinline__ float dot2(float x, float y) { return mad<C>(x, x, y * y); }
template <bool C>
__device__ __forceinline__ float fast_length(float x, float y) { return fsqrt(dot2<C>(x, y)); }
template <bool C>
__device__ __forceinline__ float fast_length_cr(float x, float y) { return sqrt_cr(dot2<C>(x, y)); }
// (x, y) / fast_length(x, y): two IEEE quotients by one rounded length
template <bool C>
__device__ __forceinline__ float2 div_by_length(float x, float y) {
    const float d = dot2<C>(x, y);
    if (__builtin_expect(sqrt_short_ok(d) && fabsf(x) >= 0x1p-70f && fabsf(y) >= 0x1p-70f, 1)) {
        const float len = sqrt_short(d);   // in [2^-48, 2^50], and |x|, |y| <= len * (1 + 2^-22)
        const float r = recip_short(len);
        return make_float2(div_short(x, len, r), div_short(y, len, r));
    }
    const float len = fsqrt(d);
    return make_float2(fdiv(x, len), fdiv(y, len));
}

// The hardware float -> integer conversions: v_cvt_u32_f32 / v_cvt_i32_f32
// truncate toward zero, saturate out-of-range values (infinities included) to
// the type's range and map NaN to 0.  C++ leaves those inputs undefined, so the
// instructions are issued directly; one instruction replaces the compare and
// select a defined C++ form needs.
__device__ __forceinline__ uint32_t cvt_u32_sat(float f) {
    uint32_t r;
    __asm__("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}
__device__ __forceinline__ int32_t cvt_i32_sat(float f) {
    int32_t r;
    __asm__("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}

// Pixel coordinate of a position: min(n-1, (uint)max(0, v+PAD_WIDTH_PP5))
// (essentials.cl:26-27, :57-58), the conversion saturating (negative and NaN
// sums give 0, so the max is the conversion's own).
__device__ __forceinline__ uint32_t pix(float v, float pp5, uint32_t nm1) {
    return min(cvt_u32_sat(v + pp5), nm1);
}

// get_array_idx (essentials.cl:56-59)
__device__ __forceinline__ uint32_t array_idx(const SlmParams& P, float x, float y) {
    return pix(y, P.pad_width_pp5, P.ny_padded - 1u) +
           P.ny_padded * pix(x, P.pad_width_pp5, P.nx_padded - 1u);
}

// (int)f: saturating truncation, NaN -> 0
__device__ __forceinline__ int32_t cvt_i32(float f) { return cvt_i32_sat(f); }

// approximate (essentials.cl:104-109): (char)(v*RES)/RES.  For RES = 128 the
// division is an exact scaling, done as a multiply.
__device__ __forceinline__ float approx(float v, float res) {
    const int32_t i = cvt_i32(v * res);
    const float c = (float)(int8_t)(i & 0xff);
    return (res == 128.0f) ? c * 0.0078125f : fdiv(c, res);
}
// approximate with the branch resolved at compile time.  R2: RES is a power of
// two (the host checks), so c/RES == c * (1/RES) exactly (c is an integer in
// [-128, 127] and RES <= 2^31, so nothing underflows) and no division is issued.
template <bool R2>
__device__ __forceinline__ float approx_r(float v, float res, float inv_res) {
    const int32_t i = cvt_i32(v * res);
    const float c = (float)(int8_t)(i & 0xff);
    if constexpr (R2) return c * inv_res;
    else return fdiv(c, res);
}
// approximate of an absolute position (the Euler tails, rungekutta.cl:115,
// :237), where v*RES is usually outside char range.  wrap: the low byte of the
// truncation, as a literal reading of (char) gives; otherwise the int32
// truncation itself (the lattice value), which is what the reference's
// recorded runs computed.
__device__ __forceinline__ float approx_abs(float v, float res, uint32_t wrap) {
    if (wrap) return approx(v, res);
    const float c = (float)cvt_i32(v * res);
    return (res == 128.0f) ? c * 0.0078125f : fdiv(c, res);
}
// compress (essentials.cl:71-74)
__device__ __forceinline__ int8_t compress1(float v, float res) { return (int8_t)(cvt_i32(v * res) & 0xff); }

// lehmer_rand_uint (rng.cl:37-45): ((u64)(uint)(s+1) * 279470273) mod M with
// M = 2^32-5, bit-identical to the 64-bit %.  With p = a*279470273 = hi*2^32 +
// lo, 2^32 == 5 (mod M) gives p == r = 5*hi + lo, formed by one 64-bit
// multiply-add; r < 2^32 + 5*2^29, so r = rhi*2^32 + rlo with rhi in {0, 1}
// and r == u = 5*rhi + rlo (mod M), u < 2^32 (rhi = 1 leaves rlo below
// 5*2^29).  One conditional subtraction of M finishes: u + 5 wraps to u - M
// exactly when u >= M, so min(u, u + 5) is the residue.  Checked against the
// 64-bit % for all 2^32 states.
__device__ __forceinline__ uint32_t lehmer(uint32_t& s) {
    const uint32_t a = s + 1u;                 // wraps in 32 bits first
    const uint64_t p = (uint64_t)a * 279470273u;
    const uint64_t r = (uint64_t)(uint32_t)(p >> 32) * 5u + (uint32_t)p;
    const uint32_t u = (uint32_t)r + 5u * (uint32_t)(r >> 32);
    s = min(u, u + 5u);
    return s;
}

// lehmer_rand_vec (rng.cl:62-67): x drawn first; (float)0xfffffffb == 2^32 so
// the division is an exact scaling by 2^-32.
__device__ __forceinline__ float2 rand_vec(uint32_t& s) {
    const float a = (float)lehmer(s);
    const float b = (float)lehmer(s);
    return make_float2(a * 2.3283064365386963e-10f - 0.5f, b * 2.3283064365386963e-10f - 0.5f);
}

__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// RNG seeding of integratefields.cl:93-99
__device__ __forceinline__ uint32_t rng_seed(uint32_t seed_idx) {
    uint32_t s = bswap32(seed_idx);
    lehmer(s);
    return bswap32(s);
}

// speed_interpolator (essentials.cl:24-45): blend of pixels (xl,yd),(xr,yd),
// (xl,yu),(xr,yu) with xl = round-half-up of x in padded coordinates, weights
// rx = x - xl + PAD in [-0.5,0.5) (an extrapolating blend), then fast_normalize.
template <bool C>
__device__ __forceinline__ float2 speed_interp(const SlmParams& P, const float2* __restrict__ uv,
                                               float x, float y) {
    const uint32_t xl = pix(x, P.pad_width_pp5, P.nx_padded - 1u);
    const uint32_t yd = pix(y, P.pad_width_pp5, P.ny_padded - 1u);
    const uint32_t xr = min(P.nx_padded - 1u, xl + 1u);
    const uint32_t yu = min(P.ny_padded - 1u, yd + 1u);
    const float rx = x - (float)xl + (float)P.pad_width;
    const float lx = 1.0f - rx;
    const float uy = y - (float)yd + (float)P.pad_width;
    const float dy = 1.0f - uy;
    const float2 a = uv[P.ny_padded * xl + yd];
    const float2 b = uv[P.ny_padded * xr + yd];
    const float2 e = uv[P.ny_padded * xl + yu];
    const float2 f = uv[P.ny_padded * xr + yu];
    const float dwn0 = mad<C>(a.x, lx, b.x * rx), dwn1 = mad<C>(a.y, lx, b.y * rx);
    const float upp0 = mad<C>(e.x, lx, f.x * rx), upp1 = mad<C>(e.y, lx, f.y * rx);
    const float s0 = mad<C>(dwn0, dy, upp0 * uy), s1 = mad<C>(dwn1, dy, upp1 * uy);
    const float inv = fdiv(1.0f, fsqrt(dot2<C>(s0, s1)));
    return make_float2(s0 * inv, s1 * inv);
}

// Two neighbouring float2 of one uv row in one 16-byte load (8-byte aligned;
// gfx950 global loads take any dword alignment).
typedef float SlmFloat4A8 __attribute__((ext_vector_type(4), aligned(8)));
__device__ __forceinline__ float4 load_uv_pair(const float2* __restrict__ p) {
    const SlmFloat4A8 v = *reinterpret_cast<const SlmFloat4A8*>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// speed_interp with the same arithmetic (so the same bits): pixels (xl, yd)
// and (xl, yu) are neighbours in memory, as are (xr, yd) and (xr, yu), so two
// 16-byte loads fetch the four corners.  On the last row (yu == yd) each pair
// is read one element lower and its upper half serves both corners.  The
// normalisation uses the short sqrt / divide forms.
template <bool C>
__device__ __forceinline__ float2 speed_interp_cr(const SlmParams& P, const float2* __restrict__ uv,
                                                  float x, float y) {
    const uint32_t top = P.ny_padded - 1u;
    const uint32_t xl = pix(x, P.pad_width_pp5, P.nx_padded - 1u);
    const uint32_t yd = pix(y, P.pad_width_pp5, top);
    const uint32_t xr = min(P.nx_padded - 1u, xl + 1u);
    const bool last = yd >= top;
    const uint32_t y0 = last ? yd - 1u : yd;
    const float4 L = load_uv_pair(uv + (P.ny_padded * xl + y0));
    const float4 R = load_uv_pair(uv + (P.ny_padded * xr + y0));
    const float ax = last ? L.z : L.x, ay = last ? L.w : L.y;
    const float bx = last ? R.z : R.x, by = last ? R.w : R.y;
    const float rx = x - (float)xl + (float)P.pad_width;
    const float lx = 1.0f - rx;
    const float uy = y - (float)yd + (float)P.pad_width;
    const float dy = 1.0f - uy;
    const float dwn0 = mad<C>(ax, lx, bx * rx), dwn1 = mad<C>(ay, lx, by * rx);
    const float upp0 = mad<C>(L.z, lx, R.z * rx), upp1 = mad<C>(L.w, lx, R.w * rx);
    const float s0 = mad<C>(dwn0, dy, upp0 * uy), s1 = mad<C>(dwn1, dy, upp1 * uy);
    const float inv = inv_sqrt_cr(dot2<C>(s0, s1));
    return make_float2(s0 * inv, s1 * inv);
}

// The adaptive step can never leave DT_MAX.  Every dxy component is c/RES with
// c in [-128, 127] (approximate), so a component of dxy2 - dxy1 is below
// 256/RES and step_error <= sqrt(8) * (128/RES) / GRID_SCALE.  When
// GRID_SCALE * ADJUSTED_MAX_ERROR exceeds that bound by a margin that covers
// the roundings of step_error and of next_dt's quotient, ADJ*DT_MAX/err >=
// DT_MAX for every err, so next_dt returns DT_MAX at every step and a trace
// started at DT_MAX keeps it.  Kernels then skip step_error and next_dt.
__host__ __device__ inline bool dt_is_fixed(const SlmParams& p) {
    if (!(p.trajectory_resolution > 0.0f)) return false;
    const double bound = 2.8284271247461903 * (128.0 / (double)p.trajectory_resolution) * (1.0 + 1.0 / 1048576.0);
    return (double)p.grid_scale * (double)p.adjusted_max_error >= bound;
}

struct StepOut {
    float2 uv1, dxy1, dxy2, next;
    uint32_t idx;
};

// compute_step_vec (computestep.cl:42-57) / compute_step_vec_jittered (:86-106).
// R2: TRAJECTORY_RESOLUTION is a power of two and inv_res its reciprocal.
// CR: the paired-load interpolation and the short sqrt / divide forms (the
// same bits).
template <bool C, bool JITTER, bool R2 = false, bool CR = false>
__device__ __forceinline__ void compute_step(const SlmParams& P, const float2* __restrict__ uv,
                                             float dt, float x, float y, uint32_t& rng, StepOut& o,
                                             float inv_res = 0.0f) {
    const float res = P.trajectory_resolution;
    const float combo = P.combo_factor;
    float2 u1 = CR ? speed_interp_cr<C>(P, uv, x, y) : speed_interp<C>(P, uv, x, y);
    if constexpr (JITTER) {
        const float2 r = rand_vec(rng);
        u1.x = mad<C>(r.x, P.jitter_magnitude, u1.x);
        u1.y = mad<C>(r.y, P.jitter_magnitude, u1.y);
        if constexpr (CR) {
            u1 = div_by_length<C>(u1.x, u1.y);
        } else {
            const float len = fast_length<C>(u1.x, u1.y);
            u1.x = fdiv(u1.x, len);
            u1.y = fdiv(u1.y, len);
        }
    }
    o.uv1 = u1;
    o.dxy1 = make_float2(approx_r<R2>(u1.x * combo * dt, res, inv_res), approx_r<R2>(u1.y * combo * dt, res, inv_res));
    float2 u2 = CR ? speed_interp_cr<C>(P, uv, x + o.dxy1.x, y + o.dxy1.y)
                   : speed_interp<C>(P, uv, x + o.dxy1.x, y + o.dxy1.y);
    if constexpr (JITTER) {
        const float2 r = rand_vec(rng);
        u2.x = mad<C>(r.x, P.jitter_magnitude, u2.x);
        u2.y = mad<C>(r.y, P.jitter_magnitude, u2.y);
        if constexpr (CR) {
            u2 = div_by_length<C>(u2.x, u2.y);
        } else {
            const float len = fast_length<C>(u2.x, u2.y);
            u2.x = fdiv(u2.x, len);
            u2.y = fdiv(u2.y, len);
        }
    }
    o.dxy2 = make_float2(approx_r<R2>(0.5f * mad<C>(u2.x * combo, dt, o.dxy1.x), res, inv_res),
                         approx_r<R2>(0.5f * mad<C>(u2.y * combo, dt, o.dxy1.y), res, inv_res));
    o.next = make_float2(x + o.dxy2.x, y + o.dxy2.y);
    o.idx = array_idx(P, o.next.x, o.next.y);
}

// step_error of runge_kutta_step_* (rungekutta.cl:56, :175)
template <bool C>
__device__ __forceinline__ float step_error(const SlmParams& P, const StepOut& o) {
    return fast_length<C>(fdiv(o.dxy2.x - o.dxy1.x, P.grid_scale), fdiv(o.dxy2.y - o.dxy1.y, P.grid_scale));
}
// adaptive dt (rungekutta.cl:69-70, :181-182)
__device__ __forceinline__ float next_dt(const SlmParams& P, float dt, float err) {
    return (err == 0.0f) ? P.dt_max : fminf(P.dt_max, fdiv(P.adjusted_max_error * dt, err));
}
// (uint)(l+0.5f) of writearray.cl:37, :78 (saturating, NaN -> 0)
__device__ __forceinline__ uint32_t slt_inc(float l) { return cvt_u32_sat(l + 0.5f); }
// sub-pixel start offset i*STEP-HALFSPAN (integratefields.cl:108-110)
template <bool C>
__device__ __forceinline__ float sub_offset(const SlmParams& P, uint32_t i) {
    return mad<C>((float)i, P.subpixel_seed_step, -P.subpixel_seed_halfspan);
}


// ------------------------------------------------------------ quad layout
//
// quad[NY_PADDED*x + y] holds the four uv values speed_interpolator reads for
// a position whose pixel is (x, y): (x,y), (xr,y), (x,yu), (xr,yu) with
// xr = min(NX_PADDED-1, x+1), yu = min(NY_PADDED-1, y+1).  The interpolation
// pixel of a position is its get_array_idx pixel, so one aligned 32-byte load
// replaces four scattered float2 loads, and the quad of a step's end pixel is
// fetched together with that pixel's mask byte.
struct alignas(32) Quad {
    float4 lo;  // (x,y).u, (x,y).v, (xr,y).u, (xr,y).v
    float4 hi;  // (x,yu).u, (x,yu).v, (xr,yu).u, (xr,yu).v
};

__device__ __forceinline__ Quad load_quad(const Quad* __restrict__ q, uint32_t idx) {
    Quad r;
    r.lo = q[idx].lo;
    r.hi = q[idx].hi;
    return r;
}

// Global-memory views of pointers a kernel loads from a descriptor.  The
// compiler cannot tell where such a pointer points and issues flat loads, which
// count against the LDS wait counter as well as the vector-memory one: every
// wait for an LDS result then also waits for the flat loads in flight.  The
// buffers are device allocations, so the global forms are exact.
typedef const __attribute__((address_space(1))) Quad* GQuadPtr;
typedef const __attribute__((address_space(1))) uint8_t* GBytePtr;
__device__ __forceinline__ GQuadPtr as_global(const Quad* p) { return (GQuadPtr)p; }
__device__ __forceinline__ GBytePtr as_global(const uint8_t* p) { return (GBytePtr)p; }
typedef float SlmF4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) SlmF4* GF4Ptr;
__device__ __forceinline__ Quad load_quad(GQuadPtr q, uint32_t idx) {
    const GF4Ptr p = (GF4Ptr)(q + idx);
    const SlmF4 a = p[0], b = p[1];
    Quad r;
    r.lo = make_float4(a.x, a.y, a.z, a.w);
    r.hi = make_float4(b.x, b.y, b.z, b.w);
    return r;
}

// speed_interpolator (essentials.cl:24-45) on a quad; same arithmetic as
// speed_interp, so bit-identical.  interp_quad_px takes the position's pixel
// (xl, yd) = (pix(x), pix(y)) when the caller already has it.
template <bool C>
__device__ __forceinline__ float2 interp_quad_px(const SlmParams& P, const Quad& Q, float x, float y, uint32_t xl,
                                                 uint32_t yd) {
    const float rx = x - (float)xl + (float)P.pad_width;
    const float lx = 1.0f - rx;
    const float uy = y - (float)yd + (float)P.pad_width;
    const float dy = 1.0f - uy;
    const float dwn0 = mad<C>(Q.lo.x, lx, Q.lo.z * rx), dwn1 = mad<C>(Q.lo.y, lx, Q.lo.w * rx);
    const float upp0 = mad<C>(Q.hi.x, lx, Q.hi.z * rx), upp1 = mad<C>(Q.hi.y, lx, Q.hi.w * rx);
    const float s0 = mad<C>(dwn0, dy, upp0 * uy), s1 = mad<C>(dwn1, dy, upp1 * uy);
    const float inv = fdiv(1.0f, fsqrt(dot2<C>(s0, s1)));
    return make_float2(s0 * inv, s1 * inv);
}
template <bool C>
__device__ __forceinline__ float2 interp_quad(const SlmParams& P, const Quad& Q, float x, float y) {
    return interp_quad_px<C>(P, Q, x, y, pix(x, P.pad_width_pp5, P.nx_padded - 1u),
                             pix(y, P.pad_width_pp5, P.ny_padded - 1u));
}

// Jitter of compute_step_vec_jittered (computestep.cl:96-98): add the scaled
// variate, renormalise.
template <bool C>
__device__ __forceinline__ float2 jitter_unit(const SlmParams& P, float2 u, uint32_t& rng) {
    const float2 r = rand_vec(rng);
    u.x = mad<C>(r.x, P.jitter_magnitude, u.x);
    u.y = mad<C>(r.y, P.jitter_magnitude, u.y);
    const float len = fast_length<C>(u.x, u.y);
    return make_float2(fdiv(u.x, len), fdiv(u.y, len));
}

// interp_quad_px and jitter_unit with the short sqrt / divide forms where
// their inputs allow (same bits).  The fields tracers (k_fields,
// k_fields_batch) use these; the operator forms above serve the kernels whose
// steps are a small part of their time.
template <bool C>
__device__ __forceinline__ float2 interp_quad_px_cr(const SlmParams& P, const Quad& Q, float x, float y, uint32_t xl,
                                                    uint32_t yd) {
    const float rx = x - (float)xl + (float)P.pad_width;
    const float lx = 1.0f - rx;
    const float uy = y - (float)yd + (float)P.pad_width;
    const float dy = 1.0f - uy;
    const float dwn0 = mad<C>(Q.lo.x, lx, Q.lo.z * rx), dwn1 = mad<C>(Q.lo.y, lx, Q.lo.w * rx);
    const float upp0 = mad<C>(Q.hi.x, lx, Q.hi.z * rx), upp1 = mad<C>(Q.hi.y, lx, Q.hi.w * rx);
    const float s0 = mad<C>(dwn0, dy, upp0 * uy), s1 = mad<C>(dwn1, dy, upp1 * uy);
    const float inv = inv_sqrt_cr(dot2<C>(s0, s1));
    return make_float2(s0 * inv, s1 * inv);
}
template <bool C>
__device__ __forceinline__ float2 jitter_unit_cr(const SlmParams& P, float2 u, uint32_t& rng) {
    const float2 r = rand_vec(rng);
    u.x = mad<C>(r.x, P.jitter_magnitude, u.x);
    u.y = mad<C>(r.y, P.jitter_magnitude, u.y);
    return div_by_length<C>(u.x, u.y);
}


// A deterministic trace whose state (position, dt, and any carried label)
// repeats is periodic from there on: every exit test of the loop is a function
// of that state and fails around the whole cycle, so the trace runs on to
// MAX_N_STEPS.  Brent's algorithm finds the period lam; whole laps are then
// skipped (n += laps*lam) and the last partial lap is stepped as before, so
// the trace ends at the same step, pixel and label.  Positions live on the
// 1/TRAJECTORY_RESOLUTION lattice, so traces caught in pits and flats do cycle.
struct CycleSkip {
    float sx, sy, sdt;
    uint32_t sown, sn, power;
    bool on;
    __device__ __forceinline__ void reset() {
        sx = sy = sdt = -1.0f;
        sown = 0u;
        sn = 0u;
        power = 1u;
        on = true;
    }
    // true when the state repeats; n then jumps over the whole laps to cap
    __device__ __forceinline__ bool check(float x, float y, float dt, uint32_t own, uint32_t& n, uint32_t cap) {
        if (!on) return false;
        if (x == sx && y == sy && dt == sdt && own == sown && n > sn) {
            const uint32_t lam = n - sn;
            n += ((cap - n) / lam) * lam;
            on = false;
            return true;
        }
        if (n - sn >= power) {
            sx = x;
            sy = y;
            sdt = dt;
            sown = own;
            sn = n;
            power <<= 1;
        }
        return false;
    }
};

// Persistent lane refill for per-seed traces.  One thread per seed leaves a
// wave running as long as its longest trace, and these traces run from 1 to
// MAX_N_STEPS steps.  Here a wave takes 64-seed chunks g, g + G, g + 2G, ...
// (g its index, G the number of waves; no shared counter, so calls on
// different streams never interact) and a lane whose trace ends starts the
// next seed of the chunk at once.  start(k) sets a lane up for seed k and
// returns whether it has a trace to run; step() advances it one iteration
// and returns false once the trace has finished (and written its result).
// Every trace is independent, so the order changes no result.
// A wave refills its idle lanes only when at least kRefillMin of them are idle
// (or none is busy).  A refill costs the wave a queue grab, ballots and the
// new traces' first loads, whatever the number of lanes it serves; refilling
// every time one lane ends made that cost a large share of the loop (the
// fields tracers ran 14-18% slower).  Lanes idle meanwhile waste about
// kRefillMin / 128 of the issue slots.  Which lane runs which trace, and
// when, changes no result.
constexpr uint32_t kRefillMin = 20;

// The fields tracers' refill batch for d x d sub-seeds per pixel, measured on
// a 2048^2 tile (the same traces, 282 steps on average, at d = 1 and d = 2):
// 28 lanes at d = 1 (8 lanes: 15% slower, 20: 4%), 20 lanes from d = 2 on
// (8 lanes: 2.5% slower, 24: 5%, 28: 3%).  Neighbouring work items of one
// pixel trace almost the same path and tend to end together, which moves the
// best batch down.
__host__ __device__ constexpr uint32_t fields_refill_batch(uint32_t d) { return d <= 1u ? 28u : 20u; }

template <typename Start, typename Step>
__device__ __forceinline__ void persist_seeds(uint32_t first, uint32_t n_items, Start start, Step step) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    const uint32_t G = gridDim.x * (blockDim.x >> 6);
    uint32_t chunk = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint32_t c_next = chunk < (n_items + 63u) / 64u ? chunk * 64u : n_items;
    uint32_t c_end = min(n_items, c_next + 64u);
    bool active = false;
    while (true) {
        uint64_t need = __ballot(!active);
        if (__popcll(need) < kRefillMin && __any(active)) need = 0;
        while (need && c_next < n_items) {
            if (c_next >= c_end) {
                chunk += G;
                if (chunk >= (n_items + 63u) / 64u) {
                    c_next = n_items;
                    break;
                }
                c_next = chunk * 64u;
                c_end = min(n_items, c_next + 64u);
            }
            const uint32_t take = min((uint32_t)__popcll(need), c_end - c_next);
            const uint32_t rank = __popcll(need & lt);
            const bool mine = ((need >> lane) & 1ull) && rank < take;
            if (mine) active = start(first + c_next + rank);
            need &= ~__ballot(mine);
            c_next += take;
        }
        if (!__any(active)) {
            if (c_next >= n_items) break;
            continue;
        }
        if (active) active = step();
    }
}

}  // namespace slm
