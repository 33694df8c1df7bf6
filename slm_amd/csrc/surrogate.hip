// surrogate.hip -- kernels of the BF16 convolutional surrogate of the fields.
//
// A U-Net learns log(1+slt), log(1+slc) and a channel mask from the DTM, on
// fields the tracer produced; it has no counterpart in the reference.
// Activations are NHWC bf16 (channels contiguous, a pixel's channels at
// base + pixel*ld so concatenations are channel slices of one buffer);
// weights are bf16 [Cout][taps][Cin] with fp32 master copies; every product
// accumulates in fp32 on MFMA (v_mfma_f32_32x32x16_bf16).
//
//   conv (3x3 pad 1, or 1x1)  implicit GEMM: M = pixels, N = Cout, K = taps*Cin;
//                             bias and ReLU fused into the epilogue
//   conv backward-data        the same kernel on dY with weights flipped and
//                             transposed (k_wflip)
//   conv backward-weight      dW = dY^T X: pixels are the reduction; on
//                             16..128-column images dY and X (with its halo)
//                             are staged once per chunk of rows (k_wgrad_halo),
//                             other shapes stage tiles of dY and im2col(X)
//                             (k_wgrad_t)
//   maxpool 2x2, nearest upsample 2x, ReLU mask, add, loss, Adam
//
// MFMA 32x32x16 operand layout (gfx950, checked on the device): lane l holds
// A[l%32][8*(l/32) .. +8], B[8*(l/32) .. +8][l%32]; the 16 results of lane l
// are D[(v/4)*8 + 4*(l/32) + v%4][l%32], v = 0..15.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>

#include "common.hpp"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float b2f(bf16 v) { return (float)v; }
__device__ __forceinline__ bf16 f2b(float v) { return (bf16)v; }   // round to nearest even

// ------------------------------------------------------------------ convolution

// One wave computes a 32-pixel x 32-channel tile; a block has WM x WN waves.
// Input channels [0, c_up) come from x2, a half-resolution image read at
// (h/2, w/2) (a fused nearest 2x upsample); the rest from x (c_up = 0: x only).
template <int TAPS, bool RELU, int WM, int WN>
__global__ __launch_bounds__(256) void k_conv(const bf16* __restrict__ x, int ldx, const bf16* __restrict__ x2,
                                              int ldx2, int c_up, const bf16* __restrict__ w,
                                              const float* __restrict__ bias, bf16* __restrict__ y, int ldy,
                                              int M, int H, int W, int Cin, int Cout) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pm = blockIdx.x * (WM * 32) + (wid % WM) * 32;
    const int pc = blockIdx.y * (WN * 32) + (wid / WM) * 32;
    if (pm >= M || pc >= Cout) return;             // wave-uniform
    const int r = lane & 31, kg = (lane >> 5) * 8;
    const int p = pm + r;
    const int pw = p % W, ph = (p / W) % H, pn = p / (W * H);
    const int co = pc + r;
    const bf16* wrow = w + (size_t)co * TAPS * Cin;
    f32x16 acc = {};
    const bf16x8 zero = {};
#pragma unroll 1
    for (int tap = 0; tap < TAPS; tap++) {
        const int hh = ph + (TAPS == 9 ? tap / 3 - 1 : 0), ww = pw + (TAPS == 9 ? tap % 3 - 1 : 0);
        const bool ok = hh >= 0 && hh < H && ww >= 0 && ww < W;
        const bf16* xp = x + (size_t)((pn * H + (ok ? hh : 0)) * W + (ok ? ww : 0)) * ldx + kg;
        const bf16* xp2 =
            x2 + (size_t)((pn * (H / 2) + (ok ? hh : 0) / 2) * (W / 2) + (ok ? ww : 0) / 2) * ldx2 + kg;
        const bf16* wp = wrow + tap * Cin + kg;

        for (int c0 = 0; c0 < Cin; c0 += 16) {
            const bf16x8 a = ok ? *(const bf16x8*)(c0 < c_up ? xp2 + c0 : xp + c0) : zero;
            const bf16x8 b = *(const bf16x8*)(wp + c0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
    }
    const float bc = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        float val = acc[v] + bc;
        if (RELU) val = fmaxf(val, 0.0f);
        y[(size_t)(pm + rr) * ldy + co] = f2b(val);
    }
}

// 3x3 convolution with LDS-staged tiles.  A block computes a BH x BW pixel
// tile of one image (BW = 32 or 16 columns) for CT output channels.  For
// each chunk of KC input channels it stages the (BH+2) x (BW+2) input halo
// and the CT x 9 x KC weights in LDS (rows padded by 8 bf16 against bank
// conflicts); then every tap reads its shifted A fragments from the halo, so
// each input element comes from L2 once per chunk instead of nine times.
// Wave w owns pixel groups w*RW .. w*RW+RW-1 (a group is 32 pixels: one
// 32-wide row, or two 16-wide rows) and all CT channels: RW x CT/32
// accumulator tiles of 32x32.
// Input channels [0, c_up) come from x2 at half resolution, as in k_conv.
// The next input chunk's halo and weights are loaded into registers while
// the current chunk computes from LDS.
template <int BW, int RW, int CT, int KC, bool RELU>
__global__ __launch_bounds__(256, 2) void k_conv_lds(const bf16* __restrict__ x, int ldx, const bf16* __restrict__ x2,
                                                     int ldx2, int c_up, const bf16* __restrict__ w,
                                                     const float* __restrict__ bias, bf16* __restrict__ y, int ldy,
                                                     int H, int W, int Cin, int Cout) {
    constexpr int GR = 32 / BW;                 // rows per pixel group
    constexpr int BH = 4 * RW * GR;             // block tile rows
    constexpr int XP = KC + 8;                  // halo pixel stride (elements)
    constexpr int WP = 9 * KC + 8;              // weight row stride (elements)
    constexpr int NJ = CT / 32;
    __shared__ __attribute__((aligned(16))) bf16 xs[(BH + 2) * (BW + 2) * XP];
    __shared__ __attribute__((aligned(16))) bf16 ws[CT * WP];
    const int tiles_w = W / BW, tiles_h = H / BH;
    const int tw = blockIdx.x % tiles_w, th = (blockIdx.x / tiles_w) % tiles_h, n = blockIdx.x / (tiles_w * tiles_h);
    const int h0 = th * BH, w0 = tw * BW, co0 = blockIdx.y * CT;
    const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
    const int r = lane & 31, kg = (lane >> 5) * 8;
    f32x16 acc[RW][NJ];
#pragma unroll
    for (int i = 0; i < RW; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[i][j] = f32x16{};
    // this lane's pixel (row, column) in the block tile, per group
    int prow[RW], pcol[RW];
#pragma unroll
    for (int i = 0; i < RW; i++) {
        const int g = wid * RW + i;
        prow[i] = g * GR + r / BW;
        pcol[i] = r % BW;
    }
    const bf16x8 zero = {};
    // staging: the halo ((BH+2) x (BW+2) pixels x KC channels) and the weights
    // (CT x 9 x KC) of one input chunk, in 16-byte vectors, NXV + NWV per
    // thread; the next chunk's vectors are loaded into registers while the
    // current chunk computes from LDS
    constexpr int XV = (BH + 2) * (BW + 2) * (KC / 8);
    constexpr int WV = CT * 9 * (KC / 8);
    constexpr int NXV = (XV + 255) / 256, NWV = (WV + 255) / 256;
    bf16x8 xr[NXV], wr[NWV];
    auto load_chunk = [&](int c0) {
#pragma unroll
        for (int i = 0; i < NXV; i++) {
            const int v = t + i * 256;
            bf16x8 val = zero;
            if (v < XV) {
                const int px = v / (KC / 8), part = v % (KC / 8);
                const int hr = px / (BW + 2), hc = px % (BW + 2);
                const int gh = h0 + hr - 1, gw = w0 + hc - 1;
                if (gh >= 0 && gh < H && gw >= 0 && gw < W) {
                    const int c = c0 + part * 8;
                    val = c < c_up
                              ? *(const bf16x8*)(x2 + ((size_t)(n * (H / 2) + gh / 2) * (W / 2) + gw / 2) * ldx2 + c)
                              : *(const bf16x8*)(x + ((size_t)(n * H + gh) * W + gw) * ldx + c);
                }
            }
            xr[i] = val;
        }
#pragma unroll
        for (int i = 0; i < NWV; i++) {
            const int v = t + i * 256;
            if (v < WV) {
                const int co = v / (9 * KC / 8), rem = v % (9 * KC / 8);
                const int tap = rem / (KC / 8), part = rem % (KC / 8);
                wr[i] = *(const bf16x8*)(w + ((size_t)(co0 + co) * 9 + tap) * Cin + c0 + part * 8);
            }
        }
    };
    auto store_chunk = [&]() {
#pragma unroll
        for (int i = 0; i < NXV; i++) {
            const int v = t + i * 256;
            if (v < XV) {
                const int px = v / (KC / 8), part = v % (KC / 8);
                *(bf16x8*)&xs[px * XP + part * 8] = xr[i];
            }
        }
#pragma unroll
        for (int i = 0; i < NWV; i++) {
            const int v = t + i * 256;
            if (v < WV) {
                const int co = v / (9 * KC / 8), rem = v % (9 * KC / 8);
                const int tap = rem / (KC / 8), part = rem % (KC / 8);
                *(bf16x8*)&ws[co * WP + tap * KC + part * 8] = wr[i];
            }
        }
    };
    load_chunk(0);
    for (int c0 = 0; c0 < Cin; c0 += KC) {
        __syncthreads();                      // the previous chunk's reads of LDS are done
        store_chunk();
        __syncthreads();
        if (c0 + KC < Cin) load_chunk(c0 + KC);
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int kk = 0; kk < KC; kk += 16) {
                bf16x8 b[NJ];
#pragma unroll
                for (int j = 0; j < NJ; j++) b[j] = *(const bf16x8*)&ws[(j * 32 + r) * WP + tap * KC + kk + kg];
#pragma unroll
                for (int i = 0; i < RW; i++) {
                    const bf16x8 a =
                        *(const bf16x8*)&xs[((prow[i] + dy) * (BW + 2) + pcol[i] + dx) * XP + kk + kg];
#pragma unroll
                    for (int j = 0; j < NJ; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[j], acc[i][j], 0, 0, 0);
                }
            }
        }
    }
    // epilogue: D row = pixel within the group, column = channel
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int co = co0 + j * 32 + r;
        const float bc = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < RW; i++) {
            const int g = wid * RW + i;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int rr = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                const int row = g * GR + rr / BW, col = rr % BW;
                float val = acc[i][j][v] + bc;
                if (RELU) val = fmaxf(val, 0.0f);
                y[((size_t)(n * H + h0 + row) * W + w0 + col) * ldy + co] = f2b(val);
            }
        }
    }
}

// wt[ci][tap'][co] = w[co][8 - tap'][ci] (3x3) or w[co][0][ci] (1x1): the
// weights of the backward-data convolution.
__global__ void k_wflip(const bf16* __restrict__ w, bf16* __restrict__ wt, int taps, int Cin, int Cout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = Cout * taps * Cin;
    if (i >= n) return;
    const int ci = i % Cin, tap = (i / Cin) % taps, co = i / (Cin * taps);
    const int tflip = taps - 1 - tap;
    wt[((size_t)ci * taps + tflip) * Cout + co] = w[i];
}

// dW[co][tap*Cin+ci] += sum over pixels of dY[p][co] * X[p + off(tap)][ci];
// block: 64 channels x 64 k, pixels [blockIdx.z*ppb, +ppb) in chunks of 64.
template <int TAPS>
__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ x, int ldx, const bf16* __restrict__ dy,
                                               int lddy, float* __restrict__ dw, int M, int H, int W, int Cin,
                                               int Cout, int ppb) {
    __shared__ __attribute__((aligned(16))) bf16 dyT[64][64 + 8];   // [co][p]
    __shared__ __attribute__((aligned(16))) bf16 xT[64][64 + 8];    // [k][p]
    const int K = TAPS * Cin;
    const int c0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    const int p_begin = blockIdx.z * ppb, p_end = min(M, p_begin + ppb);
    const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
    const int wc = (wid & 1) * 32, wk = (wid >> 1) * 32;
    const bool wave_on = (c0 + wc < Cout) && (k0 + wk < K);
    f32x16 acc = {};
    for (int pb = p_begin; pb < p_end; pb += 64) {
        // stage: 64 pixels x 64 channels of dY, 64 pixels x 64 k of im2col(X), transposed
        for (int v = t; v < 512; v += 256) {
            const int pl = v >> 3, cv = (v & 7) * 8;
            const int p = pb + pl;
            bf16x8 a = {}, b = {};
            if (p < p_end) {
                if (c0 + cv < Cout) a = *(const bf16x8*)(dy + (size_t)p * lddy + c0 + cv);
                const int k = k0 + cv;
                if (k < K) {
                    const int tap = k / Cin, ci = k % Cin;
                    const int pw = p % W, ph = (p / W) % H, pn = p / (W * H);
                    const int hh = ph + (TAPS == 9 ? tap / 3 - 1 : 0), ww = pw + (TAPS == 9 ? tap % 3 - 1 : 0);
                    if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                        b = *(const bf16x8*)(x + (size_t)((pn * H + hh) * W + ww) * ldx + ci);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                dyT[cv + j][pl] = a[j];
                xT[cv + j][pl] = b[j];
            }
        }
        __syncthreads();
        if (wave_on) {
            const int r = lane & 31, kg = (lane >> 5) * 8;
#pragma unroll
            for (int kk = 0; kk < 64; kk += 16) {
                const bf16x8 a = *(const bf16x8*)&dyT[wc + r][kk + kg];
                const bf16x8 b = *(const bf16x8*)&xT[wk + r][kk + kg];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (!wave_on) return;
    const int kcol = k0 + wk + (lane & 31);
    if (kcol >= K) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int co = c0 + wc + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        if (co < Cout) atomicAdd(&dw[(size_t)co * K + kcol], acc[v]);
    }
}

// dW[co][k] += sum over pixels of dY[p][co] * im2col(X)[p][k], k = tap*Cin + ci.
// Block tile CT channels x KT k (4 waves of 32x32), pixels [blockIdx.z*ppb,
// +ppb) in chunks of 128.  Staging: a thread loads an 8-pixel x 8-channel
// block (eight 16-byte loads), transposes it in registers and stores eight
// 16-byte rows [channel][8 pixels] to LDS, so the MFMA operands (8 pixels
// of one channel per lane) are single 16-byte LDS reads.
template <int TAPS, int CT, int KT>
__global__ __launch_bounds__(256) void k_wgrad_t(const bf16* __restrict__ x, int ldx, const bf16* __restrict__ dy,
                                                 int lddy, float* __restrict__ dw, int M, int H, int W, int Cin,
                                                 int Cout, int ppb) {
    constexpr int PC = 128, PP = PC + 8;
    __shared__ __attribute__((aligned(16))) bf16 dyT[CT * PP];
    __shared__ __attribute__((aligned(16))) bf16 xT[KT * PP];
    const int K = TAPS * Cin;
    const int c0 = blockIdx.x * CT, k0 = blockIdx.y * KT;
    const int p_begin = blockIdx.z * ppb, p_end = min(M, p_begin + ppb);
    const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
    constexpr int WCO = CT / 32;                 // waves along channels
    const int wc = (wid % WCO) * 32, wk = (wid / WCO) * 32;
    const bool wave_on = (c0 + wc < Cout) && (k0 + wk < K);
    constexpr int NB_DY = (PC / 8) * (CT / 8), NB = NB_DY + (PC / 8) * (KT / 8);
    f32x16 acc = {};
    const bf16x8 zero = {};
    for (int pb = p_begin; pb < p_end; pb += PC) {
        for (int blk = t; blk < NB; blk += 256) {
            const bool is_dy = blk < NB_DY;
            const int b2 = is_dy ? blk : blk - NB_DY;
            const int pg = b2 % (PC / 8), cg = b2 / (PC / 8);
            bf16x8 v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int p = pb + pg * 8 + i;
                v[i] = zero;
                if (p < p_end) {
                    if (is_dy) {
                        if (c0 + cg * 8 < Cout) v[i] = *(const bf16x8*)(dy + (size_t)p * lddy + c0 + cg * 8);
                    } else {
                        const int k = k0 + cg * 8;
                        if (k < K) {
                            const int tap = k / Cin, ci = k % Cin;
                            const int pw = p % W, ph = (p / W) % H, pn = p / (W * H);
                            const int hh = ph + (TAPS == 9 ? tap / 3 - 1 : 0), ww = pw + (TAPS == 9 ? tap % 3 - 1 : 0);
                            if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                                v[i] = *(const bf16x8*)(x + (size_t)((pn * H + hh) * W + ww) * ldx + ci);
                        }
                    }
                }
            }
            bf16* dst = is_dy ? dyT : xT;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                bf16x8 col;
#pragma unroll
                for (int i = 0; i < 8; i++) col[i] = v[i][j];
                *(bf16x8*)&dst[(cg * 8 + j) * PP + pg * 8] = col;
            }
        }
        __syncthreads();
        if (wave_on) {
            const int r = lane & 31, kg = (lane >> 5) * 8;
#pragma unroll
            for (int kk = 0; kk < PC; kk += 16) {
                const bf16x8 a = *(const bf16x8*)&dyT[(wc + r) * PP + kk + kg];
                const bf16x8 b = *(const bf16x8*)&xT[(wk + r) * PP + kk + kg];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (!wave_on) return;
    const int kcol = k0 + wk + (lane & 31);
    if (kcol >= K) return;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int co = c0 + wc + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        if (co < Cout) atomicAdd(&dw[(size_t)co * K + kcol], acc[v]);
    }
}

// Weight gradient with one LDS copy of X per pixel chunk (3x3 or 1x1).  A chunk is
// R = 256 / W whole rows of one image (W = 16 .. 128 columns).  Staged
// channel-major, each 8-pixel x 8-channel block read with eight 16-byte loads
// and written transposed as eight 16-byte rows:
//   dyT [CT][256 px]                               the A operands
//   xh  [CIT][R + 2 rows][8 zeros | W | 8 zeros]   X with its row and column halo
// so X leaves L2 once per chunk (plus its two halo rows) instead of once per
// tap.  Wave (cw, dh) owns output channels [32 cw, +32) and the three taps of
// row offset dh - 1: per 16-pixel step it reads its A operand, the aligned
// 8 columns of its B operand and the words either side, and forms the column
// offsets -1 and +1 by shifting bf16 pairs across neighbouring words
// (v_alignbyte), so three MFMAs share one set of aligned LDS reads.  The
// plane strides are 8 (mod 128) elements, so the 32 rows of an operand read
// start on different 16-byte bank groups.  Chunks [blockIdx.z * cpb, +cpb)
// of the N * ceil(H / R), the next one loaded into registers while the
// current one computes; partial sums leave by fp32 atomics.  For 1x1
// (T = 1) there is no halo: one wave per 32 output channels, one MFMA per step.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int T, int W, int CT, int CIT>
__global__ __launch_bounds__((T == 9 ? 192 : 64) * (CT / 32)) void k_wgrad_halo(
    const bf16* __restrict__ x, int ldx, const bf16* __restrict__ dy, int lddy, float* __restrict__ dw, int N, int H,
    int Cin, int cpb) {
    constexpr int R = 256 / W;
    constexpr int HALO = T == 9 ? 1 : 0;                // rows of halo above and below
    constexpr int WP = W + 16;
    constexpr int XROW = (R + 2 * HALO) * WP;
    constexpr int XS = XROW + ((8 - XROW % 128) + 128) % 128;
    constexpr int DS = 256 + 8;
    constexpr int NT = (T == 9 ? 192 : 64) * (CT / 32);
    constexpr int XG = W / 8 + 2;                       // staged 8-column groups per row
    static_assert(T == 1 || T == 9, "taps");
    static_assert(W % 16 == 0 && W >= 16 && W <= 256, "W");
    static_assert(XS % 128 == 8 && DS % 128 == 8, "strides");
    __shared__ __attribute__((aligned(16))) bf16 dyT[CT * DS];
    __shared__ __attribute__((aligned(16))) bf16 xh[CIT * XS];
    const int K = T * Cin;
    const int c0 = blockIdx.x * CT, ci0 = blockIdx.y * CIT;
    const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
    const int cw = T == 9 ? wid / 3 : wid, dh = T == 9 ? wid % 3 : 0;
    const int r = lane & 31, half = lane >> 5;
    const int per_image = (H + R - 1) / R;
    const int ch_end = min(N * per_image, (int)(blockIdx.z + 1) * cpb);
    const bool b_on = CIT == 32 || r < CIT;
    f32x16 acc0 = {}, acc1 = {}, acc2 = {};
    const bf16x8 zero = {};
    // staging blocks (8 pixels x 8 channels) per thread; the next chunk's are
    // loaded into registers while the current chunk computes from LDS
    constexpr int NBD = 32 * (CT / 8), NBX = (R + 2 * HALO) * XG * (CIT / 8);
    constexpr int ND = (NBD + NT - 1) / NT, NX = (NBX + NT - 1) / NT;
    bf16x8 dv[ND][8], xv[NX][8];
    auto load_chunk = [&](int ch) {
        const int n = ch / per_image, r0 = (ch % per_image) * R;
#pragma unroll
        for (int k = 0; k < ND; k++) {
            const int blk = t + k * NT;
            const int cg = blk % (CT / 8), q = (blk / (CT / 8)) * 8;
            const int row = q / W, col = q % W;
            const bool in = blk < NBD && r0 + row < H;
            const bf16* src = dy + ((size_t)(n * H + r0 + row) * W + col) * lddy + c0 + cg * 8;
#pragma unroll
            for (int i = 0; i < 8; i++) dv[k][i] = in ? *(const bf16x8*)(src + (size_t)i * lddy) : zero;
        }
#pragma unroll
        for (int k = 0; k < NX; k++) {
            const int blk = t + k * NT;
            const int cg = blk % (CIT / 8), g = (blk / (CIT / 8)) % XG, rr = blk / ((CIT / 8) * XG);
            const int row = r0 - HALO + rr, col = (g - 1) * 8;
            const bool in = blk < NBX && row >= 0 && row < H && g >= 1 && g <= W / 8;
            if (in) {
                const bf16* src = x + ((size_t)(n * H + row) * W + col) * ldx + ci0 + cg * 8;
#pragma unroll
                for (int i = 0; i < 8; i++) xv[k][i] = *(const bf16x8*)(src + (size_t)i * ldx);
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) xv[k][i] = zero;
            }
        }
    };
    auto store_chunk = [&]() {
#pragma unroll
        for (int k = 0; k < ND; k++) {
            const int blk = t + k * NT;
            if (blk >= NBD) continue;
            const int cg = blk % (CT / 8), q = (blk / (CT / 8)) * 8;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                bf16x8 c;
#pragma unroll
                for (int i = 0; i < 8; i++) c[i] = dv[k][i][j];
                *(bf16x8*)&dyT[(cg * 8 + j) * DS + q] = c;
            }
        }
#pragma unroll
        for (int k = 0; k < NX; k++) {
            const int blk = t + k * NT;
            if (blk >= NBX) continue;
            const int cg = blk % (CIT / 8), g = (blk / (CIT / 8)) % XG, rr = blk / ((CIT / 8) * XG);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                bf16x8 c;
#pragma unroll
                for (int i = 0; i < 8; i++) c[i] = xv[k][i][j];
                *(bf16x8*)&xh[(cg * 8 + j) * XS + rr * WP + g * 8] = c;
            }
        }
    };
    const int ch_begin = blockIdx.z * cpb;
    if (ch_begin < ch_end) load_chunk(ch_begin);
    for (int ch = ch_begin; ch < ch_end; ++ch) {
        __syncthreads();                      // the previous chunk's reads of LDS are done
        store_chunk();
        __syncthreads();
        if (ch + 1 < ch_end) load_chunk(ch + 1);
#pragma unroll 4
        for (int ks = 0; ks < 16; ++ks) {
            const int q = ks * 16 + half * 8, row = q / W, col = q % W;
            const bf16x8 a = *(const bf16x8*)&dyT[(cw * 32 + r) * DS + q];
            u32x4 g = {0u, 0u, 0u, 0u};
            uint32_t prev = 0u, next = 0u;
            if (b_on) {
                const bf16* bp = &xh[r * XS + (row + dh) * WP + col + 8];
                g = *(const u32x4*)bp;
                if (T == 9) {
                    prev = *(const uint32_t*)(bp - 2);
                    next = *(const uint32_t*)(bp + 8);
                }
            }
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, g), acc1, 0, 0, 0);
            if (T == 9) {
                const u32x4 gm = {__builtin_amdgcn_alignbyte(g[0], prev, 2), __builtin_amdgcn_alignbyte(g[1], g[0], 2),
                                  __builtin_amdgcn_alignbyte(g[2], g[1], 2), __builtin_amdgcn_alignbyte(g[3], g[2], 2)};
                const u32x4 gp = {__builtin_amdgcn_alignbyte(g[1], g[0], 2), __builtin_amdgcn_alignbyte(g[2], g[1], 2),
                                  __builtin_amdgcn_alignbyte(g[3], g[2], 2), __builtin_amdgcn_alignbyte(next, g[3], 2)};
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, gm), acc0, 0, 0, 0);
                acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, gp), acc2, 0, 0, 0);
            }
        }
    }
    if (!b_on) return;
    const int ci = ci0 + r;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int co = c0 + cw * 32 + (v >> 2) * 8 + half * 4 + (v & 3);
        float* d = dw + (size_t)co * K + (size_t)(dh * 3) * Cin + ci;
        if (T == 9) {
            atomicAdd(d, acc0[v]);
            atomicAdd(d + Cin, acc1[v]);
            atomicAdd(d + 2 * Cin, acc2[v]);
        } else {
            atomicAdd(d, acc1[v]);
        }
    }
}

// db[co] += sum over pixels of dY[p][co].  A thread sums 8 channels (one
// 16-byte load per pixel) over a strided set of pixels; the block's 256/(C/8)
// pixel lanes are reduced in LDS and each channel gets one atomic per block
// (the launch keeps the block count small: the atomics share Cout addresses).
__global__ __launch_bounds__(256) void k_bgrad(const bf16* __restrict__ dy, int lddy, float* __restrict__ db,
                                               int M, int Cout) {
    __shared__ float red[256 * 8];
    const int C8 = Cout / 8;                    // 4 .. 32: divides 256
    const int t = threadIdx.x, tc = t % C8, tp = t / C8, P = 256 / C8;
    float acc[8] = {};
#pragma unroll 4
    for (size_t p = (size_t)blockIdx.x * P + tp; p < (size_t)M; p += (size_t)gridDim.x * P) {
        const bf16x8 v = *(const bf16x8*)(dy + p * lddy + tc * 8);
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] += b2f(v[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) red[t * 8 + j] = acc[j];
    __syncthreads();
    if (tp == 0) {
        float sum[8] = {};
        for (int q = 0; q < P; q++)
#pragma unroll
            for (int j = 0; j < 8; j++) sum[j] += red[(q * C8 + tc) * 8 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) atomicAdd(&db[tc * 8 + j], sum[j]);
    }
}

// ------------------------------------------------------------------ elementwise

// Fused ReLU backward and bias gradient: dY *= (Y > 0) when y is given (the
// stored post-ReLU output), then the column sums of dY.  A thread handles 8
// channels (16-byte accesses) over strided pixels; lanes are reduced in LDS.
// Each block's C sums go to row blockIdx.x of partials when given (summed by
// k_colsum), else straight into db by atomics (few blocks then: the atomics
// of every block share C addresses).
__global__ __launch_bounds__(256) void k_relu_bias(const bf16* __restrict__ y, int ldy, bf16* __restrict__ dy,
                                                   int lddy, float* __restrict__ db, float* __restrict__ partials,
                                                   int M, int C) {
    __shared__ float red[256 * 8];
    const int C8 = C / 8;                       // 4 .. 32: divides 256
    const int t = threadIdx.x, tc = t % C8, tp = t / C8, P = 256 / C8;
    float acc[8] = {};
#pragma unroll 4
    for (size_t p = (size_t)blockIdx.x * P + tp; p < (size_t)M; p += (size_t)gridDim.x * P) {
        bf16x8 g = *(const bf16x8*)(dy + p * lddy + tc * 8);
        if (y) {
            const bf16x8 yv = *(const bf16x8*)(y + p * ldy + tc * 8);
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (!(b2f(yv[j]) > 0.0f)) g[j] = f2b(0.0f);
            *(bf16x8*)(dy + p * lddy + tc * 8) = g;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) acc[j] += b2f(g[j]);
    }
    if (!db) return;                            // uniform
#pragma unroll
    for (int j = 0; j < 8; j++) red[t * 8 + j] = acc[j];
    __syncthreads();
    if (tp == 0) {
        float sum[8] = {};
        for (int q = 0; q < P; q++)
#pragma unroll
            for (int j = 0; j < 8; j++) sum[j] += red[(q * C8 + tc) * 8 + j];
        if (partials) {
#pragma unroll
            for (int j = 0; j < 8; j++) partials[(size_t)blockIdx.x * C + tc * 8 + j] = sum[j];
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) atomicAdd(&db[tc * 8 + j], sum[j]);
        }
    }
}

// db[c] += sum over rows r < n_rows of partials[r][c]: block b takes columns
// [32 b, +32), its 32 row lanes stride the rows with four independent sums
// each, reduced in LDS.
__global__ __launch_bounds__(1024) void k_colsum(const float* __restrict__ partials, int n_rows, int C,
                                                 float* __restrict__ db) {
    __shared__ float red[1024];
    const int c = blockIdx.x * 32 + (threadIdx.x & 31), lane_r = threadIdx.x >> 5;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    int r = lane_r;
    for (; r + 96 < n_rows; r += 128)
#pragma unroll
        for (int u = 0; u < 4; u++) s[u] += partials[(size_t)(r + 32 * u) * C + c];
    for (; r < n_rows; r += 32) s[0] += partials[(size_t)r * C + c];
    red[threadIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    if (threadIdx.x < 32) {
        float sum = 0.f;
        for (int q = 0; q < 32; q++) sum += red[q * 32 + threadIdx.x];
        db[c] += sum;
    }
}

// dY *= (Y > 0): the ReLU backward on the stored (post-ReLU) output
__global__ void k_relu_bwd(const bf16* __restrict__ y, int ldy, bf16* __restrict__ dy, int lddy, int M, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * C) return;
    const size_t p = i / C, c = i % C;
    if (!(b2f(y[p * ldy + c]) > 0.0f)) dy[p * lddy + c] = f2b(0.0f);
}

// 2x2 max pool (stride 2): y[n][h][w][c] from x[n][2h..][2w..][c]
__global__ void k_pool(const bf16* __restrict__ x, int ldx, bf16* __restrict__ y, int ldy, int N, int H, int W,
                       int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    if (i >= (size_t)N * Ho * Wo * C) return;
    const int c = i % C;
    const size_t q = i / C;
    const int wo = q % Wo, ho = (q / Wo) % Ho, n = q / ((size_t)Wo * Ho);
    float m = -INFINITY;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            m = fmaxf(m, b2f(x[((size_t)(n * H + 2 * ho + a) * W + 2 * wo + b) * ldx + c]));
    y[q * ldy + c] = f2b(m);
}

// gradient to the first maximum of each 2x2 window (row-major), others 0;
// ACC: added to dx instead (dx + g, or dx + 0: what adding a written
// gradient to dx gives, in one pass)
template <bool ACC>
__global__ void k_pool_bwd(const bf16* __restrict__ x, int ldx, const bf16* __restrict__ y, int ldy,
                           const bf16* __restrict__ dy, int lddy, bf16* __restrict__ dx, int lddx, int N, int H,
                           int W, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    if (i >= (size_t)N * Ho * Wo * C) return;
    const int c = i % C;
    const size_t q = i / C;
    const int wo = q % Wo, ho = (q / Wo) % Ho, n = q / ((size_t)Wo * Ho);
    const float m = b2f(y[q * ldy + c]);
    const bf16 g = dy[q * lddy + c];
    bool done = false;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const size_t o = (size_t)(n * H + 2 * ho + a) * W + 2 * wo + b;
            const bool hit = !done && b2f(x[o * ldx + c]) == m;
            const bf16 r = hit ? g : f2b(0.0f);
            if constexpr (ACC) dx[o * lddx + c] = f2b(b2f(dx[o * lddx + c]) + b2f(r));
            else dx[o * lddx + c] = r;
            done = done || hit;
        }
}

// nearest 2x upsample: y[n][h][w][c] = x[n][h/2][w/2][c]  (y is H x W)
// (8 channels per thread: C, ldx, ldy multiples of 8, 16-byte aligned buffers)
__global__ void k_up(const bf16* __restrict__ x, int ldx, bf16* __restrict__ y, int ldy, int N, int H, int W, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (i >= (size_t)N * H * W * C8) return;
    const int c = (int)(i % C8) * 8;
    const size_t q = i / C8;
    const int w = q % W, h = (q / W) % H, n = q / ((size_t)W * H);
    *(bf16x8*)(y + q * ldy + c) = *(const bf16x8*)(x + ((size_t)(n * (H / 2) + h / 2) * (W / 2) + w / 2) * ldx + c);
}

// dx[n][h][w][c] = sum of the 4 dy of its 2x2 block (dy is 2H x 2W)
__global__ void k_up_bwd(const bf16* __restrict__ dy, int lddy, bf16* __restrict__ dx, int lddx, int N, int H, int W,
                         int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C) return;
    const int c = i % C;
    const size_t q = i / C;
    const int w = q % W, h = (q / W) % H, n = q / ((size_t)W * H);
    float s = 0.0f;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            s += b2f(dy[((size_t)(n * 2 * H + 2 * h + a) * (2 * W) + 2 * w + b) * lddy + c]);
    dx[q * lddx + c] = f2b(s);
}

// a += b (channel slices)
__global__ void k_add(bf16* __restrict__ a, int lda, const bf16* __restrict__ b, int ldb, int M, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * C) return;
    const size_t p = i / C, c = i % C;
    a[p * lda + c] = f2b(b2f(a[p * lda + c]) + b2f(b[p * ldb + c]));
}

// Loss over valid pixels: w_len * squared error on channel 0 (log1p slt),
// squared error on channel 1 (log1p slc) and w_bce * binary cross-entropy
// with logits on channel 2 (channel mask).  Writes the gradient (scaled by
// 1/n_valid) into dhead and adds the unweighted summed terms to out[0..2].
__global__ __launch_bounds__(256) void k_loss(const bf16* __restrict__ head, int ldh, const float* __restrict__ target,
                                              const uint8_t* __restrict__ valid, bf16* __restrict__ dhead, int lddh,
                                              int M, float inv_n, float w_len, float w_bce, float* __restrict__ out) {
    __shared__ float red[4][3];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float l0 = 0.f, l1 = 0.f, l2 = 0.f;
    if (p < M) {
        const bool ok = valid[p] != 0;
        const float a = b2f(head[(size_t)p * ldh + 0]), b = b2f(head[(size_t)p * ldh + 1]);
        const float z = b2f(head[(size_t)p * ldh + 2]);
        const float ta = target[3 * (size_t)p], tb = target[3 * (size_t)p + 1], tc = target[3 * (size_t)p + 2];
        float g0 = 0.f, g1 = 0.f, g2 = 0.f;
        if (ok) {
            l0 = (a - ta) * (a - ta);
            l1 = (b - tb) * (b - tb);
            // BCE with logits, stable: max(z,0) - z*t + log(1+exp(-|z|))
            l2 = fmaxf(z, 0.f) - z * tc + log1pf(expf(-fabsf(z)));
            g0 = w_len * 2.f * (a - ta) * inv_n;
            g1 = 2.f * (b - tb) * inv_n;
            g2 = w_bce * (1.f / (1.f + expf(-z)) - tc) * inv_n;
        }
        // the gradient row: channels 0..2, zeros up to min(lddh, 32), 16-byte stores
        bf16x8 v = {};
        v[0] = f2b(g0);
        v[1] = f2b(g1);
        v[2] = f2b(g2);
        bf16* row = dhead + (size_t)p * lddh;
        *(bf16x8*)row = v;
        const bf16x8 zero = {};
        const int c_end = min(lddh, 32);
        for (int c = 8; c < c_end; c += 8) *(bf16x8*)(row + c) = zero;
    }
    for (int off = 32; off > 0; off >>= 1) {
        l0 += __shfl_down(l0, off);
        l1 += __shfl_down(l1, off);
        l2 += __shfl_down(l2, off);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wid][0] = l0;
        red[wid][1] = l1;
        red[wid][2] = l2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        atomicAdd(&out[k], red[0][k] + red[1][k] + red[2][k] + red[3][k]);
    }
}

// Adam (Kingma & Ba 2015) on fp32 master weights, bf16 copy for the kernels.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16* __restrict__ pb, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = f2b(pi);
}

// k_adam, also writing each listed layer's weights flipped and transposed for
// the backward-data convolution (k_wflip's wt from the new bf16 value)
constexpr int kFlipMax = 32;
struct FlipTable {
    SlmNnFlip l[kFlipMax];
    int n;
};
__global__ void k_adam_flip(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, bf16* __restrict__ pb, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2, FlipTable T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    const bf16 w = f2b(pi);
    pb[i] = w;
    for (int k = 0; k < T.n; k++) {
        const SlmNnFlip& L = T.l[k];
        const int nw = (int)(L.cout * L.taps * L.cin);
        const int j = i - (int)L.offset;
        if (j < 0 || j >= nw) continue;
        const int cin = (int)L.cin, taps = (int)L.taps, cout = (int)L.cout;
        const int ci = j % cin, tap = (j / cin) % taps, co = j / (cin * taps);
        ((bf16*)L.wt)[((size_t)ci * taps + (taps - 1 - tap)) * cout + co] = w;
        break;
    }
}

// ------------------------------------------------------------------ tiles
//
// A tile is an S x S window of a grid (the tracer's padded grid, x-major:
// pixel (i, j) at i*ny + j) whose top-left corner tiles[2t], tiles[2t+1] may
// lie outside the grid; pixels outside the grid or masked are invalid.

__device__ __forceinline__ bool tile_pixel(const int* tiles, int t, int S, int q, int nx, int ny, size_t& idx) {
    const int i = tiles[2 * t] + q / S, j = tiles[2 * t + 1] + q % S;
    if (i < 0 || j < 0 || i >= nx || j >= ny) return false;
    idx = (size_t)i * ny + j;
    return true;
}

// mean[t] = mean of h over the tile's valid pixels (0 when none)
__global__ __launch_bounds__(256) void k_tile_mean(const float* __restrict__ h, const uint8_t* __restrict__ mask,
                                                   int nx, int ny, const int* __restrict__ tiles, int S,
                                                   float* __restrict__ mean) {
    __shared__ float ssum[4];
    __shared__ int scnt[4];
    const int t = blockIdx.x;
    float s = 0.f;
    int c = 0;
    for (int q = threadIdx.x; q < S * S; q += 256) {
        size_t idx;
        if (tile_pixel(tiles, t, S, q, nx, ny, idx) && !mask[idx]) {
            s += h[idx];
            c++;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off);
        c += __shfl_down(c, off);
    }
    if ((threadIdx.x & 63) == 0) {
        ssum[threadIdx.x >> 6] = s;
        scnt[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float st = ssum[0] + ssum[1] + ssum[2] + ssum[3];
        const int ct = scnt[0] + scnt[1] + scnt[2] + scnt[3];
        mean[t] = ct ? st / (float)ct : 0.f;
    }
}

// Network input, 16 bf16 channels per pixel: (h - mean)/h_scale, u, v,
// valid, then three second-order features computed in fp32 before the bf16
// rounding (which would leave h too coarse to difference): the divergence
// of uv (central differences, per pixel), 10 x the 5-point Laplacian of h
// (metres per pixel^2) and 5 x the slope |grad h| (metres per pixel); the
// three are 0 where a 4-neighbour is masked or outside the grid.  Channels
// 7..15 are 0.
__global__ void k_tile_inputs(const float* __restrict__ h, const float* __restrict__ uv,
                              const uint8_t* __restrict__ mask, int nx, int ny, const int* __restrict__ tiles,
                              int n_tiles, int S, const float* __restrict__ mean, float inv_scale,
                              bf16* __restrict__ x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n_tiles * S * S) return;
    const int t = (int)(i / ((size_t)S * S)), q = (int)(i % ((size_t)S * S));
    size_t idx;
    bf16x8 lo = {}, hi = {};
    if (tile_pixel(tiles, t, S, q, nx, ny, idx) && !mask[idx]) {
        lo[0] = f2b((h[idx] - mean[t]) * inv_scale);
        lo[1] = f2b(uv[2 * idx]);
        lo[2] = f2b(uv[2 * idx + 1]);
        lo[3] = f2b(1.0f);
        const int gi = (int)(idx / ny), gj = (int)(idx % ny);
        if (gi > 0 && gj > 0 && gi + 1 < nx && gj + 1 < ny) {
            const size_t n0 = idx - ny, n1 = idx + ny, n2 = idx - 1, n3 = idx + 1;
            if (!mask[n0] && !mask[n1] && !mask[n2] && !mask[n3]) {
                const float div = 0.5f * (uv[2 * n1] - uv[2 * n0]) + 0.5f * (uv[2 * n3 + 1] - uv[2 * n2 + 1]);
                const float lap = h[n0] + h[n1] + h[n2] + h[n3] - 4.0f * h[idx];
                const float gx = 0.5f * (h[n1] - h[n0]), gy = 0.5f * (h[n3] - h[n2]);
                lo[4] = f2b(div);
                lo[5] = f2b(10.0f * lap);
                lo[6] = f2b(5.0f * sqrtf(gx * gx + gy * gy));
            }
        }
    }
    *(bf16x8*)(x + i * 16) = lo;
    *(bf16x8*)(x + i * 16 + 8) = hi;
}

// Training targets from the raw downstream fields: log1p(slt/d^2),
// log1p(slc/d^2), [slc/d^2 >= threshold]; valid = in grid and not masked.
__global__ void k_tile_targets(const uint32_t* __restrict__ slc, const uint32_t* __restrict__ slt,
                               const uint8_t* __restrict__ mask, int nx, int ny, const int* __restrict__ tiles,
                               int n_tiles, int S, float inv_d2, float threshold, float* __restrict__ target,
                               uint8_t* __restrict__ valid) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n_tiles * S * S) return;
    const int t = (int)(i / ((size_t)S * S)), q = (int)(i % ((size_t)S * S));
    size_t idx;
    float a = 0.f, b = 0.f, c = 0.f;
    uint8_t ok = 0;
    if (tile_pixel(tiles, t, S, q, nx, ny, idx) && !mask[idx]) {
        const float dens = (float)slc[idx] * inv_d2;
        a = log1pf((float)slt[idx] * inv_d2);
        b = log1pf(dens);
        c = dens >= threshold ? 1.f : 0.f;
        ok = 1;
    }
    target[3 * i] = a;
    target[3 * i + 1] = b;
    target[3 * i + 2] = c;
    valid[i] = ok;
}

// Stitch: the core [halo, S-halo)^2 of each tile goes to out[k][pixel]
// (planar float32, k = 0..2: head channels 0..2) where it lies in the grid.
__global__ void k_tile_stitch(const bf16* __restrict__ head, int ldh, const int* __restrict__ tiles, int n_tiles,
                              int S, int halo, int nx, int ny, float* __restrict__ out) {
    const int C = S - 2 * halo;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n_tiles * C * C) return;
    const int t = (int)(i / ((size_t)C * C)), q = (int)(i % ((size_t)C * C));
    const int a = halo + q / C, b = halo + q % C;
    const int gi = tiles[2 * t] + a, gj = tiles[2 * t + 1] + b;
    if (gi < 0 || gj < 0 || gi >= nx || gj >= ny) return;
    const size_t src = ((size_t)t * S * S + (size_t)a * S + b) * ldh;
    const size_t dst = (size_t)gi * ny + gj, nxy = (size_t)nx * ny;
    out[dst] = b2f(head[src]);
    out[nxy + dst] = b2f(head[src + 1]);
    out[2 * nxy + dst] = b2f(head[src + 2]);
}

// Traced raw fields -> the surrogate's planar fields over a grid: out[0] =
// log1p(slt/d^2), out[1] = log1p(slc/d^2), out[2] = +1 where slc/d^2 >=
// threshold else -1 (a channel "logit"); masked pixels 0, 0, -1.
__global__ void k_fields_log(const uint32_t* __restrict__ slc, const uint32_t* __restrict__ slt,
                             const uint8_t* __restrict__ mask, size_t nxy, float inv_d2, float threshold,
                             float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    float a = 0.f, b = 0.f, c = -1.f;
    if (!mask[i]) {
        const float dens = (float)slc[i] * inv_d2;
        a = log1pf((float)slt[i] * inv_d2);
        b = log1pf(dens);
        c = dens >= threshold ? 1.f : -1.f;
    }
    out[i] = a;
    out[nxy + i] = b;
    out[2 * nxy + i] = c;
}

inline dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

int check_conv(int M, int Cin, int Cout, int ldx, int ldy, int taps) {
    if (taps != 1 && taps != 9) return slm_fail(SLM_E_INVALID_ARG, "taps %d is not 1 or 9", taps);
    if (M <= 0 || M % 32) return slm_fail(SLM_E_INVALID_ARG, "pixel count %d not a positive multiple of 32", M);
    if (Cin <= 0 || Cin % 16) return slm_fail(SLM_E_INVALID_ARG, "Cin %d not a positive multiple of 16", Cin);
    if (Cout <= 0 || Cout % 32) return slm_fail(SLM_E_INVALID_ARG, "Cout %d not a positive multiple of 32", Cout);
    if (ldx < Cin || ldx % 8 || ldy < Cout) return slm_fail(SLM_E_INVALID_ARG, "bad strides %d %d", ldx, ldy);
    return SLM_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

namespace {

int conv_impl(const void* x, int ldx, const void* x2, int ldx2, int c_up, const void* w, const float* bias, void* y,
              int ldy, int N, int H, int W, int Cin, int Cout, int taps, int relu, void* stream, uint32_t flags) {
    const int M = N * H * W;
    int rc = check_conv(M, Cin, Cout, ldx, ldy, taps);
    if (rc) return rc;
    if (!x || !w || !y) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (((uintptr_t)x) % 16) return slm_fail(SLM_E_INVALID_ARG, "x not 16-byte aligned");
    if (((uintptr_t)w) % 16) return slm_fail(SLM_E_INVALID_ARG, "w not 16-byte aligned");
    if (c_up) {
        if (!x2 || c_up < 0 || c_up > Cin || c_up % 32 || ldx2 < c_up || ldx2 % 8 || ((uintptr_t)x2) % 16 || H % 2 ||
            W % 2)
            return slm_fail(SLM_E_INVALID_ARG, "bad upsampled source (c_up %d, ldx2 %d)", c_up, ldx2);
    } else {
        x2 = x;
        ldx2 = ldx;
    }
    hipStream_t s = (hipStream_t)stream;
    const bf16* x2b = (const bf16*)x2;
    const bf16* xb = (const bf16*)x;
    const bf16* wb = (const bf16*)w;
    bf16* yb = (bf16*)y;
    // LDS-tiled path: 3x3, rows of 32 (or exactly 16) columns, whole block tiles
    const int bw = (W % 32 == 0) ? 32 : (W == 16 ? 16 : 0);
    const int ct = (Cout % 64 == 0) ? 64 : 32;
    const int rw = (bw == 32 && ct == 32) ? 4 : 2;
    const int bh = bw ? 4 * rw * (32 / bw) : 0;
    if (taps == 9 && bw && H % bh == 0 && !(flags & SLM_FLAG_NN_DIRECT)) {
        const dim3 grid((unsigned)(N * (H / bh) * (W / bw)), (unsigned)(Cout / ct));
        const bool k32 = Cin % 32 == 0;
#define LDS_L(BW, RW, CT, KC)                                                                                      \
    do {                                                                                                           \
        if (relu) hipLaunchKernelGGL((k_conv_lds<BW, RW, CT, KC, true>), grid, dim3(256), 0, s, xb, ldx, x2b,      \
                                     ldx2, c_up, wb, bias, yb, ldy, H, W, Cin, Cout);                              \
        else hipLaunchKernelGGL((k_conv_lds<BW, RW, CT, KC, false>), grid, dim3(256), 0, s, xb, ldx, x2b, ldx2,     \
                                c_up, wb, bias, yb, ldy, H, W, Cin, Cout);                                         \
    } while (0)
        if (bw == 32) {
            if (ct == 64) { if (k32) LDS_L(32, 2, 64, 32); else LDS_L(32, 2, 64, 16); }
            else { if (k32) LDS_L(32, 4, 32, 32); else LDS_L(32, 4, 32, 16); }
        } else {
            if (ct == 64) { if (k32) LDS_L(16, 2, 64, 32); else LDS_L(16, 2, 64, 16); }
            else { if (k32) LDS_L(16, 2, 32, 32); else LDS_L(16, 2, 32, 16); }
        }
#undef LDS_L
        return slm_finish(stream, flags);
    }
    if (Cout >= 64) {
        const dim3 grid((M + 63) / 64, (Cout + 63) / 64);
#define CONV_L(T, R) hipLaunchKernelGGL((k_conv<T, R, 2, 2>), grid, dim3(256), 0, s, xb, ldx, x2b, ldx2, c_up, wb, bias, yb, ldy, M, H, W, Cin, Cout)
        if (taps == 9) { if (relu) CONV_L(9, true); else CONV_L(9, false); }
        else { if (relu) CONV_L(1, true); else CONV_L(1, false); }
#undef CONV_L
    } else {
        const dim3 grid((M + 127) / 128, (Cout + 31) / 32);
#define CONV_L(T, R) hipLaunchKernelGGL((k_conv<T, R, 4, 1>), grid, dim3(256), 0, s, xb, ldx, x2b, ldx2, c_up, wb, bias, yb, ldy, M, H, W, Cin, Cout)
        if (taps == 9) { if (relu) CONV_L(9, true); else CONV_L(9, false); }
        else { if (relu) CONV_L(1, true); else CONV_L(1, false); }
#undef CONV_L
    }
    return slm_finish(stream, flags);
}

}  // namespace

extern "C" int slm_nn_conv(const void* x, int ldx, const void* w, const float* bias, void* y, int ldy, int N, int H,
                           int W, int Cin, int Cout, int taps, int relu, void* stream, uint32_t flags) {
    return conv_impl(x, ldx, nullptr, 0, 0, w, bias, y, ldy, N, H, W, Cin, Cout, taps, relu, stream, flags);
}

extern "C" int slm_nn_conv_up(const void* x, int ldx, const void* x2, int ldx2, int c_up, const void* w,
                              const float* bias, void* y, int ldy, int N, int H, int W, int Cin, int Cout, int taps,
                              int relu, void* stream, uint32_t flags) {
    if (c_up <= 0) return slm_fail(SLM_E_INVALID_ARG, "c_up must be positive");
    return conv_impl(x, ldx, x2, ldx2, c_up, w, bias, y, ldy, N, H, W, Cin, Cout, taps, relu, stream, flags);
}

extern "C" int slm_nn_wflip(const void* w, void* wt, int taps, int Cin, int Cout, void* stream, uint32_t flags) {
    if (!w || !wt || (taps != 1 && taps != 9) || Cin <= 0 || Cout <= 0)
        return slm_fail(SLM_E_INVALID_ARG, "bad wflip arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_wflip, g1((size_t)Cout * taps * Cin), dim3(256), 0, s, (const bf16*)w, (bf16*)wt, taps, Cin,
                       Cout);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_conv_wgrad(const void* x, int ldx, const void* dy, int lddy, float* dw, float* db, int N, int H,
                                 int W, int Cin, int Cout, int taps, void* stream, uint32_t flags) {
    const int M = N * H * W;
    int rc = check_conv(M, Cin, Cout, ldx, lddy, taps);
    if (rc) return rc;
    if (!x || !dy || !dw) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    if (Cout > 256) return slm_fail(SLM_E_INVALID_ARG, "Cout %d above 256", Cout);
    if (lddy % 8 || ((uintptr_t)dy) % 16 || ((uintptr_t)x) % 16)
        return slm_fail(SLM_E_INVALID_ARG, "dy/x must be 16-byte aligned with lddy a multiple of 8");
    hipStream_t s = (hipStream_t)stream;
    const int K = taps * Cin;
    if (flags & SLM_FLAG_NN_DIRECT) {
        const int tiles = ((Cout + 63) / 64) * ((K + 63) / 64);
        // enough pixel splits to fill the chip (~1024 blocks), at least 64 pixels each
        int splits = (1024 + tiles - 1) / tiles;
        int ppb = (M + splits - 1) / splits;
        ppb = ((ppb + 63) / 64) * 64;
        splits = (M + ppb - 1) / ppb;
        const dim3 grid((Cout + 63) / 64, (K + 63) / 64, splits);
        if (taps == 9)
            hipLaunchKernelGGL(k_wgrad<9>, grid, dim3(256), 0, s, (const bf16*)x, ldx, (const bf16*)dy, lddy, dw, M, H,
                               W, Cin, Cout, ppb);
        else
            hipLaunchKernelGGL(k_wgrad<1>, grid, dim3(256), 0, s, (const bf16*)x, ldx, (const bf16*)dy, lddy, dw, M, H,
                               W, Cin, Cout, ppb);
    } else if ((W == 16 || W == 32 || W == 64 || W == 128) && (Cin % 32 == 0 || Cin == 16) &&
               (size_t)N * H * W == (size_t)M) {
        // LDS-staged kernel: tiles of 32 or 64 output x 16 or 32 input channels x the taps;
        // at least 4 chunks per block (every block ends in a full tile of atomics),
        // about 512 blocks when there is work for them
        const int ct = Cout % 64 == 0 ? 64 : 32, cit = Cin % 32 == 0 ? 32 : 16;
        const int R = 256 / W;
        const int chunks = N * ((H + R - 1) / R);
        const int tiles = (Cout / ct) * (Cin / cit);
        int splits = std::max(1, std::min((512 + tiles - 1) / tiles, chunks / 4));
        const int cpb = (chunks + splits - 1) / splits;
        splits = (chunks + cpb - 1) / cpb;
        const dim3 grid(Cout / ct, Cin / cit, splits), block((taps == 9 ? 192 : 64) * (ct / 32));
#define WH_L(TT, WW, C, CI) hipLaunchKernelGGL((k_wgrad_halo<TT, WW, C, CI>), grid, block, 0, s, (const bf16*)x, ldx, (const bf16*)dy, lddy, dw, N, H, Cin, cpb)
#define WH_C(TT, WW) \
        if (ct == 64) { if (cit == 32) WH_L(TT, WW, 64, 32); else WH_L(TT, WW, 64, 16); } \
        else { if (cit == 32) WH_L(TT, WW, 32, 32); else WH_L(TT, WW, 32, 16); }
#define WH_W(TT) \
        if (W == 16) { WH_C(TT, 16) } else if (W == 32) { WH_C(TT, 32) } else if (W == 64) { WH_C(TT, 64) } \
        else { WH_C(TT, 128) }
        if (taps == 9) { WH_W(9) } else { WH_W(1) }
#undef WH_W
#undef WH_C
#undef WH_L
    } else {
        // 64 x 64 tiles, or 32 x 128 when Cout is 32; pixel splits of whole 128-pixel
        // chunks, enough to fill the chip (~1024 blocks)
        const bool narrow = Cout % 64 != 0;
        const int ct = narrow ? 32 : 64, kt = narrow ? 128 : 64;
        const int tiles = (Cout / ct) * ((K + kt - 1) / kt);
        int splits = (1024 + tiles - 1) / tiles;
        int ppb = (M + splits - 1) / splits;
        ppb = ((ppb + 127) / 128) * 128;
        splits = (M + ppb - 1) / ppb;
        const dim3 grid(Cout / ct, (K + kt - 1) / kt, splits);
#define WG_L(T, C, KK) hipLaunchKernelGGL((k_wgrad_t<T, C, KK>), grid, dim3(256), 0, s, (const bf16*)x, ldx, (const bf16*)dy, lddy, dw, M, H, W, Cin, Cout, ppb)
        if (taps == 9) { if (narrow) WG_L(9, 32, 128); else WG_L(9, 64, 64); }
        else { if (narrow) WG_L(1, 32, 128); else WG_L(1, 64, 64); }
#undef WG_L
    }
    if (db) {
        // few blocks: every block ends in Cout atomics on the same addresses
        const int P = 256 / (Cout / 8);
        const int blocks = std::min(256, (M + P - 1) / P);
        hipLaunchKernelGGL(k_bgrad, dim3(blocks), dim3(256), 0, s, (const bf16*)dy, lddy, db, M, Cout);
    }
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_relu_bwd(const void* y, int ldy, void* dy, int lddy, int M, int C, void* stream, uint32_t flags) {
    if (!y || !dy || M < 0 || C <= 0) return slm_fail(SLM_E_INVALID_ARG, "bad relu_bwd arguments");
    hipLaunchKernelGGL(k_relu_bwd, g1((size_t)M * C), dim3(256), 0, (hipStream_t)stream, (const bf16*)y, ldy, (bf16*)dy,
                       lddy, M, C);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_pool(const void* x, int ldx, void* y, int ldy, int N, int H, int W, int C, void* stream,
                           uint32_t flags) {
    if (!x || !y || H % 2 || W % 2) return slm_fail(SLM_E_INVALID_ARG, "bad pool arguments");
    hipLaunchKernelGGL(k_pool, g1((size_t)N * (H / 2) * (W / 2) * C), dim3(256), 0, (hipStream_t)stream,
                       (const bf16*)x, ldx, (bf16*)y, ldy, N, H, W, C);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_pool_bwd(const void* x, int ldx, const void* y, int ldy, const void* dy, int lddy, void* dx,
                               int lddx, int N, int H, int W, int C, void* stream, uint32_t flags) {
    if (!x || !y || !dy || !dx || H % 2 || W % 2) return slm_fail(SLM_E_INVALID_ARG, "bad pool_bwd arguments");
    if (flags & SLM_FLAG_NO_ZERO)
        hipLaunchKernelGGL(k_pool_bwd<true>, g1((size_t)N * (H / 2) * (W / 2) * C), dim3(256), 0,
                           (hipStream_t)stream, (const bf16*)x, ldx, (const bf16*)y, ldy, (const bf16*)dy, lddy,
                           (bf16*)dx, lddx, N, H, W, C);
    else
        hipLaunchKernelGGL(k_pool_bwd<false>, g1((size_t)N * (H / 2) * (W / 2) * C), dim3(256), 0,
                           (hipStream_t)stream, (const bf16*)x, ldx, (const bf16*)y, ldy, (const bf16*)dy, lddy,
                           (bf16*)dx, lddx, N, H, W, C);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_up(const void* x, int ldx, void* y, int ldy, int N, int H, int W, int C, void* stream,
                         uint32_t flags) {
    if (!x || !y || H % 2 || W % 2 || C % 8 || ldx % 8 || ldy % 8 || ((uintptr_t)x) % 16 || ((uintptr_t)y) % 16)
        return slm_fail(SLM_E_INVALID_ARG, "bad upsample arguments");
    hipLaunchKernelGGL(k_up, g1((size_t)N * H * W * (C / 8)), dim3(256), 0, (hipStream_t)stream, (const bf16*)x, ldx,
                       (bf16*)y, ldy, N, H, W, C);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_up_bwd(const void* dy, int lddy, void* dx, int lddx, int N, int H, int W, int C, void* stream,
                             uint32_t flags) {
    if (!dy || !dx) return slm_fail(SLM_E_INVALID_ARG, "bad upsample backward arguments");
    hipLaunchKernelGGL(k_up_bwd, g1((size_t)N * H * W * C), dim3(256), 0, (hipStream_t)stream, (const bf16*)dy, lddy,
                       (bf16*)dx, lddx, N, H, W, C);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_add(void* a, int lda, const void* b, int ldb, int M, int C, void* stream, uint32_t flags) {
    if (!a || !b) return slm_fail(SLM_E_INVALID_ARG, "NULL buffer");
    hipLaunchKernelGGL(k_add, g1((size_t)M * C), dim3(256), 0, (hipStream_t)stream, (bf16*)a, lda, (const bf16*)b, ldb,
                       M, C);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_loss(const void* head, int ldh, const float* target, const uint8_t* valid, void* dhead, int lddh,
                           int M, float inv_n, float w_len, float w_bce, float* out, void* stream, uint32_t flags) {
    if (!head || !target || !valid || !dhead || !out || ldh < 3 || lddh < 8 || lddh % 8 || ((uintptr_t)dhead) % 16)
        return slm_fail(SLM_E_INVALID_ARG, "bad loss arguments (dhead: 16-byte aligned, lddh a multiple of 8)");
    hipLaunchKernelGGL(k_loss, g1((size_t)M), dim3(256), 0, (hipStream_t)stream, (const bf16*)head, ldh, target, valid,
                       (bf16*)dhead, lddh, M, inv_n, w_len, w_bce, out);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_adam(float* p, const float* g, float* m, float* v, void* pb, int n, float lr, float b1, float b2,
                           float eps, int step, void* stream, uint32_t flags) {
    if (!p || !g || !m || !v || n <= 0 || step <= 0) return slm_fail(SLM_E_INVALID_ARG, "bad adam arguments");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, g1((size_t)n), dim3(256), 0, (hipStream_t)stream, p, g, m, v, (bf16*)pb, n, lr, b1, b2,
                       eps, c1, c2);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_adam_flip(float* p, const float* g, float* m, float* v, void* pb, int n, float lr, float b1,
                                float b2, float eps, int step, const SlmNnFlip* layers, int n_layers, void* stream,
                                uint32_t flags) {
    if (!p || !g || !m || !v || !pb || n <= 0 || step <= 0 || n_layers < 0 || n_layers > kFlipMax ||
        (n_layers && !layers))
        return slm_fail(SLM_E_INVALID_ARG, "bad adam_flip arguments");
    FlipTable T;
    T.n = n_layers;
    for (int k = 0; k < n_layers; k++) {
        const SlmNnFlip& L = layers[k];
        if (!L.wt || (L.taps != 1 && L.taps != 9) || !L.cin || !L.cout ||
            L.offset + (uint64_t)L.cout * L.taps * L.cin > (uint64_t)n)
            return slm_fail(SLM_E_INVALID_ARG, "adam_flip layer %d: bad shape, offset or buffer", k);
        T.l[k] = L;
    }
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam_flip, g1((size_t)n), dim3(256), 0, (hipStream_t)stream, p, g, m, v, (bf16*)pb, n, lr, b1,
                       b2, eps, c1, c2, T);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_tile_inputs(const float* h, const float* uv, const uint8_t* mask, int nx, int ny,
                                  const int* tiles, int n_tiles, int S, float h_scale, float* mean_ws, void* x,
                                  void* stream, uint32_t flags) {
    if (!h || !uv || !mask || !tiles || !mean_ws || !x || nx <= 0 || ny <= 0 || n_tiles < 0 || S <= 0 ||
        !(h_scale > 0.f))
        return slm_fail(SLM_E_INVALID_ARG, "bad tile_inputs arguments");
    if (((uintptr_t)x) % 16) return slm_fail(SLM_E_INVALID_ARG, "x not 16-byte aligned");
    if (!n_tiles) return slm_finish(stream, flags);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_tile_mean, dim3(n_tiles), dim3(256), 0, s, h, mask, nx, ny, tiles, S, mean_ws);
    hipLaunchKernelGGL(k_tile_inputs, g1((size_t)n_tiles * S * S), dim3(256), 0, s, h, uv, mask, nx, ny, tiles,
                       n_tiles, S, (const float*)mean_ws, 1.0f / h_scale, (bf16*)x);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_tile_targets(const uint32_t* slc, const uint32_t* slt, const uint8_t* mask, int nx, int ny,
                                   const int* tiles, int n_tiles, int S, int density, float threshold,
                                   float* target, uint8_t* valid, void* stream, uint32_t flags) {
    if (!slc || !slt || !mask || !tiles || !target || !valid || nx <= 0 || ny <= 0 || n_tiles < 0 || S <= 0 ||
        density <= 0)
        return slm_fail(SLM_E_INVALID_ARG, "bad tile_targets arguments");
    if (!n_tiles) return slm_finish(stream, flags);
    hipLaunchKernelGGL(k_tile_targets, g1((size_t)n_tiles * S * S), dim3(256), 0, (hipStream_t)stream, slc, slt,
                       mask, nx, ny, tiles, n_tiles, S, 1.0f / (float)(density * density), threshold, target, valid);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_tile_stitch(const void* head, int ldh, const int* tiles, int n_tiles, int S, int halo, int nx,
                                  int ny, float* out, void* stream, uint32_t flags) {
    if (!head || !tiles || !out || ldh < 3 || n_tiles < 0 || S <= 0 || halo < 0 || 2 * halo >= S || nx <= 0 ||
        ny <= 0)
        return slm_fail(SLM_E_INVALID_ARG, "bad tile_stitch arguments");
    if (!n_tiles) return slm_finish(stream, flags);
    const int C = S - 2 * halo;
    hipLaunchKernelGGL(k_tile_stitch, g1((size_t)n_tiles * C * C), dim3(256), 0, (hipStream_t)stream,
                       (const bf16*)head, ldh, tiles, n_tiles, S, halo, nx, ny, out);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_fields_log(const uint32_t* slc, const uint32_t* slt, const uint8_t* mask, size_t nxy,
                                 int density, float threshold, float* out, void* stream, uint32_t flags) {
    if (!slc || !slt || !mask || !out || density <= 0) return slm_fail(SLM_E_INVALID_ARG, "bad fields_log arguments");
    if (!nxy) return slm_finish(stream, flags);
    hipLaunchKernelGGL(k_fields_log, g1(nxy), dim3(256), 0, (hipStream_t)stream, slc, slt, mask, nxy,
                       1.0f / (float)(density * density), threshold, out);
    return slm_finish(stream, flags);
}

extern "C" int slm_nn_relu_bias_grad(const void* y, int ldy, void* dy, int lddy, float* db, int M, int C,
                                     float* partials, int n_partials, void* stream, uint32_t flags) {
    if (!dy || M < 0 || C <= 0 || C % 32 || C > 256 || lddy % 8 || ((uintptr_t)dy) % 16 ||
        (y && (ldy % 8 || ((uintptr_t)y) % 16)) || (partials && n_partials < C))
        return slm_fail(SLM_E_INVALID_ARG, "bad relu_bias_grad arguments");
    if (!M || (!y && !db)) return slm_finish(stream, flags);
    hipStream_t s = (hipStream_t)stream;
    const int P = 256 / (C / 8);
    if (db && partials) {
        // two blocks per CU keep enough loads in flight; their sums are reduced by k_colsum
        const int blocks = std::min(std::min(512, n_partials / C), (M + P - 1) / P);
        hipLaunchKernelGGL(k_relu_bias, dim3(blocks), dim3(256), 0, s, (const bf16*)y, ldy, (bf16*)dy, lddy, db,
                           partials, M, C);
        hipLaunchKernelGGL(k_colsum, dim3(C / 32), dim3(1024), 0, s, (const float*)partials, blocks, C, db);
    } else {
        const int blocks = std::min(db ? 128 : 512, (M + P - 1) / P);
        hipLaunchKernelGGL(k_relu_bias, dim3(blocks), dim3(256), 0, s, (const bf16*)y, ldy, (bf16*)dy, lddy, db,
                           (float*)nullptr, M, C);
    }
    return slm_finish(stream, flags);
}
