// Learned k-step surrogate: kernels.
//
// The network (no counterpart in the reference) maps a 1-channel state to
// the state k steps ahead:
//   z1 = conv5(x; 1->32)+b1, a1 = gelu(z1);  p = avgpool2(a1)
//   z2 = conv5(p; 32->64)+b2, a2 = gelu(z2);  u = upsample2(a2)
//   z3 = conv5(u; 64->32)+b3, s = gelu(z3) + a1
//   r  = conv5(s; 32->1)+b4,  y = clip(x + r, 0, 1)
// conv5 = 5x5 cross-correlation with circular padding, gelu = z*Phi(z).
//
// Layout: activations are NHWC, bf16 where a convolution reads them (a1, p,
// a2, s, dz3, dz2 and xb = bf16(x)) and for gelu' of the pre-activations,
// kept for the backward pass (half the bytes of the largest tensors the
// step moves); FP32 for the backward input gradients dp and dz1.
//
// Hot kernels run on MFMA as an implicit GEMM, D[co][pixel] = sum over
// (tap, ci) of W[co][tap][ci] * X[pixel+tap][ci], with
// v_mfma_f32_16x16x32_bf16: A = 16 output channels x 32 input channels of one
// tap (one 16-byte load per lane from [co][tap][ci] weights), B = 32 input
// channels x 16 consecutive pixels (one 16-byte load per lane from NHWC).
// The upsampled input of layer 3 is never stored: the kernel reads the
// half-resolution tensor at (y/2, x/2).  The backward input gradients of
// layers 2 and 3 are the same kernel over tap-flipped, channel-swapped
// weights.  Layer 1 (one input channel) forms its 25-tap im2col fragments
// in registers from an LDS tile of x; layer 4 (one output channel) puts the
// 25 taps in the MFMA's N dimension and adds the shifted products.  The weight gradients
// (reductions over pixels) use fixed-order partial sums.
#include <algorithm>
#include <cmath>

#include "ah_internal.h"
#include "surrogate.h"

namespace ah {
namespace surr {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// Two floats as packed bf16 (element a in the low half), rounded to nearest
// even by one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
__device__ __forceinline__ uint32_t bf16_bits(float f) { return pack2(f, 0.0f) & 0xffffu; }
__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// Four floats as bf16 (round to nearest even), element 0 in the low half.
__device__ __forceinline__ uint2 pack4(const float (&v)[4]) { return make_uint2(pack2(v[0], v[1]), pack2(v[2], v[3])); }
// Eight bf16 (one 16-byte load) as floats.
__device__ __forceinline__ void unpack8(const uint4& u, float (&v)[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[2 * k] = __uint_as_float(w[k] << 16);
        v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
    }
}

// gelu(z) = z Phi(z) and, from the same erf, gelu'(z) = Phi(z) + z phi(z).
// erf by Abramowitz & Stegun 7.1.26 (Hastings): erf(x) = 1 - P(t) exp(-x^2),
// t = 1 / (1 + 0.3275911 x), |error| <= 1.5e-7 for x >= 0.  With x = |z|/sqrt(2)
// the exponential is exp(-z^2/2), which gelu' needs as well, so one exp
// serves both.
__device__ __forceinline__ float gelu_and_grad(float z, float& dg) {
    const float E = __expf(-0.5f * z * z);
    const float x = fabsf(z) * 0.70710678118654752f;
    const float t = __frcp_rn(fmaf(0.3275911f, x, 1.0f));
    const float P = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f),
                             0.254829592f);
    const float h = 0.5f * P * E;                 // 1 - Phi(|z|)
    const float phi = z >= 0.0f ? 1.0f - h : h;  // Phi(z)
    dg = fmaf(z * E, 0.3989422804014327f, phi);
    return z * phi;
}
__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// ---------------------------------------------------------------- MFMA conv
// Bias, z, GELU (+ skip) and the bf16 activation of one wave's outputs: lane
// holds channels 16m + 4q .. +3 of pixel x0 + 16b + l16 of row y.
template <int MT, int NB>
__device__ __forceinline__ void conv_epilogue(const ConvArgs& a, const f32x4 (&acc)[MT][NB], int n, int y, int x0,
                                              int q, int l16) {
    const int H = a.H, W = a.W, Cout = a.Cout;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int co = 16 * m + 4 * q;
            const size_t o = (((size_t)n * H + y) * W + x0 + 16 * b + l16) * Cout + co;
            float z[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) z[r] = acc[m][b][r] + (a.bias ? a.bias[co + r] : 0.0f);
            if (a.mode == CONV_GELU || a.mode == CONV_GELU_ADD) {
                float v[4], dg[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = gelu_and_grad(z[r], dg[r]);
                if (a.dgelu) *reinterpret_cast<uint2*>(a.dgelu + o) = pack4(dg);
                if (a.mode == CONV_GELU_ADD) {
                    const uint2 ad = *reinterpret_cast<const uint2*>(a.add + o);
                    v[0] += __uint_as_float(ad.x << 16);
                    v[1] += __uint_as_float(ad.x & 0xffff0000u);
                    v[2] += __uint_as_float(ad.y << 16);
                    v[3] += __uint_as_float(ad.y & 0xffff0000u);
                }
                *reinterpret_cast<uint2*>(a.a + o) = pack4(v);
            } else if (a.z) {
                *reinterpret_cast<float4*>(a.z + o) = make_float4(z[0], z[1], z[2], z[3]);
            }
        }
}

// Layer 1 (5x5, 1 -> 32): the MFMA product over a 32-channel im2col
// (channel t < 25 = bf16(x) at tap t, the rest zero), without the im2col in
// memory.  Wave: one output row, 64 consecutive pixels (4 blocks of 16), 32
// output channels; workgroup: 4 waves = 4 rows; grid (W/64, H/4, N).  The
// workgroup stages its 8 x 68 window of bf16(x) in LDS, and lane (q, l16)
// gathers its B fragment, taps 8q .. 8q+7 of pixel l16 of a block, from
// there.  Besides gelu'(z1) (when a.dgelu is set) and a1 = bf16(gelu(z1)) it writes
// xb = bf16(x) for its own 4 x 64 pixels and the 2 x 2 average pool of a1,
// pool[y/2][x/2] = bf16(((a(x,y) + a(x+1,y)) + a(x,y+1)) + a(x+1,y+1)) / 4):
// rows y and y+1 (waves 2j, 2j+1) meet in LDS, pixels x and x+1 (lanes l16
// and l16 ^ 1) by a lane exchange.
__global__ void __launch_bounds__(256) k_conv1(ConvArgs a, const float* __restrict__ x, uint16_t* __restrict__ xb,
                                               uint16_t* __restrict__ pool) {
    constexpr int NB = 4, TR = 8, TC = 16 * NB + 4;
    __shared__ uint16_t tile[TR * TC];
    __shared__ float odd_rows[2 * 2 * NB * 4 * 64];  // waves 1 and 3: their bf16 activations
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int n = blockIdx.z, y0 = blockIdx.y * 4, y = y0 + wave, x0 = blockIdx.x * (16 * NB);
    const int H = a.H, W = a.W;
    const float* xs = x + (size_t)n * H * W;
    for (int i = threadIdx.x; i < TR * TC; i += 256) {
        const int r = i / TC, c = i - r * TC;
        const int yy = wrap(y0 - 2 + r, H), xx = wrap(x0 - 2 + c, W);
        const uint16_t h = (uint16_t)bf16_bits(xs[(size_t)yy * W + xx]);
        tile[i] = h;
        if (r >= 2 && r < 6 && c >= 2 && c < TC - 2) xb[((size_t)n * H + yy) * W + xx] = h;
    }
    __syncthreads();
    bf16x8 A[2];
#pragma unroll
    for (int m = 0; m < 2; ++m)
        A[m] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(a.Wt + (size_t)(16 * m + l16) * 32 + 8 * q));
    f32x4 acc[2][NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        // Tap t = (dy + 2) * 5 + (dx + 2) of pixel (wave, 16 b + l16) is
        // window pixel (wave + t / 5, 16 b + l16 + t % 5).
        uint32_t h[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t v2[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int t = 8 * q + 2 * j + e;
                v2[e] = t < 25 ? (uint32_t)tile[(wave + t / 5) * TC + 16 * b + l16 + t % 5] : 0u;
            }
            h[j] = v2[0] | (v2[1] << 16);
        }
        const bf16x8 B = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
#pragma unroll
        for (int m = 0; m < 2; ++m)
            acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
    }
    // Lane holds channels 16m + 4q .. +3 of pixel x0 + 16b + l16 of row y.
    float act[2][NB][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int co = 16 * m + 4 * q;
            const size_t o = (((size_t)n * H + y) * W + x0 + 16 * b + l16) * 32 + co;
            float dg[4], v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = gelu_and_grad(acc[m][b][r] + a.bias[co + r], dg[r]);
            const uint2 h = pack4(v);
            act[m][b][0] = __uint_as_float(h.x << 16);
            act[m][b][1] = __uint_as_float(h.x & 0xffff0000u);
            act[m][b][2] = __uint_as_float(h.y << 16);
            act[m][b][3] = __uint_as_float(h.y & 0xffff0000u);
            if (a.dgelu) *reinterpret_cast<uint2*>(a.dgelu + o) = pack4(dg);
            *reinterpret_cast<uint2*>(a.a + o) = h;
        }
    float* sc = odd_rows + (wave >> 1) * (2 * NB * 4 * 64);
    if (wave & 1) {
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) sc[((m * NB + b) * 4 + r) * 64 + lane] = act[m][b][r];
    }
    __syncthreads();
    if (!(wave & 1)) {
        const int Hh = H / 2, Wh = W / 2, yh = y >> 1;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                float pv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float v00 = act[m][b][r], v10 = __shfl_xor(v00, 1);
                    const float* row1 = sc + ((m * NB + b) * 4 + r) * 64;
                    const float v01 = row1[lane], v11 = row1[lane ^ 1];
                    pv[r] = 0.25f * (((v00 + v10) + v01) + v11);
                }
                const uint2 h = pack4(pv);
                if (!(l16 & 1)) {
                    const int xh = (x0 + 16 * b + l16) >> 1, co = 16 * m + 4 * q;
                    const size_t o = (((size_t)n * Hh + yh) * Wh + xh) * 32 + co;
                    *reinterpret_cast<uint2*>(pool + o) = h;
                }
            }
    }
}

// The 5x5 convolutions over 32 or 64 input channels.  Wave: one output
// row, NB blocks of 16 pixels, MT blocks of 16 output channels; workgroup:
// 4 waves = 4 rows.  Per 32-channel slice the workgroup loads its 8 x (16 NB
// + 4)-pixel input window (4 output rows plus the 2-row, 2-column halo) once,
// and every tap's B fragments come from LDS (read from the L1 cache for
// every tap they were 25 loads per input).  An upsampled layer (UP) stages the 4 x 34
// half-resolution pixels under that window and reads window pixel (r, c)
// at (r/2, c/2).  Pixel stride 40 bf16 (80 bytes): a wave's 16-pixel
// fragment reads fall on distinct banks.
namespace cl {
constexpr int TR = 8, PS = 40;  // window rows, bf16 per pixel (32 + 8)
}

// NB blocks of 16 pixels per wave (a window of 16 NB + 4 columns).
template <int MT, bool UP, int NB>
__global__ void __launch_bounds__(256) k_conv_lds(ConvArgs a) {
    using namespace cl;
    constexpr int TC = 16 * NB + 4;
    constexpr int SR = UP ? TR / 2 : TR, SC = UP ? TC / 2 : TC;  // staged rows, columns
    __shared__ __align__(16) uint16_t tile[SR * SC * PS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int n = blockIdx.z, y0 = blockIdx.y * 4, y = y0 + wave, x0 = blockIdx.x * (16 * NB);
    const int H = a.H, W = a.W, Cin = a.Cin;
    const int Hx = UP ? H / 2 : H, Wx = UP ? W / 2 : W;
    const uint16_t* X = a.X + (size_t)n * Hx * Wx * Cin;
    // First staged input row and column: the window starts at (y0 - 2, x0 - 2),
    // at (y0/2 - 1, x0/2 - 1) in the half-resolution input (y0, x0 even).
    const int ys = UP ? y0 / 2 - 1 : y0 - 2, xs = UP ? x0 / 2 - 1 : x0 - 2;
    f32x4 acc[MT][NB];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[m][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int c0 = 0; c0 < Cin; c0 += 32) {
        if (c0) __syncthreads();  // every wave has read the previous slice
        for (int i = threadIdx.x; i < SR * SC * 4; i += 256) {
            const int pix = i >> 2, part = i & 3, r = pix / SC, c = pix - r * SC;
            const int yy = wrap(ys + r, Hx), xx = wrap(xs + c, Wx);
            const uint4 v = *reinterpret_cast<const uint4*>(X + ((size_t)yy * Wx + xx) * Cin + c0 + 8 * part);
            *reinterpret_cast<uint4*>(tile + pix * PS + 8 * part) = v;
        }
        // Weight fragments rotate through NS register sets: tap t uses set
        // t % NS while the loads for tap t + NS - 1 are in flight, so NS - 1
        // taps of MFMAs hide each load.  Taps 0..23 run in groups of NS (NS
        // divides 24) with the set index known at compile time and no branch
        // in the group (a rolled loop that copied set to set waited for the
        // newest load before every copy); tap 24 follows the loop.
        constexpr int NS = MT <= 2 ? 3 : 2;
        static_assert(24 % NS == 0, "tap groups");
        auto weights = [&](int t, bf16x8 (&A)[MT]) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
                A[m] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(a.Wt + ((size_t)(16 * m + l16) * 25 + t) * Cin + c0 + 8 * q));
        };
        // Block b of the window lies 16 b pixels (8 b half-resolution
        // pixels) right of block 0.
        constexpr int BSTEP = (UP ? 8 : 16) * PS;
        auto tap = [&](int t, const bf16x8 (&A)[MT]) {
            const int dy = t / 5 - 2, dx = t - 5 * (t / 5) - 2;
            const int r = wave + 2 + dy, c = 2 + dx + l16;  // window pixel of block 0
            const uint16_t* b0 = tile + (UP ? (r >> 1) * SC + (c >> 1) : r * SC + c) * PS + 8 * q;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const bf16x8 B = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(b0 + b * BSTEP));
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][b], 0, 0, 0);
            }
        };
        bf16x8 Aw[NS][MT];
#pragma unroll
        for (int j = 0; j < NS - 1; ++j) weights(j, Aw[j]);
        __syncthreads();
#pragma unroll 1
        for (int t0 = 0; t0 < 24; t0 += NS) {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                // With NS = 3 the last group's final load repeats tap 24 (unused).
                weights(min(t0 + j + NS - 1, 24), Aw[(j + NS - 1) % NS]);
                // Keep the loads here: the scheduler would otherwise sink them
                // next to their first use, a tap or less ahead.
                __builtin_amdgcn_sched_barrier(0);
                tap(t0 + j, Aw[j]);
            }
        }
        tap(24, Aw[0]);
    }
    constexpr bool can_pool = !UP && MT * NB * 64 * 4 * 2 * sizeof(float) <= sizeof(tile);
    if (can_pool && a.mode == CONV_POOL_DGELU) {
        // Rows y0 + 2j and y0 + 2j + 1 (waves 2j, 2j + 1) meet in LDS, then
        // lanes l16 and l16 ^ 1 (pixels x, x + 1) meet by a lane exchange.
        float* sc = reinterpret_cast<float*>(tile);
        __syncthreads();  // every wave has read the last slice
        if (wave & 1) {
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        sc[(((wave >> 1) * MT + m) * NB + b) * 256 + r * 64 + lane] = acc[m][b][r];
        }
        __syncthreads();
        if (!(wave & 1)) {
            const int Hh = H / 2, Wh = W / 2, yh = y >> 1;
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float col = acc[m][b][r] + sc[(((wave >> 1) * MT + m) * NB + b) * 256 + r * 64 + lane];
                        v[r] = col + __shfl_xor(col, 1);
                    }
                    if (!(l16 & 1)) {
                        const int xh = (x0 + 16 * b + l16) >> 1, co = 16 * m + 4 * q;
                        const size_t o = (((size_t)n * Hh + yh) * Wh + xh) * a.Cout + co;
                        const uint2 dg = *reinterpret_cast<const uint2*>(a.dgelu_in + o);
                        const float g[4] = {v[0] * bf16f(dg.x), v[1] * bf16f(dg.x >> 16), v[2] * bf16f(dg.y),
                                            v[3] * bf16f(dg.y >> 16)};
                        *reinterpret_cast<uint2*>(a.dpool + o) = pack4(g);
                    }
                }
        }
        return;
    }
    conv_epilogue<MT, NB>(a, acc, n, y, x0, q, l16);
}

// Layer 4 (5x5, 32 -> 1) and the output: r = conv(a3, W4) + b4 and
// y = clip(x + r, 0, 1).  With one output channel, a product over (pixel,
// channel) x (channel, output) would leave 15 of an MFMA's 16 columns zero;
// here the taps are the N dimension instead: Y[p][t] = sum over ci of
// a3[p][ci] W4[t][ci] for every window pixel p (16 pixels x 16 taps per
// MFMA, the 32 channels in K), and r(y, x) = sum over taps (dy, dx) of
// Y[(y + dy, x + dx)][(dy, dx)], added from LDS in tap order.  Workgroup:
// 4 output rows x 64 columns; window rows y0 - 2 .. y0 + 5 and columns
// x0 - 2 .. x0 + 77 (5 blocks of 16, of which 68 columns are read).  LDS
// keeps only the (window row, dy) pairs some output row of the tile uses:
// window row r serves dy + 2 = e in [max(0, r - 3), min(4, r)], 5 taps per
// e, 100 slots in all.
namespace ot {
constexpr int ROWS = 4, COLS = 64, WR = ROWS + 4, NBLK = 5, UNITS = WR * NBLK, PER = UNITS / 4;
constexpr int YS = 84;  // floats per slot: 80 columns, padded so 8 slots of one store meet distinct banks
__host__ __device__ constexpr int e_lo(int r) { return r > 3 ? r - 3 : 0; }
__host__ __device__ constexpr int e_hi(int r) { return r < 4 ? r : 4; }
constexpr int slot0_sum(int r) {
    int s = 0;
    for (int i = 0; i < r; ++i) s += 5 * (e_hi(i) - e_lo(i) + 1);
    return s;
}
constexpr int SLOTS = slot0_sum(WR);
// First slot of window row r (0 <= r <= 8), in closed form.
__host__ __device__ constexpr int slot0(int r) { return r <= 4 ? 5 * r * (r + 1) / 2 : SLOTS - 5 * (8 - r) * (9 - r) / 2; }
static_assert(SLOTS == 100 && UNITS % 4 == 0, "layer-4 window");
static_assert(slot0(1) == slot0_sum(1) && slot0(3) == slot0_sum(3) && slot0(4) == slot0_sum(4) &&
                  slot0(5) == slot0_sum(5) && slot0(7) == slot0_sum(7) && slot0(8) == slot0_sum(8),
              "slot table");
}  // namespace ot

__global__ void __launch_bounds__(256) k_out_taps(const uint16_t* __restrict__ a3, const uint16_t* __restrict__ W4b,
                                                  const float* __restrict__ b4, const float* __restrict__ x,
                                                  float* __restrict__ r, float* __restrict__ y, int H, int W) {
    using namespace ot;
    __shared__ __align__(16) float Y[SLOTS * YS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int n = blockIdx.z, y0 = blockIdx.y * ROWS, x0 = blockIdx.x * COLS;
    const uint16_t* A3 = a3 + (size_t)n * H * W * 32;
    // A: 16 window pixels x channels 8q..8q+7; unit u = wave + 4i is window
    // row u / NBLK, block u % NBLK.
    bf16x8 A[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int u = wave + 4 * i, rw = u / NBLK, b = u - NBLK * rw;
        const int gy = wrap(y0 - 2 + rw, H), gx = wrap(x0 - 2 + 16 * b + l16, W);
        A[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(A3 + ((size_t)gy * W + gx) * 32 + 8 * q));
    }
    // B: W4 transposed, tap 16g + l16 in N (taps 25..31 zero), channels in K.
    bf16x8 B[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        const int t = 16 * g + l16;
        const uint4 v = t < 25 ? *reinterpret_cast<const uint4*>(W4b + t * 32 + 8 * q) : make_uint4(0, 0, 0, 0);
        B[g] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int u = wave + 4 * i, rw = u / NBLK, b = u - NBLK * rw;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            // Lane holds Y of window pixels 16b + 4q .. +3 for tap 16g + l16.
            const f32x4 d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B[g], f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
            const int t = 16 * g + l16, e = t / 5;
            if (t < 25 && e >= e_lo(rw) && e <= e_hi(rw))
                *reinterpret_cast<f32x4*>(Y + (slot0(rw) + t - 5 * e_lo(rw)) * YS + 16 * b + 4 * q) = d;
        }
    }
    __syncthreads();
    // Output row y0 + wave, column x0 + lane: taps in order, dy then dx.
    float s = 0.0f;
#pragma unroll
    for (int e = 0; e < 5; ++e) {
        const int rw = wave + e;
        const float* row = Y + (slot0(rw) + 5 * (e - e_lo(rw))) * YS + lane;
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) s += row[dx * YS + dx];
    }
    const size_t o = ((size_t)n * H + y0 + wave) * W + x0 + lane;
    const float rr = s + b4[0];
    r[o] = rr;
    y[o] = fminf(fmaxf(x[o] + rr, 0.0f), 1.0f);
}

// MSE loss: dr = 2 (y - t) / count where 0 < x + r < 1, else 0; per-block
// FP64 partial sums of (y - t)^2.
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ y, const float* __restrict__ tgt,
                                              const float* __restrict__ x, const float* __restrict__ r,
                                              float* __restrict__ dr, double* __restrict__ part, size_t total) {
    __shared__ double sh[256];
    double acc = 0.0;
    const float scale = 2.0f / (float)total;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const float d = y[i] - tgt[i];
        acc += (double)d * (double)d;
        if (dr) {
            const float pre = x[i] + r[i];
            dr[i] = (pre > 0.0f && pre < 1.0f) ? scale * d : 0.0f;
        }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ds = d(loss)/ds, the layer-4 input gradient: ds[p][ci] = sum over taps t
// of W4[t][ci] dr[p - tap_t].  It is not stored: the two kernels below form
// it where they need it.  Workgroup = one 64-pixel row segment (grid W/64,
// H, N), thread = (pixel tid / 4, channels 8 (tid % 4) .. +7).  The segment's
// 5 x 68 dr window and W4 sit in LDS; each thread adds its 25 taps in order,
// one fused multiply-add each, so both kernels form the same values.
namespace dz {
constexpr int SEG = 64, COLS = SEG + 4;

__device__ __forceinline__ void stage(float* w, float* drt, const float* __restrict__ W4, const float* __restrict__ dr,
                                      int n, int y, int x0, int H, int W) {
    for (int i = threadIdx.x; i < 25 * 32; i += 256) w[i] = W4[i];
    for (int i = threadIdx.x; i < 5 * COLS; i += 256) {
        const int r = i / COLS, c = i - r * COLS;
        drt[i] = dr[((size_t)n * H + wrap(y - 2 + r, H)) * W + wrap(x0 - 2 + c, W)];
    }
    __syncthreads();
}

// Tap t = (dy + 2) * 5 + (dx + 2) reads dr at (y - dy, x - dx): window row
// 2 - dy, column p + 2 - dx.
__device__ __forceinline__ void ds8(const float* w, const float* drt, int p, int c8, float (&acc)[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0f;
#pragma unroll
    for (int t = 0; t < 25; ++t) {
        const int dy = t / 5 - 2, dx = t % 5 - 2;
        const float d = drt[(2 - dy) * COLS + p + 2 - dx];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = __fmaf_rn(w[t * 32 + 8 * c8 + k], d, acc[k]);
    }
}
}  // namespace dz

// dz3 = bf16(ds * gelu'(z3)), g3 = gelu'(z3) as the forward pass stored it (bf16).
__global__ void __launch_bounds__(256) k_dz3_dr(const float* __restrict__ dr, const float* __restrict__ W4,
                                                const uint16_t* __restrict__ g3, uint16_t* __restrict__ dz3, int N,
                                                int H, int W) {
    __shared__ float w[25 * 32];
    __shared__ float drt[5 * dz::COLS];
    const int n = blockIdx.z, y = blockIdx.y, x0 = blockIdx.x * dz::SEG;
    dz::stage(w, drt, W4, dr, n, y, x0, H, W);
    const int p = threadIdx.x >> 2, c8 = threadIdx.x & 3;
    const size_t pix = ((size_t)n * H + y) * W + x0 + p;
    float ds[8];
    dz::ds8(w, drt, p, c8, ds);
    float g[8];
    unpack8(*reinterpret_cast<const uint4*>(g3 + pix * 32 + 8 * c8), g);
    uint32_t h[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = pack2(ds[2 * k] * g[2 * k], ds[2 * k + 1] * g[2 * k + 1]);
    *reinterpret_cast<uint4*>(dz3 + pix * 32 + 8 * c8) = make_uint4(h[0], h[1], h[2], h[3]);
}

// ---------------------------------------------------------------- weight grads
// dW[co][t][ci] = sum over pixels of dZ[pix][co] * X[pix + tap_t][ci], the
// bias gradient the sum of dZ; per-workgroup partials in fixed ranges of
// pixels, summed in workgroup order (results do not depend on timing).
namespace wg {
constexpr int SEG = 64, ROWS = 5, COLS = SEG + 4, THREADS = 512, GROUPS = 256;
// Workgroups of the thin-layer kernels: 800 outputs each, so many more
// workgroups than the MFMA kernels' (256 left 4 waves per CU).
constexpr int THIN_GROUPS = 2048;
static_assert((size_t)GROUPS * 64 * 25 * 32 <= WGRAD_PART_FLOATS && (size_t)THIN_GROUPS * 800 <= WGRAD_PART_FLOATS &&
                  GROUPS * 64 <= WGRAD_BPART_FLOATS && THIN_GROUPS * 32 <= WGRAD_BPART_FLOATS,
              "partials fit the workspace");
}

template <int COUT, int CIN, int UP>
__global__ void __launch_bounds__(wg::THREADS) k_wgrad_mfma(WGradArgs a, float* __restrict__ part,
                                                            float* __restrict__ bpart) {
    using namespace wg;
    static_assert((COUT / 16) * (CIN / 16) == THREADS / 64, "one 16x16 block per wave");
    static_assert(COLS % 2 == 0 && 32 * (SEG / 32 - 1) + 8 * 3 + 16 <= COLS + 4, "tile columns");
    // X tile transposed, [ci][row][column] (CS bf16 per row), so a lane's 8
    // consecutive pixels of one channel are contiguous; dZ per pixel (ZS).
    constexpr int CS = COLS + 4, ZS = COUT + 8;
    __shared__ __attribute__((aligned(16))) uint16_t xt[CIN * ROWS * CS];
    __shared__ __attribute__((aligned(16))) uint16_t dz[SEG * ZS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int cob = wave / (CIN / 16), cib = wave % (CIN / 16);
    const int H = a.H, W = a.W, Hx = UP ? H / 2 : H, Wx = UP ? W / 2 : W;
    const int segs_per_row = W / SEG;
    const long nseg = (long)a.N * H * segs_per_row;
    const long s0 = nseg * blockIdx.x / gridDim.x, s1 = nseg * (blockIdx.x + 1) / gridDim.x;
    f32x4 acc[25];
#pragma unroll
    for (int t = 0; t < 25; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bacc = 0.0f;  // bias gradient of channel threadIdx.x (< COUT) over this workgroup's pixels
    // A segment's global loads are issued before the previous segment's
    // MFMAs and written to LDS after them (registers in between), so their
    // latency hides behind the MFMAs.  dZ: 64 pixels x COUT, contiguous in
    // global memory; X tile: rows y-2..y+2, columns x0-2..x0+65, wrapped on
    // the torus, transposed on the way in: a thread takes 8 channels of two
    // adjacent pixels and writes each channel's pair as one 32-bit word.
    constexpr int DZ_V = SEG * COUT / 8, X_V = ROWS * (COLS / 2) * (CIN / 8);
    constexpr int NDZ = (DZ_V + THREADS - 1) / THREADS, NX = (X_V + THREADS - 1) / THREADS;
    uint4 pdz[NDZ], px0[NX], px1[NX];
    auto fetch = [&](long sg) {
        const int xb = (int)(sg % segs_per_row);
        const long ry = sg / segs_per_row;
        const int y = (int)(ry % H), n = (int)(ry / H), x0 = xb * SEG;
        const uint16_t* src = a.dz_h + (((size_t)n * H + y) * W + x0) * COUT;
#pragma unroll
        for (int i = 0; i < NDZ; ++i) {
            const int v = threadIdx.x + i * THREADS;
            if (v < DZ_V) {
                const int px = v / (COUT / 8), part8 = v % (COUT / 8);
                pdz[i] = *reinterpret_cast<const uint4*>(src + (size_t)px * COUT + 8 * part8);
            }
        }
#pragma unroll
        for (int i = 0; i < NX; ++i) {
            const int v = threadIdx.x + i * THREADS;
            if (v < X_V) {
                const int pair = v % (ROWS * (COLS / 2)), part8 = v / (ROWS * (COLS / 2));
                const int r = pair / (COLS / 2), c = 2 * (pair % (COLS / 2));
                int sy = wrap(y + r - 2, H), sx0 = wrap(x0 + c - 2, W), sx1 = wrap(x0 + c - 1, W);
                if (UP) {
                    sy >>= 1;
                    sx0 >>= 1;
                    sx1 >>= 1;
                }
                const uint16_t* row = a.X + ((size_t)n * Hx + sy) * Wx * a.xstride + a.xoff + 8 * part8;
                px0[i] = *reinterpret_cast<const uint4*>(row + (size_t)sx0 * a.xstride);
                px1[i] = *reinterpret_cast<const uint4*>(row + (size_t)sx1 * a.xstride);
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int i = 0; i < NDZ; ++i) {
            const int v = threadIdx.x + i * THREADS;
            if (v < DZ_V) {
                const int px = v / (COUT / 8), part8 = v % (COUT / 8);
                *reinterpret_cast<uint4*>(dz + px * ZS + 8 * part8) = pdz[i];
            }
        }
#pragma unroll
        for (int i = 0; i < NX; ++i) {
            const int v = threadIdx.x + i * THREADS;
            if (v < X_V) {
                const int pair = v % (ROWS * (COLS / 2)), part8 = v / (ROWS * (COLS / 2));
                const int r = pair / (COLS / 2), c = 2 * (pair % (COLS / 2));
                const uint32_t e0[4] = {px0[i].x, px0[i].y, px0[i].z, px0[i].w};
                const uint32_t e1[4] = {px1[i].x, px1[i].y, px1[i].z, px1[i].w};
                uint32_t* dst = reinterpret_cast<uint32_t*>(xt + ((size_t)(8 * part8) * ROWS + r) * CS + c);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    dst[(2 * j) * ROWS * CS / 2] = (e0[j] & 0xffffu) | (e1[j] << 16);
                    dst[(2 * j + 1) * ROWS * CS / 2] = (e0[j] >> 16) | (e1[j] & 0xffff0000u);
                }
            }
        }
    };
    if (s0 < s1) fetch(s0);
    for (long sg = s0; sg < s1; ++sg) {
        __syncthreads();  // the previous segment's LDS reads are done
        store();
        __syncthreads();
        if (sg + 1 < s1) fetch(sg + 1);
        if ((int)threadIdx.x < COUT)
            for (int px = 0; px < SEG; ++px) bacc += bf16f(dz[px * ZS + threadIdx.x]);
#pragma unroll
        for (int kc = 0; kc < SEG / 32; ++kc) {
            const int p0 = 32 * kc + 8 * q;  // this lane's 8 pixels: p0 .. p0 + 7
            uint32_t ab[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                ab[j] = (uint32_t)dz[(p0 + 2 * j) * ZS + 16 * cob + l16] |
                        ((uint32_t)dz[(p0 + 2 * j + 1) * ZS + 16 * cob + l16] << 16);
            const bf16x8 A = __builtin_bit_cast(bf16x8, make_uint4(ab[0], ab[1], ab[2], ab[3]));
#pragma unroll
            for (int dy = 0; dy < 5; ++dy) {
                // Columns p0 .. p0+15 of channel 16 cib + l16, tile row dy (two
                // aligned 16-byte reads) serve the 5 taps of this row: tap dx
                // (tile column offset, tap + 2) needs columns p0+dx .. p0+dx+7.
                const uint4* src =
                    reinterpret_cast<const uint4*>(xt + ((size_t)(16 * cib + l16) * ROWS + dy) * CS + p0);
                const uint4 w0 = src[0], w1 = src[1];
                const uint32_t w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                for (int dx = 0; dx < 5; ++dx) {
                    uint32_t bb[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        bb[k] = dx % 2 == 0 ? w[dx / 2 + k] : (w[(dx - 1) / 2 + k] >> 16) | (w[(dx + 1) / 2 + k] << 16);
                    const bf16x8 B = __builtin_bit_cast(bf16x8, make_uint4(bb[0], bb[1], bb[2], bb[3]));
                    acc[dy * 5 + dx] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[dy * 5 + dx], 0, 0, 0);
                }
            }
        }
    }
    if ((int)threadIdx.x < COUT) bpart[(size_t)blockIdx.x * COUT + threadIdx.x] = bacc;
    // Lane holds dW[co = 16 cob + 4q + r][t][ci = 16 cib + l16].
    float* dst = part + (size_t)blockIdx.x * COUT * 25 * CIN;
#pragma unroll
    for (int t = 0; t < 25; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[((size_t)(16 * cob + 4 * q + r) * 25 + t) * CIN + 16 * cib + l16] = acc[t][r];
}

// Layer 3's weight gradient through the parity split of its nearest-
// upsampled input u (u[y][x] = a2[y >> 1][x >> 1]).  Output pixel (2i + py,
// 2j + px) meets a2 at (i + A(py, dy), j + A(px, dx)) through tap (dy, dx)
// (0..4, tap + 2), A(p, d) = (p + d - 2) >> 1 in {-1, 0, 1}, so
//   dW3[co][dy][dx][ci] = sum over parities (py, px) of
//                         G[py][px][A(py, dy)][A(px, dx)][co][ci],
//   G[py][px][a][b] = sum over half-resolution (i, j) of
//                     dz3[2i + py][2j + px][co] * a2[i + a][j + b][ci]:
// 36 blocks over a quarter of the pixels each, 9 MFMA blocks per pixel
// where the direct form takes 25; the inputs are the same bf16 values.
// Grid (UPG, 4 parities); a workgroup walks a fixed range of 64-pixel
// half-resolution row segments of its parity class, stages that row's dz3
// pixels transposed ([co][k]) and a2 rows i-1..i+1, columns j0-1..j0+64
// transposed ([ci][row][column]), prefetching the next segment into
// registers during the MFMAs; wave w owns the (co, ci) block (w / 4, w % 4)
// for all 9 (a, b).  Partials: part[workgroup][parity][a][b][co][ci],
// bias partials bpart[workgroup][parity][co].
namespace wu {
constexpr int SEG = 64, COUT = 32, CIN = 64, THREADS = 512, UPG = 128;
static_assert((size_t)UPG * 4 * 9 * COUT * CIN <= WGRAD_PART_FLOATS && UPG * 4 * COUT <= WGRAD_BPART_FLOATS,
              "partials fit the workspace");
constexpr int ROWS = 3, COLS = SEG + 2, CS = COLS + 6, DS = SEG + 8;  // a2 tile; dz3 row stride
static_assert(8 * 3 + 32 + 16 <= CS, "B windows stay inside a tile row");
}

__global__ void __launch_bounds__(wu::THREADS) k_wgrad_up(WGradArgs a, float* __restrict__ part,
                                                          float* __restrict__ bpart) {
    using namespace wu;
    __shared__ __attribute__((aligned(16))) uint16_t xt[CIN * ROWS * CS];
    __shared__ __attribute__((aligned(16))) uint16_t dzt[COUT * DS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int cob = wave / (CIN / 16), cib = wave % (CIN / 16);
    const int par = blockIdx.y, py = par >> 1, px = par & 1;
    const int H = a.H, W = a.W, Hh = H / 2, Wh = W / 2;
    const int segs_per_row = Wh / SEG;
    const long nseg = (long)a.N * Hh * segs_per_row;
    const long s0 = nseg * blockIdx.x / gridDim.x, s1 = nseg * (blockIdx.x + 1) / gridDim.x;
    f32x4 acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bacc = 0.0f;
    // dz3: SEG pixels (2i + py, 2(j0 + k) + px) x 32 channels, 8 channels
    // per load; a2: ROWS x COLS pixels x 64 channels, 8 channels of two
    // adjacent pixels per task, written as one 32-bit word per channel.
    constexpr int DZ_V = SEG * COUT / 8, X_V = ROWS * (COLS / 2) * (CIN / 8);
    constexpr int NDZ = (DZ_V + THREADS - 1) / THREADS, NX = (X_V + THREADS - 1) / THREADS;
    uint4 pdz[NDZ], px0[NX], px1[NX];
    auto fetch = [&](long sg) {
        const int jb = (int)(sg % segs_per_row);
        const long ri = sg / segs_per_row;
        const int i = (int)(ri % Hh), n = (int)(ri / Hh), j0 = jb * SEG;
        const uint16_t* zrow = a.dz_h + (((size_t)n * H + 2 * i + py) * W) * COUT;
#pragma unroll
        for (int u = 0; u < NDZ; ++u) {
            const int v = threadIdx.x + u * THREADS;
            if (v < DZ_V) {
                const int k = v / (COUT / 8), part8 = v % (COUT / 8);
                pdz[u] = *reinterpret_cast<const uint4*>(zrow + (size_t)(2 * (j0 + k) + px) * COUT + 8 * part8);
            }
        }
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int v = threadIdx.x + u * THREADS;
            if (v < X_V) {
                const int pair = v % (ROWS * (COLS / 2)), part8 = v / (ROWS * (COLS / 2));
                const int r = pair / (COLS / 2), c = 2 * (pair % (COLS / 2));
                const int sy = wrap(i + r - 1, Hh), sx0 = wrap(j0 + c - 1, Wh), sx1 = wrap(j0 + c, Wh);
                const uint16_t* row = a.X + ((size_t)n * Hh + sy) * Wh * a.xstride + a.xoff + 8 * part8;
                px0[u] = *reinterpret_cast<const uint4*>(row + (size_t)sx0 * a.xstride);
                px1[u] = *reinterpret_cast<const uint4*>(row + (size_t)sx1 * a.xstride);
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < NDZ; ++u) {
            const int v = threadIdx.x + u * THREADS;
            if (v < DZ_V) {
                const int k = v / (COUT / 8), part8 = v % (COUT / 8);
                const uint32_t e[4] = {pdz[u].x, pdz[u].y, pdz[u].z, pdz[u].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    dzt[(8 * part8 + 2 * j) * DS + k] = (uint16_t)(e[j] & 0xffffu);
                    dzt[(8 * part8 + 2 * j + 1) * DS + k] = (uint16_t)(e[j] >> 16);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NX; ++u) {
            const int v = threadIdx.x + u * THREADS;
            if (v < X_V) {
                const int pair = v % (ROWS * (COLS / 2)), part8 = v / (ROWS * (COLS / 2));
                const int r = pair / (COLS / 2), c = 2 * (pair % (COLS / 2));
                const uint32_t e0[4] = {px0[u].x, px0[u].y, px0[u].z, px0[u].w};
                const uint32_t e1[4] = {px1[u].x, px1[u].y, px1[u].z, px1[u].w};
                uint32_t* dst = reinterpret_cast<uint32_t*>(xt + ((size_t)(8 * part8) * ROWS + r) * CS + c);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    dst[(2 * j) * ROWS * CS / 2] = (e0[j] & 0xffffu) | (e1[j] << 16);
                    dst[(2 * j + 1) * ROWS * CS / 2] = (e0[j] >> 16) | (e1[j] & 0xffff0000u);
                }
            }
        }
    };
    if (s0 < s1) fetch(s0);
    for (long sg = s0; sg < s1; ++sg) {
        __syncthreads();  // the previous segment's LDS reads are done
        store();
        __syncthreads();
        if (sg + 1 < s1) fetch(sg + 1);
        if ((int)threadIdx.x < COUT)
            for (int k = 0; k < SEG; ++k) bacc += bf16f(dzt[threadIdx.x * DS + k]);
#pragma unroll
        for (int kc = 0; kc < SEG / 32; ++kc) {
            const int p0 = 32 * kc + 8 * q;  // this lane's 8 half-resolution pixels
            const bf16x8 A = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(dzt + (16 * cob + l16) * DS + p0));
#pragma unroll
            for (int ar = 0; ar < 3; ++ar) {
                // Tile row ar (a = ar - 1), columns p0 .. p0 + 15; offset b
                // (= bc - 1) needs columns p0 + bc .. p0 + bc + 7.
                const uint4* src =
                    reinterpret_cast<const uint4*>(xt + ((size_t)(16 * cib + l16) * ROWS + ar) * CS + p0);
                const uint4 w0 = src[0], w1 = src[1];
                const uint32_t w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                for (int bc = 0; bc < 3; ++bc) {
                    uint32_t bb[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        bb[k] = bc % 2 == 0 ? w[bc / 2 + k] : (w[(bc - 1) / 2 + k] >> 16) | (w[(bc + 1) / 2 + k] << 16);
                    const bf16x8 B = __builtin_bit_cast(bf16x8, make_uint4(bb[0], bb[1], bb[2], bb[3]));
                    acc[ar * 3 + bc] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[ar * 3 + bc], 0, 0, 0);
                }
            }
        }
    }
    const size_t grp = (size_t)blockIdx.x * 4 + par;
    if ((int)threadIdx.x < COUT) bpart[grp * COUT + threadIdx.x] = bacc;
    // Lane holds G[a][b][co = 16 cob + 4q + r][ci = 16 cib + l16].
    float* dst = part + grp * 9 * COUT * CIN;
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int r = 0; r < 4; ++r) dst[((size_t)k * COUT + 16 * cob + 4 * q + r) * CIN + 16 * cib + l16] = acc[k][r];
}

// dW3[co][t][ci] from the G partials: thread = output (co, t, ci); sums over
// workgroups, then parities, in that fixed order.
__global__ void k_reduce_up(const float* __restrict__ part, int groups, float* __restrict__ out) {
    using namespace wu;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= COUT * 25 * CIN) return;
    const int ci = o % CIN, t = (o / CIN) % 25, co = o / (CIN * 25);
    const int dy = t / 5, dx = t % 5;
    float acc = 0.0f;
    for (int g = 0; g < groups; ++g)
#pragma unroll
        for (int par = 0; par < 4; ++par) {
            const int py = par >> 1, px = par & 1;
            const int ab = (((py + dy - 2) >> 1) + 1) * 3 + (((px + dx - 2) >> 1) + 1);
            acc += part[((((size_t)g * 4 + par) * 9 + ab) * COUT + co) * CIN + ci];
        }
    out[o] = acc;
}

// The thin layers' weight gradients on MFMA.  With U the fp32 output
// gradient and V the bf16 layer input, out[a][t][b] = sum over pixels p of
// U[p][a] V[p + tap_t][b] is a product over K = pixels:
//   layer 4 (A = 1, B = 32): M = tap t (25 of 32 rows), N = channel b;
//     A-operand U[q - tap_t] (pixel q = p + tap_t), B-operand V[q][b];
//   layer 1 (A = 32, B = 1): M = channel a, N = tap t (25 of 32 columns);
//     A-operand U[p][a], B-operand V[p + tap_t].
// U enters as two bf16 terms, hi = bf16(U) and lo = bf16(U - hi), so each
// product keeps about 16 bits of U; V is bf16 already.  A workgroup walks a
// fixed range of 64-pixel row segments (partials in workgroup order, so the
// result does not depend on timing); per segment it
// stages U and V in LDS (the operand read along pixels as [channel][pixel],
// the shifted one as its 5 x 68 tile) and each of its 4 waves adds two
// 32-pixel chunks into its own 16 x 16 output block (mb = wave / 2, nb =
// wave % 2).  The bias gradient's partial (the sum of U over the workgroup's
// pixels, segment by segment in pixel order) goes to bpart.  Layer 1's U,
// dz1, is formed in LDS by the same kernel (k_dz1_wgrad) and never stored.
namespace wt {
constexpr int VS = wg::SEG + 8;  // layer 4: bf16 per channel row of V^T
constexpr int US = wg::SEG + 4;  // layer 1: fp32 per channel row of U^T
}

__device__ __forceinline__ void split8(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        h[j] = pack2(v[2 * j], v[2 * j + 1]);
        l[j] = pack2(v[2 * j] - __uint_as_float(h[j] << 16), v[2 * j + 1] - __uint_as_float(h[j] & 0xffff0000u));
    }
    hi = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
    lo = __builtin_bit_cast(bf16x8, make_uint4(l[0], l[1], l[2], l[3]));
}

// Layer 4: U = dr (one channel), V = s (32 channels).
__global__ void __launch_bounds__(256) k_wgrad_thin4(WGradArgs a, float* __restrict__ part, float* __restrict__ bpart) {
    using wg::SEG;
    using wg::ROWS;
    using wg::COLS;
    // U tile [ROWS][COLS] fp32 and V^T [32][VS] bf16.
    __shared__ __attribute__((aligned(16))) float uf[ROWS * COLS];
    __shared__ __attribute__((aligned(16))) uint16_t vh[32 * wt::VS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int mb = wave >> 1, nb = wave & 1;
    const int H = a.H, W = a.W, segs_per_row = W / SEG;
    const long nseg = (long)a.N * H * segs_per_row;
    const long s0 = nseg * blockIdx.x / gridDim.x, s1 = nseg * (blockIdx.x + 1) / gridDim.x;
    // This lane's tap: row m of the A operand.
    const int t = 16 * mb + l16;
    const bool tap_ok = t < 25;
    const int dy = tap_ok ? t / 5 - 2 : 0, dx = tap_ok ? t % 5 - 2 : 0;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bacc = 0.0f;
    for (long sg = s0; sg < s1; ++sg) {
        const int xb = (int)(sg % segs_per_row);
        const long ry = sg / segs_per_row;
        const int y = (int)(ry % H), n = (int)(ry / H), x0 = xb * SEG;
        __syncthreads();  // the previous segment's reads are done
        for (int i = threadIdx.x; i < ROWS * COLS; i += 256) {
            const int r = i / COLS, c = i % COLS;
            uf[i] = a.dz_f[((size_t)n * H + wrap(y + r - 2, H)) * W + wrap(x0 + c - 2, W)];
        }
        {   // V^T: thread = (pixel, 8 channels)
            const int p = threadIdx.x >> 2, c8 = threadIdx.x & 3;
            const uint4 v = *reinterpret_cast<const uint4*>(a.X + (((size_t)n * H + y) * W + x0 + p) * a.xstride +
                                                            a.xoff + 8 * c8);
            const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                vh[(8 * c8 + 2 * j) * wt::VS + p] = (uint16_t)(e[j] & 0xffffu);
                vh[(8 * c8 + 2 * j + 1) * wt::VS + p] = (uint16_t)(e[j] >> 16);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int p = 0; p < SEG; ++p) bacc += uf[2 * COLS + 2 + p];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int k0 = 32 * c + 8 * q;  // this lane's first pixel of the chunk
            // A: U at (y - dy, x0 + k0 + j - dx), tile row 2 - dy, column k0 + j - dx + 2.
            float u8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) u8[j] = tap_ok ? uf[(2 - dy) * COLS + k0 + j - dx + 2] : 0.0f;
            const bf16x8 B =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(vh + (16 * nb + l16) * wt::VS + k0));
            bf16x8 hi, lo;
            split8(u8, hi, lo);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, B, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo, B, acc, 0, 0, 0);
        }
    }
    if (threadIdx.x == 0) bpart[blockIdx.x] = bacc;
    // D: rows 16 mb + 4 q + r (tap), column 16 nb + l16 (channel).
    float* dst = part + (size_t)blockIdx.x * 800;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = 16 * mb + 4 * q + r;
        if (m < 25) dst[m * 32 + 16 * nb + l16] = acc[r];
    }
}

// Layer 1: per segment, dz1 = (ds + dp[p/2] / 4) * gelu'(z1) (ds formed from
// the segment's 5 x 68 tile of dr as in k_dz3_dr; dp the pooled path's
// gradient at half resolution; gelu'(z1) as the forward pass kept it) goes
// into LDS as U^T [channel][pixel]; V = bf16(x) from its 5 x 68 tile.
__global__ void __launch_bounds__(256) k_dz1_wgrad(const float* __restrict__ dr, const float* __restrict__ W4,
                                                   const float* __restrict__ dp, const uint16_t* __restrict__ g1,
                                                   const uint16_t* __restrict__ xb, int N, int H, int W,
                                                   float* __restrict__ part, float* __restrict__ bpart) {
    using wg::SEG;
    using wg::ROWS;
    using wg::COLS;
    static_assert(dz::SEG == SEG && dz::COLS == COLS && ROWS == 5, "segment tiles");
    __shared__ float w4[25 * 32];
    __shared__ float drt[ROWS * COLS];
    __shared__ __attribute__((aligned(16))) float uf[32 * wt::US];
    __shared__ __attribute__((aligned(16))) uint16_t vh[ROWS * COLS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int mb = wave >> 1, nb = wave & 1;
    const int p = threadIdx.x >> 2, c8 = threadIdx.x & 3;  // dz1: pixel p, channels 8 c8 .. +7
    const int segs_per_row = W / SEG;
    const long nseg = (long)N * H * segs_per_row;
    const long s0 = nseg * blockIdx.x / gridDim.x, s1 = nseg * (blockIdx.x + 1) / gridDim.x;
    // This lane's tap: column n of the B operand.
    const int t = 16 * nb + l16;
    const bool tap_ok = t < 25;
    const int dy = tap_ok ? t / 5 - 2 : 0, dx = tap_ok ? t % 5 - 2 : 0;
    for (int i = threadIdx.x; i < 25 * 32; i += 256) w4[i] = W4[i];
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bacc = 0.0f;
    for (long sg = s0; sg < s1; ++sg) {
        const int xs = (int)(sg % segs_per_row);
        const long ry = sg / segs_per_row;
        const int y = (int)(ry % H), n = (int)(ry / H), x0 = xs * SEG;
        // This thread's dp and gelu'(z1), in flight during the staging.
        const size_t pix = ((size_t)n * H + y) * W + x0 + p;
        const float4* pp =
            reinterpret_cast<const float4*>(dp + (((size_t)n * (H / 2) + y / 2) * (W / 2) + (x0 + p) / 2) * 32 + 8 * c8);
        const float4 pa = pp[0], pb = pp[1];
        const uint4 gu = *reinterpret_cast<const uint4*>(g1 + pix * 32 + 8 * c8);
        __syncthreads();  // the previous segment's reads are done (and w4 is written)
        for (int i = threadIdx.x; i < ROWS * COLS; i += 256) {
            const int r = i / COLS, c = i - r * COLS;
            const size_t o = ((size_t)n * H + wrap(y + r - 2, H)) * W + wrap(x0 + c - 2, W);
            drt[i] = dr[o];
            vh[i] = xb[o];
        }
        __syncthreads();
        {
            float ds[8], g[8];
            dz::ds8(w4, drt, p, c8, ds);
            unpack8(gu, g);
            const float d[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
#pragma unroll
            for (int k = 0; k < 8; ++k) uf[(8 * c8 + k) * wt::US + p] = (ds[k] + 0.25f * d[k]) * g[k];
        }
        __syncthreads();
        if ((int)threadIdx.x < 32)
            for (int i = 0; i < SEG; ++i) bacc += uf[threadIdx.x * wt::US + i];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int k0 = 32 * c + 8 * q;  // this lane's first pixel of the chunk
            const float4* up = reinterpret_cast<const float4*>(uf + (16 * mb + l16) * wt::US + k0);
            const float4 u0 = up[0], u1 = up[1];
            const float u8[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
            // B: V at (y + dy, x0 + k0 + j + dx), tile row dy + 2, column k0 + j + dx + 2.
            uint32_t h[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t v0 = tap_ok ? vh[(dy + 2) * COLS + k0 + 2 * j + dx + 2] : 0u;
                const uint32_t v1 = tap_ok ? vh[(dy + 2) * COLS + k0 + 2 * j + 1 + dx + 2] : 0u;
                h[j] = v0 | (v1 << 16);
            }
            const bf16x8 B = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
            bf16x8 hi, lo;
            split8(u8, hi, lo);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi, B, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo, B, acc, 0, 0, 0);
        }
    }
    if ((int)threadIdx.x < 32) bpart[(size_t)blockIdx.x * 32 + threadIdx.x] = bacc;
    // D: rows 16 mb + 4 q + r (channel), column 16 nb + l16 (tap).
    float* dst = part + (size_t)blockIdx.x * 800;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int nn = 16 * nb + l16;
        if (nn < 25) dst[(16 * mb + 4 * q + r) * 25 + nn] = acc[r];
    }
}

// Few outputs over many chunks (biases, the thin layers' weights):
// out[c] = sum over chunks of part[chunk][c], one workgroup per output:
// thread t adds chunks t, t + 256, ... in order, then a fixed pairwise tree
// over the 256 threads (bitwise reproducible).
__global__ void __launch_bounds__(256) k_reduce_few(const float* __restrict__ part, int chunks, int C,
                                                    float* __restrict__ out) {
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float acc = 0.0f;
    for (int k = threadIdx.x; k < chunks; k += 256) acc += part[(size_t)k * C + c];
    sh[threadIdx.x] = acc;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// out[i] = sum over chunks of part[chunk][i], in chunk order.
__global__ void k_reduce_parts(const float* __restrict__ part, int chunks, int n, float* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int c = 0; c < chunks; ++c) acc += part[(size_t)c * n + i];
    out[i] = acc;
}

// ---------------------------------------------------------------- Adam
// theta -= lr * mhat / (sqrt(vhat) + eps); bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t.
__global__ void k_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Device copies of the master weights in the layouts the kernels read:
// W1x [32][1][32] (im2col channel t), W2 [64][25][32], W3 [32][25][64] as
// bf16; W2T [32][25][64] and W3T [64][25][32] tap-flipped and
// channel-swapped for the backward input gradients; W4b [25][32] bf16 for
// the output layer and W4 [25][32] as fp32 of the same bf16 values.
__global__ void k_pack_weights(const float* __restrict__ theta, PackedWeights w) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 32 * 32) {  // W1x[co][0][t]
        const int co = i / 32, t = i % 32;
        w.W1x[i] = t < 25 ? (uint16_t)bf16_bits(theta[OFF_W1 + co * 25 + t]) : (uint16_t)0;
    }
    if (i < 64 * 25 * 32) {  // W2[co][t][ci] and W2T[ci][24-t][co]
        const int ci = i % 32, t = (i / 32) % 25, co = i / (32 * 25);
        const uint16_t h = (uint16_t)bf16_bits(theta[OFF_W2 + i]);
        w.W2[i] = h;
        w.W2T[((size_t)ci * 25 + (24 - t)) * 64 + co] = h;
    }
    if (i < 32 * 25 * 64) {  // W3[co][t][ci] and W3T[ci][24-t][co]
        const int ci = i % 64, t = (i / 64) % 25, co = i / (64 * 25);
        const uint16_t h = (uint16_t)bf16_bits(theta[OFF_W3 + i]);
        w.W3[i] = h;
        w.W3T[((size_t)ci * 25 + (24 - t)) * 32 + co] = h;
    }
    if (i < 25 * 32) {
        const uint32_t h = bf16_bits(theta[OFF_W4 + i]);
        w.W4[i] = __uint_as_float(h << 16);
        w.W4b[i] = (uint16_t)h;
    }
}

// Crop a (side x side) window at (cx, cy) of a torus state into dst.
__global__ void k_crop(const float* __restrict__ state, int W, int H, const int2* __restrict__ corners, int side,
                       float* __restrict__ dst) {
    const int c = blockIdx.y;
    const int2 cc = corners[c];
    const size_t n = (size_t)side * side;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int xx = (int)(i % side), yy = (int)(i / side);
        dst[(size_t)c * n + i] = state[(size_t)((cc.y + yy) % H) * W + (cc.x + xx) % W];
    }
}

// Gather batch items: dst[b] = src[index[b]], each `n` floats.
__global__ void k_gather(const float* __restrict__ src, const int* __restrict__ index, size_t n,
                         float* __restrict__ dst) {
    const int b = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[(size_t)b * n + i] = src[(size_t)index[b] * n + i];
}

// FP64 partials of sum (a-b)^2 and sum b^2 (relative L2 of a against b).
__global__ void __launch_bounds__(256) k_rel_l2(const float* __restrict__ a, const float* __restrict__ b, size_t n,
                                                double* __restrict__ part) {
    __shared__ double s0[256], s1[256];
    double d2 = 0.0, b2 = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double d = (double)a[i] - (double)b[i];
        d2 += d * d;
        b2 += (double)b[i] * (double)b[i];
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = b2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s0[threadIdx.x] += s0[threadIdx.x + s];
            s1[threadIdx.x] += s1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

inline unsigned blocks(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void conv1(hipStream_t s, Profiler* p, const ConvArgs& a, const float* x, uint16_t* xb, uint16_t* pool) {
    if (a.W % 64 || a.H % 4 || a.Cout != 32 || !a.bias || !a.a || !xb || !pool)
        throw Error{AH_ERR_INVALID, "surrogate layer 1: unsupported shape"};
    Launch L(p, s, "surr_conv1");
    hipLaunchKernelGGL(k_conv1, dim3(a.W / 64, a.H / 4, a.N), dim3(256), 0, s, a, x, xb, pool);
    AH_HIP(hipGetLastError());
}

void conv(hipStream_t s, Profiler* p, const ConvArgs& a, const char* name) {
    if (a.W % 64 || a.H % 4 || a.Cin % 32 || a.ks != 5 || (a.Cout != 32 && a.Cout != 64))
        throw Error{AH_ERR_INVALID, "surrogate conv: unsupported shape"};
    if (a.mode == CONV_POOL_DGELU && (a.ks != 5 || a.up))
        throw Error{AH_ERR_INVALID, "surrogate conv: the pooled epilogue needs a 5x5 full-resolution input"};
    Launch L(p, s, name);
    dim3 grid(a.W / 64, a.H / 4, a.N);
    {
        // The upsampled 32-channel layer: 128 pixels per wave when the width
        // allows, so each weight fragment feeds 16 MFMAs (1008 -> 803 us per
        // call at batch 64 x 256^2; the half-resolution 32-channel input
        // gradient lost the same change, 300 -> 394 us).
        const dim3 grid8(a.W / 128, a.H / 4, a.N);
        if (a.Cout == 64 && a.up) hipLaunchKernelGGL((k_conv_lds<4, true, 4>), grid, dim3(256), 0, s, a);
        else if (a.Cout == 64) hipLaunchKernelGGL((k_conv_lds<4, false, 4>), grid, dim3(256), 0, s, a);
        else if (a.up && a.W % 128 == 0) hipLaunchKernelGGL((k_conv_lds<2, true, 8>), grid8, dim3(256), 0, s, a);
        else if (a.up) hipLaunchKernelGGL((k_conv_lds<2, true, 4>), grid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_conv_lds<2, false, 4>), grid, dim3(256), 0, s, a);
    }
    AH_HIP(hipGetLastError());
}

void out_layer(hipStream_t s, Profiler* p, const uint16_t* sa, const uint16_t* W4b, const float* b4, const float* x,
               float* r, float* y, int N, int H, int W) {
    if (W % ot::COLS || H % ot::ROWS || N < 1) throw Error{AH_ERR_INVALID, "surrogate layer 4: unsupported shape"};
    Launch L(p, s, "surr_out_layer");
    hipLaunchKernelGGL(k_out_taps, dim3(W / ot::COLS, H / ot::ROWS, N), dim3(256), 0, s, sa, W4b, b4, x, r, y, H, W);
    AH_HIP(hipGetLastError());
}

int loss_blocks(size_t total) { return (int)std::min<size_t>(1024, (total + 255) / 256); }

void loss(hipStream_t s, Profiler* p, const float* y, const float* t, const float* x, const float* r, float* dr,
          double* part, size_t total) {
    Launch L(p, s, "surr_loss");
    hipLaunchKernelGGL(k_loss, dim3(loss_blocks(total)), dim3(256), 0, s, y, t, x, r, dr, part, total);
    AH_HIP(hipGetLastError());
}

void dz3_from_dr(hipStream_t s, Profiler* p, const float* dr, const float* W4, const uint16_t* g3, uint16_t* out, int N,
                 int H, int W) {
    if (W % dz::SEG) throw Error{AH_ERR_INVALID, "surrogate dz3: width % 64 != 0"};
    Launch L(p, s, "surr_dz3");
    hipLaunchKernelGGL(k_dz3_dr, dim3(W / dz::SEG, H, N), dim3(256), 0, s, dr, W4, g3, out, N, H, W);
    AH_HIP(hipGetLastError());
}

void dz1_wgrad(hipStream_t s, Profiler* p, const float* dr, const float* W4, const float* dp, const uint16_t* g1,
               const uint16_t* xb, int N, int H, int W, float* part, float* out, float* bpart, float* bout) {
    if (W % wg::SEG || H % 2 || N < 1) throw Error{AH_ERR_INVALID, "surrogate layer 1: width % 64 != 0 or odd height"};
    const int chunks = wg::THIN_GROUPS;
    {
        Launch L(p, s, "surr_dz1_wgrad");
        hipLaunchKernelGGL(k_dz1_wgrad, dim3(chunks), dim3(256), 0, s, dr, W4, dp, g1, xb, N, H, W, part, bpart);
        AH_HIP(hipGetLastError());
    }
    {
        Launch L(p, s, "surr_reduce");
        hipLaunchKernelGGL(k_reduce_few, dim3(32 * 25), dim3(256), 0, s, part, chunks, 32 * 25, out);
        AH_HIP(hipGetLastError());
    }
    Launch L(p, s, "surr_reduce_bias");
    hipLaunchKernelGGL(k_reduce_few, dim3(32), dim3(256), 0, s, bpart, chunks, 32, bout);
    AH_HIP(hipGetLastError());
}

void wgrad(hipStream_t s, Profiler* p, const WGradArgs& a, float* part, float* out, float* bpart, float* bout) {
    const int nw = a.Cout * 25 * a.Cin;
    int chunks = 0;  // partial blocks of the chosen kernel
    const bool up = a.dz_h && a.xoff == 0 && a.up && a.Cout == wu::COUT && a.Cin == wu::CIN && a.H % 2 == 0 &&
                    (a.W / 2) % wu::SEG == 0;
    const bool mfma = a.dz_h && a.xoff == 0 && a.W % wg::SEG == 0 && a.Cout == 64 && a.Cin == 32 && !a.up;
    const bool thin = a.dz_f && a.W % wg::SEG == 0 && !a.up && a.Cout == 1 && a.Cin == 32;
    if (thin) {
        chunks = wg::THIN_GROUPS;
        Launch L(p, s, "surr_wgrad_thin");
        hipLaunchKernelGGL(k_wgrad_thin4, dim3(chunks), dim3(256), 0, s, a, part, bpart);
        AH_HIP(hipGetLastError());
    } else if (up) {
        chunks = 4 * wu::UPG;  // bias partials: workgroup x parity
        {
            Launch L(p, s, "surr_wgrad_up");
            hipLaunchKernelGGL(k_wgrad_up, dim3(wu::UPG, 4), dim3(wu::THREADS), 0, s, a, part, bpart);
            AH_HIP(hipGetLastError());
        }
        Launch L(p, s, "surr_reduce");
        hipLaunchKernelGGL(k_reduce_up, dim3(blocks(nw)), dim3(256), 0, s, part, wu::UPG, out);
        AH_HIP(hipGetLastError());
    } else if (mfma) {
        chunks = wg::GROUPS;
        Launch L(p, s, "surr_wgrad_mfma");
        hipLaunchKernelGGL((k_wgrad_mfma<64, 32, 0>), dim3(chunks), dim3(wg::THREADS), 0, s, a, part, bpart);
        AH_HIP(hipGetLastError());
    } else {
        throw Error{AH_ERR_INVALID, "surrogate weight gradient: unsupported layer shape"};
    }
    if (!up) {
        Launch L(p, s, "surr_reduce");
        if (thin) hipLaunchKernelGGL(k_reduce_few, dim3(nw), dim3(256), 0, s, part, chunks, nw, out);
        else hipLaunchKernelGGL(k_reduce_parts, dim3(blocks(nw)), dim3(256), 0, s, part, chunks, nw, out);
        AH_HIP(hipGetLastError());
    }
    // The bias gradient (the sum of dZ over the pixels): each kernel adds it
    // up from the dZ it stages (per-workgroup partials in bpart), one
    // reducing workgroup per channel.
    Launch L(p, s, "surr_reduce_bias");
    hipLaunchKernelGGL(k_reduce_few, dim3(a.Cout), dim3(256), 0, s, bpart, chunks, a.Cout, bout);
    AH_HIP(hipGetLastError());
}

void adam(hipStream_t s, Profiler* p, float* theta, float* m, float* v, const float* g, int n, float lr, float b1,
          float b2, float eps, float c1, float c2) {
    Launch L(p, s, "surr_adam");
    hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, s, theta, m, v, g, n, lr, b1, b2, eps, c1, c2);
    AH_HIP(hipGetLastError());
}

void pack_weights(hipStream_t s, Profiler* p, const float* theta, const PackedWeights& w) {
    Launch L(p, s, "surr_pack_weights");
    hipLaunchKernelGGL(k_pack_weights, dim3(blocks(64 * 25 * 32)), dim3(256), 0, s, theta, w);
    AH_HIP(hipGetLastError());
}

void crop(hipStream_t s, const float* state, int W, int H, const int2* corners, int count, int side, float* dst) {
    hipLaunchKernelGGL(k_crop, dim3(64, count), dim3(256), 0, s, state, W, H, corners, side, dst);
    AH_HIP(hipGetLastError());
}

void gather(hipStream_t s, const float* src, const int* index, int count, size_t n, float* dst) {
    hipLaunchKernelGGL(k_gather, dim3(64, count), dim3(256), 0, s, src, index, n, dst);
    AH_HIP(hipGetLastError());
}

int rel_l2_blocks(size_t n) { return (int)std::min<size_t>(1024, (n + 255) / 256); }

void rel_l2(hipStream_t s, const float* a, const float* b, size_t n, double* part) {
    hipLaunchKernelGGL(k_rel_l2, dim3(rel_l2_blocks(n)), dim3(256), 0, s, a, b, n, part);
    AH_HIP(hipGetLastError());
}

}  // namespace surr
}  // namespace ah
