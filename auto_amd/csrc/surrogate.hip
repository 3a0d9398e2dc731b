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
// a2, s, dz3, dz2 and the first layer's im2col), FP32 where only elementwise
// code reads them (pre-activations z and the backward input gradients).
//
// Hot kernels run on MFMA as an implicit GEMM, D[co][pixel] = sum over
// (tap, ci) of W[co][tap][ci] * X[pixel+tap][ci], with
// v_mfma_f32_16x16x32_bf16: A = 16 output channels x 32 input channels of one
// tap (one 16-byte load per lane from [co][tap][ci] weights), B = 32 input
// channels x 16 consecutive pixels (one 16-byte load per lane from NHWC).
// The upsampled input of layer 3 is never stored: the kernel reads the
// half-resolution tensor at (y/2, x/2).  The backward input gradients of
// layers 2 and 3 are the same kernel over tap-flipped, channel-swapped
// weights.  The 1-channel output layer and the weight gradients (reductions
// over pixels) are VALU kernels with fixed-order partial sums.
#include <algorithm>
#include <cmath>

#include "ah_internal.h"
#include "surrogate.h"

namespace ah {
namespace surr {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.3989422804014327f;
}
__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// ---------------------------------------------------------------- im2col
// x fp32 [N][H][W] -> X1 bf16 [N][H][W][32]: channel t < 25 holds x at tap t.
__global__ void k_im2col1(const float* __restrict__ x, uint16_t* __restrict__ X1, int N, int H, int W) {
    const size_t total = (size_t)N * H * W * 32;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = (int)(i & 31);
    const size_t pix = i >> 5;
    const int xx = (int)(pix % W), yy = (int)((pix / W) % H), n = (int)(pix / ((size_t)W * H));
    uint16_t v = 0;
    if (c < 25) {
        const int dy = c / 5 - 2, dx = c % 5 - 2;
        v = (uint16_t)bf16_bits(x[((size_t)n * H + wrap(yy + dy, H)) * W + wrap(xx + dx, W)]);
    }
    X1[i] = v;
}

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
            if (a.z) *reinterpret_cast<float4*>(a.z + o) = make_float4(z[0], z[1], z[2], z[3]);
            if (a.mode == CONV_GELU || a.mode == CONV_GELU_ADD) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = gelu(z[r]);
                if (a.mode == CONV_GELU_ADD) {
                    const uint2 ad = *reinterpret_cast<const uint2*>(a.add + o);
                    v[0] += __uint_as_float(ad.x << 16);
                    v[1] += __uint_as_float(ad.x & 0xffff0000u);
                    v[2] += __uint_as_float(ad.y << 16);
                    v[3] += __uint_as_float(ad.y & 0xffff0000u);
                }
                *reinterpret_cast<uint2*>(a.a + o) =
                    make_uint2(bf16_bits(v[0]) | (bf16_bits(v[1]) << 16), bf16_bits(v[2]) | (bf16_bits(v[3]) << 16));
            }
        }
}

// Wave: one output row, 64 consecutive pixels (4 blocks of 16), MT blocks of
// 16 output channels.  Workgroup: 4 waves = 4 rows.  Grid (W/64, H/4, N).
// Operands straight from global memory (the 1x1 first layer).
template <int MT>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, l16 = lane & 15;
    const int n = blockIdx.z, y = blockIdx.y * 4 + wave, x0 = blockIdx.x * 64;
    const int H = a.H, W = a.W, Cin = a.Cin, taps = a.ks * a.ks, half = a.ks / 2;
    const int Hx = a.up ? H / 2 : H, Wx = a.up ? W / 2 : W;
    const uint16_t* X = a.X + (size_t)n * Hx * Wx * Cin;
    f32x4 acc[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[m][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = 0; t < taps; ++t) {
        const int dy = t / a.ks - half, dx = t % a.ks - half;
        int yy = wrap(y + dy, H);
        if (a.up) yy >>= 1;
        int px[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int xx = wrap(x0 + 16 * b + l16 + dx, W);
            if (a.up) xx >>= 1;
            px[b] = (yy * Wx + xx) * Cin;
        }
        for (int c0 = 0; c0 < Cin; c0 += 32) {
            bf16x8 A[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m)
                A[m] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(a.Wt + ((size_t)(16 * m + l16) * taps + t) * Cin + c0 + 8 * q));
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bf16x8 B = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(X + px[b] + c0 + 8 * q));
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][b], 0, 0, 0);
            }
        }
    }
    conv_epilogue<MT, 4>(a, acc, n, y, x0, q, l16);
}

// The 5x5 convolutions: the same waves and MFMAs as k_conv, with the input
// staged in LDS.  Per 32-channel slice the workgroup loads its 8 x 68-pixel
// input window (4 output rows plus the 2-row, 2-column halo) once, and every
// tap's B fragments come from LDS: k_conv read them from the L1 cache for
// every tap, 25 times per input.  An upsampled layer (UP) stages the 4 x 34
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
        __syncthreads();
        for (int t = 0; t < 25; ++t) {
            const int dy = t / 5 - 2, dx = t % 5 - 2;
            bf16x8 A[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m)
                A[m] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(a.Wt + ((size_t)(16 * m + l16) * 25 + t) * Cin + c0 + 8 * q));
            const int r = wave + 2 + dy, c = 2 + dx + l16;  // window pixel of block 0
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int pix = UP ? (r >> 1) * SC + ((c + 16 * b) >> 1) : r * SC + c + 16 * b;
                const bf16x8 B = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(tile + pix * PS + 8 * q));
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[m][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][b], 0, 0, 0);
            }
        }
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
                        const float4 zz = *reinterpret_cast<const float4*>(a.zp + o);
                        const float g[4] = {v[0] * gelu_grad(zz.x), v[1] * gelu_grad(zz.y), v[2] * gelu_grad(zz.z),
                                            v[3] * gelu_grad(zz.w)};
                        *reinterpret_cast<uint2*>(a.dpool + o) = make_uint2(bf16_bits(g[0]) | (bf16_bits(g[1]) << 16),
                                                                            bf16_bits(g[2]) | (bf16_bits(g[3]) << 16));
                    }
                }
        }
        return;
    }
    conv_epilogue<MT, NB>(a, acc, n, y, x0, q, l16);
}

// ---------------------------------------------------------------- pooling
// a1 bf16 [N][H][W][32] -> p bf16 [N][H/2][W/2][32], mean of the 2x2 block.
__global__ void k_pool2(const uint16_t* __restrict__ a, uint16_t* __restrict__ p, int N, int H, int W, int C) {
    const int Hh = H / 2, Wh = W / 2;
    const size_t total = (size_t)N * Hh * Wh * C;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = (int)(i % C);
    const size_t pix = i / C;
    const int xh = (int)(pix % Wh), yh = (int)((pix / Wh) % Hh), n = (int)(pix / ((size_t)Wh * Hh));
    const size_t b = (((size_t)n * H + 2 * yh) * W + 2 * xh) * C + c;
    const float s = bf16f(a[b]) + bf16f(a[b + C]) + bf16f(a[b + (size_t)W * C]) + bf16f(a[b + (size_t)W * C + C]);
    p[i] = (uint16_t)bf16_bits(0.25f * s);
}

// ---------------------------------------------------------------- layer 4
// r = conv5(s; W4 32->1) + b4, y = clip(x + r, 0, 1).  W4 fp32 [25][32]
// (already bf16-rounded values).  One thread per pixel.
__global__ void k_out_layer(const uint16_t* __restrict__ s, const float* __restrict__ W4, const float* __restrict__ b4p,
                            const float* __restrict__ x, float* __restrict__ r, float* __restrict__ y, int N, int H,
                            int W) {
    __shared__ float w[25 * 32];
    for (int i = threadIdx.x; i < 25 * 32; i += blockDim.x) w[i] = W4[i];
    __syncthreads();
    const size_t total = (size_t)N * H * W;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int xx = (int)(i % W), yy = (int)((i / W) % H), n = (int)(i / ((size_t)W * H));
    float acc = *b4p;
    for (int t = 0; t < 25; ++t) {
        const int dy = t / 5 - 2, dx = t % 5 - 2;
        const uint4* sp = reinterpret_cast<const uint4*>(
            s + (((size_t)n * H + wrap(yy + dy, H)) * W + wrap(xx + dx, W)) * 32);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint4 v = sp[g];
            const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc += w[t * 32 + 8 * g + 2 * j] * __uint_as_float(e[j] << 16);
                acc += w[t * 32 + 8 * g + 2 * j + 1] * __uint_as_float(e[j] & 0xffff0000u);
            }
        }
    }
    r[i] = acc;
    y[i] = fminf(fmaxf(x[i] + acc, 0.0f), 1.0f);
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

// ds = d(loss)/ds for layer 4: ds[p][ci] = sum_t W4[t][ci] dr[p - tap_t].
// One thread per pixel: its 25 dr values serve all 32 channels (taps in the
// same order as before, so each channel's sum is unchanged).
__global__ void __launch_bounds__(256) k_out_input_grad(const float* __restrict__ dr, const float* __restrict__ W4,
                                                        float* __restrict__ ds, int N, int H, int W) {
    __shared__ float w[25 * 32];
    for (int i = threadIdx.x; i < 25 * 32; i += blockDim.x) w[i] = W4[i];
    __syncthreads();
    const size_t total = (size_t)N * H * W;
    const size_t pix = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= total) return;
    const int xx = (int)(pix % W), yy = (int)((pix / W) % H), n = (int)(pix / ((size_t)W * H));
    float acc[32];
#pragma unroll
    for (int ci = 0; ci < 32; ++ci) acc[ci] = 0.0f;
    for (int t = 0; t < 25; ++t) {
        const int dy = t / 5 - 2, dx = t % 5 - 2;
        const float d = dr[((size_t)n * H + wrap(yy - dy, H)) * W + wrap(xx - dx, W)];
#pragma unroll
        for (int ci = 0; ci < 32; ++ci) acc[ci] += w[t * 32 + ci] * d;
    }
    float4* o = reinterpret_cast<float4*>(ds + pix * 32);
#pragma unroll
    for (int g = 0; g < 8; ++g) o[g] = make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
}

// dz3 = bf16(ds * gelu'(z3))
__global__ void k_dz3(const float* __restrict__ ds, const float* __restrict__ z3, uint16_t* __restrict__ dz3,
                      size_t total) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) dz3[i] = (uint16_t)bf16_bits(ds[i] * gelu_grad(z3[i]));
}

// dz2 = bf16((sum of du over the 2x2 children) * gelu'(z2)), half resolution, C = 64.

// dz1 = (ds + dp[p/2] / 4) * gelu'(z1), fp32, C = 32.
__global__ void k_dz1(const float* __restrict__ ds, const float* __restrict__ dp, const float* __restrict__ z1,
                      float* __restrict__ dz1, int N, int H, int W, int C) {
    const size_t total = (size_t)N * H * W * C;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = (int)(i % C);
    const size_t pix = i / C;
    const int xx = (int)(pix % W), yy = (int)((pix / W) % H), n = (int)(pix / ((size_t)W * H));
    const size_t ph = (((size_t)n * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c;
    dz1[i] = (ds[i] + 0.25f * dp[ph]) * gelu_grad(z1[i]);
}

// ---------------------------------------------------------------- weight grads
// dW[co][t][ci] = sum over pixels of dZ[pix][co] * X[pix + tap_t][ci].
// Thread = 4 output channels x one (t, ci); blockIdx.y = pixel chunk; the
// chunk partials go to part[chunk][co][t][ci] for a fixed-order reduction.
// dZ is fp32 (dz_f) or bf16 (dz_h); X is bf16 NHWC with channel stride
// xstride (the first layer reads channel 12, tap (0,0), of its im2col), at
// half resolution when `up` (the layer-3 input u = upsample2(a2)).
__global__ void __launch_bounds__(256) k_wgrad(WGradArgs a, float* __restrict__ part) {
    const int Cq = (a.Cout + 3) / 4;
    const int nout = Cq * 25 * a.Cin;
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= nout) return;
    const int ci = o % a.Cin, t = (o / a.Cin) % 25, cq = o / (a.Cin * 25);
    const int dy = t / 5 - 2, dx = t % 5 - 2;
    const size_t npix = (size_t)a.N * a.H * a.W;
    const size_t per = (npix + gridDim.y - 1) / gridDim.y;
    const size_t p0 = per * blockIdx.y, p1 = p0 + per < npix ? p0 + per : npix;
    const int Hx = a.up ? a.H / 2 : a.H, Wx = a.up ? a.W / 2 : a.W;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // Pixel coordinates advance with the pixel index (one division per chunk).
    int xx = (int)(p0 % a.W), yy = (int)((p0 / a.W) % a.H), n = (int)(p0 / ((size_t)a.W * a.H));
    for (size_t pix = p0; pix < p1; ++pix) {
        int sy = wrap(yy + dy, a.H), sx = wrap(xx + dx, a.W);
        if (a.up) {
            sy >>= 1;
            sx >>= 1;
        }
        const float xv = bf16f(a.X[(((size_t)n * Hx + sy) * Wx + sx) * a.xstride + a.xoff + ci]);
        if (a.Cout == 1) {
            acc[0] += a.dz_f[pix] * xv;
        } else if (a.dz_f) {
            const float4 g = *reinterpret_cast<const float4*>(a.dz_f + pix * a.Cout + 4 * cq);
            acc[0] += g.x * xv;
            acc[1] += g.y * xv;
            acc[2] += g.z * xv;
            acc[3] += g.w * xv;
        } else {
            const uint2 g = *reinterpret_cast<const uint2*>(a.dz_h + pix * a.Cout + 4 * cq);
            acc[0] += __uint_as_float(g.x << 16) * xv;
            acc[1] += __uint_as_float(g.x & 0xffff0000u) * xv;
            acc[2] += __uint_as_float(g.y << 16) * xv;
            acc[3] += __uint_as_float(g.y & 0xffff0000u) * xv;
        }
        if (++xx == a.W) {
            xx = 0;
            if (++yy == a.H) {
                yy = 0;
                ++n;
            }
        }
    }
    float* dst = part + (size_t)blockIdx.y * a.Cout * 25 * a.Cin;
    const int nco = a.Cout == 1 ? 1 : 4;
    for (int j = 0; j < nco; ++j) dst[((size_t)(4 * cq + j) * 25 + t) * a.Cin + ci] = acc[j];
}

// Weight gradients of the two 32 <-> 64 channel layers on MFMA: a
// transposed implicit GEMM, dW[co][t][ci] = sum over pixels p of
// dZ[p][co] X[p + tap_t][ci], with the pixels as the MFMA's K dimension.
// v_mfma_f32_16x16x32_bf16: A = 16 output channels x 32 pixels of dZ, B = 32
// pixels x 16 input channels of X shifted by the tap, D = 16 x 16 of dW.
// A workgroup (8 waves) walks a fixed range of 64-pixel row segments; per
// segment it stages dZ (64 pixels x COUT) and the X tile the 25 taps reach
// (5 rows x 68 columns x CIN, read at half resolution when UP) in LDS, and
// wave w accumulates the (16 co x 16 ci) block w of all 25 taps.  A lane's
// operands are 8 pixels of one channel, read from the pixel-major LDS copies
// 2 bytes at a time (the row strides are padded by 8 channels so the 4 pixel
// groups of a read land on distinct banks).  Partials go to
// part[workgroup][co][t][ci] and are summed in workgroup order afterwards:
// the result does not depend on timing.
namespace wg {
constexpr int SEG = 64, ROWS = 5, COLS = SEG + 4, THREADS = 512, GROUPS = 256;
// Workgroups of the thin-layer kernels: 800 outputs each, so many more
// workgroups than the MFMA kernels' (256 left 4 waves per CU).
constexpr int THIN_GROUPS = 2048;
}

template <int COUT, int CIN, int UP>
__global__ void __launch_bounds__(wg::THREADS) k_wgrad_mfma(WGradArgs a, float* __restrict__ part,
                                                            float* __restrict__ bpart) {
    using namespace wg;
    static_assert((COUT / 16) * (CIN / 16) == THREADS / 64, "one 16x16 block per wave");
    constexpr int XS = CIN + 8, ZS = COUT + 8;  // LDS strides per pixel (bf16)
    __shared__ __attribute__((aligned(16))) uint16_t xt[ROWS * COLS * XS];
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
    for (long sg = s0; sg < s1; ++sg) {
        const int xb = (int)(sg % segs_per_row);
        const long ry = sg / segs_per_row;
        const int y = (int)(ry % H), n = (int)(ry / H), x0 = xb * SEG;
        __syncthreads();  // the previous segment's LDS reads are done
        // dZ: 64 pixels x COUT, contiguous in global memory.
        {
            const uint16_t* src = a.dz_h + (((size_t)n * H + y) * W + x0) * COUT;
            for (int v = threadIdx.x; v < SEG * COUT / 8; v += THREADS) {
                const int px = v / (COUT / 8), part8 = v % (COUT / 8);
                *reinterpret_cast<uint4*>(dz + px * ZS + 8 * part8) =
                    *reinterpret_cast<const uint4*>(src + (size_t)px * COUT + 8 * part8);
            }
        }
        // X tile: rows y-2..y+2, columns x0-2..x0+65, wrapped on the torus.
        for (int v = threadIdx.x; v < ROWS * COLS * CIN / 8; v += THREADS) {
            const int pidx = v / (CIN / 8), part8 = v % (CIN / 8);
            const int r = pidx / COLS, c = pidx % COLS;
            int sy = wrap(y + r - 2, H), sx = wrap(x0 + c - 2, W);
            if (UP) {
                sy >>= 1;
                sx >>= 1;
            }
            *reinterpret_cast<uint4*>(xt + pidx * XS + 8 * part8) = *reinterpret_cast<const uint4*>(
                a.X + (((size_t)n * Hx + sy) * Wx + sx) * a.xstride + a.xoff + 8 * part8);
        }
        __syncthreads();
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
            for (int t = 0; t < 25; ++t) {
                const int dy = t / 5, dx = t % 5;  // tile row and column offsets (tap + 2)
                const uint16_t* xb0 = xt + (dy * COLS + p0 + dx) * XS + 16 * cib + l16;
                uint32_t bb[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) bb[j] = (uint32_t)xb0[(2 * j) * XS] | ((uint32_t)xb0[(2 * j + 1) * XS] << 16);
                const bf16x8 B = __builtin_bit_cast(bf16x8, make_uint4(bb[0], bb[1], bb[2], bb[3]));
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
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

// Weight gradients of the two thin layers, out[a][t][b] = sum over pixels p
// of U[p][a] V[p + tap_t][b], with U the fp32 output gradient and V the bf16
// layer input: layer 1 (A = 32 output channels, B = 1: V is the input x,
// channel 12 of its im2col) and layer 4 (A = 1, B = 32).  A workgroup walks a
// fixed range of 64-pixel row segments, stages U and the 5 x 68 V tile the
// taps reach in LDS as fp32, and thread (t, quad) accumulates the four
// outputs a = 4 quad .. (A = 32) or b = 4 quad .. (B = 32) of tap t over the
// segment's pixels in order.  Partials go to part[workgroup][a][t][b].
template <int A, int B>
__global__ void __launch_bounds__(256) k_wgrad_thin(WGradArgs a, float* __restrict__ part, float* __restrict__ bpart) {
    using wg::SEG;
    using wg::ROWS;
    using wg::COLS;
    static_assert((A == 32 && B == 1) || (A == 1 && B == 32), "thin layers");
    __shared__ __attribute__((aligned(16))) float u[SEG * A];
    __shared__ __attribute__((aligned(16))) float v[ROWS * COLS * B];
    const int H = a.H, W = a.W, segs_per_row = W / SEG;
    const long nseg = (long)a.N * H * segs_per_row;
    const long s0 = nseg * blockIdx.x / gridDim.x, s1 = nseg * (blockIdx.x + 1) / gridDim.x;
    const int t = threadIdx.x / 8, quad = threadIdx.x % 8;  // t < 25 for the 200 working threads
    const bool active = t < 25;
    const int dy = t / 5, dx = t % 5;  // tile offsets of the tap (tap + 2)
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float bacc = 0.0f;  // bias gradient of output channel threadIdx.x (< A) over this workgroup's pixels
    for (long sg = s0; sg < s1; ++sg) {
        const int xb = (int)(sg % segs_per_row);
        const long ry = sg / segs_per_row;
        const int y = (int)(ry % H), n = (int)(ry / H), x0 = xb * SEG;
        __syncthreads();  // the previous segment's reads are done
        for (int i = threadIdx.x; i < SEG * A; i += 256)
            u[i] = a.dz_f[(((size_t)n * H + y) * W + x0) * A + i];
        if constexpr (B == 1) {
            for (int i = threadIdx.x; i < ROWS * COLS; i += 256) {
                const int r = i / COLS, cc = i % COLS;
                const int sy = wrap(y + r - 2, H), sx = wrap(x0 + cc - 2, W);
                v[i] = bf16f(a.X[(((size_t)n * H + sy) * W + sx) * a.xstride + a.xoff]);
            }
        } else {
            // 16 bytes (8 channels) per load, widened to two float4.
            for (int i = threadIdx.x; i < ROWS * COLS * (B / 8); i += 256) {
                const int pidx = i / (B / 8), c8 = i % (B / 8), r = pidx / COLS, cc = pidx % COLS;
                const int sy = wrap(y + r - 2, H), sx = wrap(x0 + cc - 2, W);
                const uint4 h = *reinterpret_cast<const uint4*>(
                    a.X + (((size_t)n * H + sy) * W + sx) * a.xstride + a.xoff + 8 * c8);
                float4* o = reinterpret_cast<float4*>(v + pidx * B + 8 * c8);
                o[0] = make_float4(__uint_as_float(h.x << 16), __uint_as_float(h.x & 0xffff0000u),
                                   __uint_as_float(h.y << 16), __uint_as_float(h.y & 0xffff0000u));
                o[1] = make_float4(__uint_as_float(h.z << 16), __uint_as_float(h.z & 0xffff0000u),
                                   __uint_as_float(h.w << 16), __uint_as_float(h.w & 0xffff0000u));
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < A)
            for (int p = 0; p < SEG; ++p) bacc += u[p * A + threadIdx.x];
        if (active) {
            for (int p = 0; p < SEG; ++p) {
                const int vp = (dy * COLS + p + dx) * B;  // V at pixel p + tap
                if constexpr (A == 32) {
                    const float4 uu = *reinterpret_cast<const float4*>(u + p * A + 4 * quad);
                    const float vv = v[vp];
                    acc[0] += uu.x * vv;
                    acc[1] += uu.y * vv;
                    acc[2] += uu.z * vv;
                    acc[3] += uu.w * vv;
                } else {
                    const float uu = u[p];
                    const float4 vv = *reinterpret_cast<const float4*>(v + vp + 4 * quad);
                    acc[0] += uu * vv.x;
                    acc[1] += uu * vv.y;
                    acc[2] += uu * vv.z;
                    acc[3] += uu * vv.w;
                }
            }
        }
    }
    if ((int)threadIdx.x < A) bpart[(size_t)blockIdx.x * A + threadIdx.x] = bacc;
    if (active) {
        float* dst = part + (size_t)blockIdx.x * A * 25 * B;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (A == 32) dst[(size_t)(4 * quad + k) * 25 + t] = acc[k];
            else dst[(size_t)t * B + 4 * quad + k] = acc[k];
        }
    }
}

// Bias gradient partials: part[chunk][c] = sum over the chunk's pixels of
// dZ[pix][c].  C divides 256: thread j sums channel j % C over the chunk's
// pixels j / C, j / C + 256 / C, ...; the 256 / C lanes of a channel are then
// added in lane order through LDS (a fixed order: bitwise reproducible).
__global__ void __launch_bounds__(256) k_bgrad(const float* __restrict__ dz_f, const uint16_t* __restrict__ dz_h,
                                               int C, size_t npix, float* __restrict__ part) {
    __shared__ float sh[256];
    const size_t per = (npix + gridDim.x - 1) / gridDim.x;
    const size_t p0 = per * blockIdx.x, p1 = p0 + per < npix ? p0 + per : npix;
    const int c = threadIdx.x % C, stride = 256 / C;
    float acc = 0.0f;
    for (size_t pix = p0 + threadIdx.x / C; pix < p1; pix += stride)
        acc += dz_f ? dz_f[pix * C + c] : bf16f(dz_h[pix * C + c]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    if ((int)threadIdx.x < C) {
        float t = 0.0f;
        for (int j = 0; j < stride; ++j) t += sh[j * C + threadIdx.x];
        part[(size_t)blockIdx.x * C + threadIdx.x] = t;
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
// channel-swapped for the backward input gradients; W4 [25][32] as fp32 of
// the bf16-rounded values.
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
    if (i < 25 * 32) w.W4[i] = __uint_as_float(bf16_bits(theta[OFF_W4 + i]) << 16);
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

void im2col1(hipStream_t s, Profiler* p, const float* x, uint16_t* X1, int N, int H, int W) {
    Launch L(p, s, "surr_im2col1");
    hipLaunchKernelGGL(k_im2col1, dim3(blocks((size_t)N * H * W * 32)), dim3(256), 0, s, x, X1, N, H, W);
    AH_HIP(hipGetLastError());
}

void conv(hipStream_t s, Profiler* p, const ConvArgs& a, const char* name) {
    if (a.W % 64 || a.H % 4 || a.Cin % 32 || (a.Cout != 32 && a.Cout != 64) || (a.ks != 1 && a.ks != 5))
        throw Error{AH_ERR_INVALID, "surrogate conv: unsupported shape"};
    if (a.mode == CONV_POOL_DGELU && (a.ks != 5 || a.up))
        throw Error{AH_ERR_INVALID, "surrogate conv: the pooled epilogue needs a 5x5 full-resolution input"};
    Launch L(p, s, name);
    dim3 grid(a.W / 64, a.H / 4, a.N);
    if (a.ks == 5) {
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
    } else if (a.Cout == 64) {
        hipLaunchKernelGGL(k_conv<4>, grid, dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_conv<2>, grid, dim3(256), 0, s, a);
    }
    AH_HIP(hipGetLastError());
}

void pool2(hipStream_t s, Profiler* p, const uint16_t* a, uint16_t* out, int N, int H, int W, int C) {
    Launch L(p, s, "surr_pool2");
    hipLaunchKernelGGL(k_pool2, dim3(blocks((size_t)N * H * W * C / 4)), dim3(256), 0, s, a, out, N, H, W, C);
    AH_HIP(hipGetLastError());
}

void out_layer(hipStream_t s, Profiler* p, const uint16_t* sa, const float* W4, const float* b4, const float* x,
               float* r, float* y, int N, int H, int W) {
    Launch L(p, s, "surr_out_layer");
    hipLaunchKernelGGL(k_out_layer, dim3(blocks((size_t)N * H * W)), dim3(256), 0, s, sa, W4, b4, x, r, y, N, H, W);
    AH_HIP(hipGetLastError());
}

int loss_blocks(size_t total) { return (int)std::min<size_t>(1024, (total + 255) / 256); }

void loss(hipStream_t s, Profiler* p, const float* y, const float* t, const float* x, const float* r, float* dr,
          double* part, size_t total) {
    Launch L(p, s, "surr_loss");
    hipLaunchKernelGGL(k_loss, dim3(loss_blocks(total)), dim3(256), 0, s, y, t, x, r, dr, part, total);
    AH_HIP(hipGetLastError());
}

void out_input_grad(hipStream_t s, Profiler* p, const float* dr, const float* W4, float* ds, int N, int H, int W) {
    Launch L(p, s, "surr_out_input_grad");
    hipLaunchKernelGGL(k_out_input_grad, dim3(blocks((size_t)N * H * W)), dim3(256), 0, s, dr, W4, ds, N, H, W);
    AH_HIP(hipGetLastError());
}

void dz3(hipStream_t s, Profiler* p, const float* ds, const float* z3, uint16_t* out, size_t total) {
    Launch L(p, s, "surr_dz3");
    hipLaunchKernelGGL(k_dz3, dim3(blocks(total)), dim3(256), 0, s, ds, z3, out, total);
    AH_HIP(hipGetLastError());
}

void dz1(hipStream_t s, Profiler* p, const float* ds, const float* dp, const float* z1, float* out, int N, int H, int W,
         int C) {
    Launch L(p, s, "surr_dz1");
    hipLaunchKernelGGL(k_dz1, dim3(blocks((size_t)N * H * W * C)), dim3(256), 0, s, ds, dp, z1, out, N, H, W, C);
    AH_HIP(hipGetLastError());
}

void wgrad(hipStream_t s, Profiler* p, const WGradArgs& a, float* part, int chunks, float* out, float* bpart,
           float* bout) {
    const int Cq = (a.Cout + 3) / 4;
    const int nout = Cq * 25 * a.Cin, nw = a.Cout * 25 * a.Cin;
    const bool mfma = a.dz_h && a.xoff == 0 && a.W % wg::SEG == 0 &&
                      ((a.Cout == 32 && a.Cin == 64 && a.up) || (a.Cout == 64 && a.Cin == 32 && !a.up));
    const bool thin = a.dz_f && a.W % wg::SEG == 0 && !a.up &&
                      ((a.Cout == 32 && a.Cin == 1) || (a.Cout == 1 && a.Cin == 32));
    if (thin) {
        chunks = wg::THIN_GROUPS;
        Launch L(p, s, "surr_wgrad_thin");
        if (a.Cout == 32)
            hipLaunchKernelGGL((k_wgrad_thin<32, 1>), dim3(chunks), dim3(256), 0, s, a, part, bpart);
        else
            hipLaunchKernelGGL((k_wgrad_thin<1, 32>), dim3(chunks), dim3(256), 0, s, a, part, bpart);
        AH_HIP(hipGetLastError());
    } else if (mfma) {
        chunks = wg::GROUPS;
        Launch L(p, s, "surr_wgrad_mfma");
        if (a.Cout == 32)
            hipLaunchKernelGGL((k_wgrad_mfma<32, 64, 1>), dim3(chunks), dim3(wg::THREADS), 0, s, a, part, bpart);
        else
            hipLaunchKernelGGL((k_wgrad_mfma<64, 32, 0>), dim3(chunks), dim3(wg::THREADS), 0, s, a, part, bpart);
        AH_HIP(hipGetLastError());
    } else {
        Launch L(p, s, "surr_wgrad");
        hipLaunchKernelGGL(k_wgrad, dim3(blocks(nout), chunks), dim3(256), 0, s, a, part);
        AH_HIP(hipGetLastError());
    }
    {
        Launch L(p, s, "surr_reduce");
        if (thin) hipLaunchKernelGGL(k_reduce_few, dim3(nw), dim3(256), 0, s, part, chunks, nw, out);
        else hipLaunchKernelGGL(k_reduce_parts, dim3(blocks(nw)), dim3(256), 0, s, part, chunks, nw, out);
        AH_HIP(hipGetLastError());
    }
    // The bias gradient (the sum of dZ over the pixels): the thin and MFMA
    // kernels add it up from the dZ they stage (per-workgroup partials in
    // bpart); the generic kernel leaves it to a pass of its own.
    if (thin || mfma) {
        Launch L(p, s, "surr_reduce_bias");
        hipLaunchKernelGGL(k_reduce_few, dim3(a.Cout), dim3(256), 0, s, bpart, chunks, a.Cout, bout);
        AH_HIP(hipGetLastError());
    } else {
        bgrad(s, p, a.dz_f, a.dz_h, a.Cout, (size_t)a.N * a.H * a.W, part, 1024, bout);
    }
}

void bgrad(hipStream_t s, Profiler* p, const float* dz_f, const uint16_t* dz_h, int C, size_t npix, float* part,
           int chunks, float* out) {
    {
        Launch L(p, s, "surr_bgrad");
        if (C < 1 || 256 % C) throw Error{AH_ERR_INVALID, "bias gradient: channels must divide 256"};
        hipLaunchKernelGGL(k_bgrad, dim3(chunks), dim3(256), 0, s, dz_f, dz_h, C, npix, part);
        AH_HIP(hipGetLastError());
    }
    // One workgroup per channel: a lone thread per channel walking all the
    // chunks in turn took 183 us per call on its chain of dependent loads.
    Launch L(p, s, "surr_reduce_bias");
    hipLaunchKernelGGL(k_reduce_few, dim3(C), dim3(256), 0, s, part, chunks, C, out);
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
