// Surrogate network: shapes, parameter layout and kernel launchers.
#pragma once

#include "ah_internal.h"

namespace ah {
namespace surr {

constexpr int C1 = 32, C2 = 64;
// Flat FP32 parameter vector: W1 [32][25][1], b1 [32], W2 [64][25][32], b2 [64],
// W3 [32][25][64], b3 [32], W4 [1][25][32], b4 [1].
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + C1 * 25;
constexpr int OFF_W2 = OFF_B1 + C1;
constexpr int OFF_B2 = OFF_W2 + C2 * 25 * C1;
constexpr int OFF_W3 = OFF_B2 + C2;
constexpr int OFF_B3 = OFF_W3 + C1 * 25 * C2;
constexpr int OFF_W4 = OFF_B3 + C1;
constexpr int OFF_B4 = OFF_W4 + 25 * C1;
constexpr int N_PARAMS = OFF_B4 + 1;  // 104129

// CONV_POOL_DGELU (the layer-3 input gradient): the output, summed over
// 2 x 2 pixel blocks (the adjoint of the nearest upsampling) and multiplied
// by dgelu_in (gelu' of layer 2's pre-activation), goes to dpool at half
// resolution; nothing else is written.
enum ConvMode { CONV_LINEAR = 0, CONV_GELU = 1, CONV_GELU_ADD = 2, CONV_POOL_DGELU = 3 };

struct ConvArgs {
    const uint16_t* X;   // bf16 NHWC input, Cin channels (half resolution when up)
    int N, H, W;         // output size (full size of the input view)
    int Cin, up, ks;     // Cin multiple of 32; ks 1 or 5
    const uint16_t* Wt;  // bf16 [Cout][ks*ks][Cin]
    int Cout;            // 32 or 64
    const float* bias;   // [Cout] or null
    int mode;            // ConvMode
    float* z;            // fp32 [N][H][W][Cout] or null: the output of CONV_LINEAR
    uint16_t* dgelu;     // bf16 [N][H][W][Cout] or null: gelu'(z) of the GELU modes (kept for training)
    uint16_t* a;         // bf16 [N][H][W][Cout]: gelu(z) (+ add)
    const uint16_t* add; // bf16 [N][H][W][Cout] for CONV_GELU_ADD
    const uint16_t* dgelu_in;  // bf16 [N][H/2][W/2][Cout] for CONV_POOL_DGELU
    uint16_t* dpool;     // bf16 [N][H/2][W/2][Cout] for CONV_POOL_DGELU
};

struct WGradArgs {
    const float* dz_f;     // fp32 [pix][Cout] (or [pix] when Cout == 1)
    const uint16_t* dz_h;  // bf16 [pix][Cout] when dz_f is null
    const uint16_t* X;     // bf16 NHWC layer input, channel stride xstride
    int xstride, xoff;     // channel ci of pixel p at X[p*xstride + xoff + ci]
    int N, H, W, Cin, Cout, up;
};

struct PackedWeights {
    uint16_t* W1x;  // [32][1][32]
    uint16_t* W2;   // [64][25][32]
    uint16_t* W2T;  // [32][25][64]
    uint16_t* W3;   // [32][25][64]
    uint16_t* W3T;  // [64][25][32]
    uint16_t* W4b;  // [25][32]
    float* W4;      // [25][32], the same bf16 values as fp32
};

// Layer 1 (5x5, 1 -> 32, bias, GELU) straight from the fp32 input x; also
// writes xb = bf16(x), the input the layer-1 weight gradient reads, and the
// 2 x 2 average pool of the activations (bf16, half resolution).
void conv1(hipStream_t s, Profiler* p, const ConvArgs& a, const float* x, uint16_t* xb, uint16_t* pool);
void conv(hipStream_t s, Profiler* p, const ConvArgs& a, const char* name);
// Layer 4 (5x5, 32 -> 1) from the bf16 NHWC activations sa: r = conv + b4
// and y = clip(x + r, 0, 1), fp32 [N][H][W] each.
void out_layer(hipStream_t s, Profiler* p, const uint16_t* sa, const uint16_t* W4b, const float* b4, const float* x,
               float* r, float* y, int N, int H, int W);
int loss_blocks(size_t total);
void loss(hipStream_t s, Profiler* p, const float* y, const float* t, const float* x, const float* r, float* dr,
          double* part, size_t total);
// The layer-4 input gradient ds = sum over taps of W4 * dr is never stored:
// dz3_from_dr forms it and writes dz3 = bf16(ds * g3); dz1_wgrad forms the
// same values again, dz1 = (ds + dp / 4) * g1 (never stored), and from it
// the layer-1 weight and bias gradients into out and bout, with g1 and g3
// the gelu' values the forward pass kept (bf16) and xb = bf16(x).
void dz3_from_dr(hipStream_t s, Profiler* p, const float* dr, const float* W4, const uint16_t* g3, uint16_t* out, int N,
                 int H, int W);
void dz1_wgrad(hipStream_t s, Profiler* p, const float* dr, const float* W4, const float* dp, const uint16_t* g1,
               const uint16_t* xb, int N, int H, int W, float* part, float* out, float* bpart, float* bout);
// Weight gradient `out` and bias gradient `bout` of one layer; `part`
// (WGRAD_PART_FLOATS) and `bpart` (WGRAD_BPART_FLOATS) hold per-workgroup
// partials.  Layer 4 (32 -> 1, fp32 dz_f), layer 2 (32 -> 64) and layer 3
// (64 -> 32 over the upsampled input) each have their kernel (layer 1's is
// dz1_wgrad); other shapes are an error.
constexpr size_t WGRAD_PART_FLOATS = (size_t)256 * 64 * 25 * 32;
constexpr size_t WGRAD_BPART_FLOATS = (size_t)2048 * 64;
void wgrad(hipStream_t s, Profiler* p, const WGradArgs& a, float* part, float* out, float* bpart, float* bout);
void adam(hipStream_t s, Profiler* p, float* theta, float* m, float* v, const float* g, int n, float lr, float b1,
          float b2, float eps, float c1, float c2);
void pack_weights(hipStream_t s, Profiler* p, const float* theta, const PackedWeights& w);
void crop(hipStream_t s, const float* state, int W, int H, const int2* corners, int count, int side, float* dst);
void gather(hipStream_t s, const float* src, const int* index, int count, size_t n, float* dst);
int rel_l2_blocks(size_t n);
void rel_l2(hipStream_t s, const float* a, const float* b, size_t n, double* part);

}  // namespace surr
}  // namespace ah
