// The view pass: the state as rgba8 pixels through the rule's colour ramp.
//
// Per output pixel, n x n sub-samples (n = clamp(ceil(scale), 1, 4)) of the
// world at origin + (pixel + (s + 0.5)/n) * scale; samples outside the world
// stay black; colours are averaged as sqrt(mean(c^2)) ("in light rather than
// in gamma"), as VIEW_WGSL does (crates/auto-core/src/sim.rs:17-41).  The
// palettes are the rules' shade() functions (spectral.rs:346-350, 407-413,
// 473-478) over program.rs:148-156's ramp and mix3.  Output is rgba8 with
// alpha 255, row-major [out_h][out_w][4]; unorm conversion rounds to nearest.
#include <algorithm>
#include <cmath>

#include "ah_internal.h"

namespace ah {

namespace {

struct Palette {
    float3 c[5];
    int stops;  // 5 for ramp, 2 for mix3
};

__device__ __forceinline__ float3 mix3(float3 a, float3 b, float t) {
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return make_float3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}

__device__ __forceinline__ float3 shade(const Palette& p, float s) {
    if (p.stops == 2) return mix3(p.c[0], p.c[1], s);
    const float t = fminf(fmaxf(s, 0.0f), 1.0f) * 4.0f;
    if (t < 1.0f) return mix3(p.c[0], p.c[1], t);
    if (t < 2.0f) return mix3(p.c[1], p.c[2], t - 1.0f);
    if (t < 3.0f) return mix3(p.c[2], p.c[3], t - 2.0f);
    return mix3(p.c[3], p.c[4], t - 3.0f);
}

__device__ __forceinline__ uint32_t unorm8(float v) {
    return (uint32_t)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f + 0.5f);
}

__global__ void __launch_bounds__(256) k_render(const float* __restrict__ state, int W, int H, Palette pal,
                                                float ox, float oy, float scale, int n, int out_w, int out_h,
                                                uint32_t* __restrict__ rgba) {
    const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (px >= out_w || py >= out_h) return;
    float3 acc = make_float3(0.0f, 0.0f, 0.0f);
    for (int sy = 0; sy < n; ++sy)
        for (int sx = 0; sx < n; ++sx) {
            const float x = ox + ((float)px + ((float)sx + 0.5f) / (float)n) * scale;
            const float y = oy + ((float)py + ((float)sy + 0.5f) / (float)n) * scale;
            if (x >= 0.0f && y >= 0.0f && x < (float)W && y < (float)H) {
                const int cx = min((int)x, W - 1), cy = min((int)y, H - 1);
                const float3 c = shade(pal, state[(size_t)cy * W + cx]);
                acc.x += c.x * c.x;
                acc.y += c.y * c.y;
                acc.z += c.z * c.z;
            }
        }
    const float inv = 1.0f / (float)(n * n);
    rgba[(size_t)py * out_w + px] = unorm8(sqrtf(acc.x * inv)) | (unorm8(sqrtf(acc.y * inv)) << 8) |
                                    (unorm8(sqrtf(acc.z * inv)) << 16) | (255u << 24);
}

Palette palette(uint32_t rule) {
    Palette p{};
    if (rule == AH_RULE_SMOOTHLIFE) {
        p.stops = 5;
        p.c[0] = make_float3(0.02f, 0.02f, 0.03f);
        p.c[1] = make_float3(0.3f, 0.07f, 0.25f);
        p.c[2] = make_float3(0.85f, 0.25f, 0.25f);
        p.c[3] = make_float3(0.98f, 0.7f, 0.3f);
        p.c[4] = make_float3(1.0f, 0.98f, 0.85f);
    } else if (rule == AH_RULE_MNCA) {
        p.stops = 2;
        p.c[0] = make_float3(0.02f, 0.02f, 0.07f);
        p.c[1] = make_float3(0.75f, 0.95f, 0.85f);
    } else {
        p.stops = 5;
        p.c[0] = make_float3(0.01f, 0.01f, 0.05f);
        p.c[1] = make_float3(0.08f, 0.12f, 0.45f);
        p.c[2] = make_float3(0.1f, 0.6f, 0.7f);
        p.c[3] = make_float3(0.95f, 0.85f, 0.4f);
        p.c[4] = make_float3(1.0f, 0.98f, 0.9f);
    }
    return p;
}

}  // namespace

void launch_render(hipStream_t s, Profiler* prof, const float* state, int W, int H, uint32_t rule, float center_x,
                   float center_y, float scale, int out_w, int out_h, uint32_t* rgba) {
    // Camera as Sim::render forms it (sim.rs:462-467).
    const float ox = center_x - (float)out_w * scale / 2.0f, oy = center_y - (float)out_h * scale / 2.0f;
    const int n = std::min(4, std::max(1, (int)std::ceil(scale)));
    Launch L(prof, s, "render");
    hipLaunchKernelGGL(k_render, dim3((out_w + 15) / 16, (out_h + 15) / 16), dim3(256), 0, s, state, W, H,
                       palette(rule), ox, oy, scale, n, out_w, out_h, rgba);
    AH_HIP(hipGetLastError());
}

}  // namespace ah
