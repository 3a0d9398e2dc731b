// Set-up kernels (FP64 where it matters) and the state statistics reduction.
//
//  - twiddle tables: exp(-2*pi*i*t/n) evaluated in FP64, rounded to FP32
//    (the reference forms cos/sin of an f32 angle per butterfly instead,
//    crates/auto-core/src/automata/spectral.rs:47-48);
//  - kernel taps: Chan's multi-ring shell of exp-bump cores, SmoothLife's
//    antialiased disc and ring, or an MNCA band, normalised to sum 1 in FP64
//    (the reference's shapes and box sum: spectral.rs:75-83,288-303,357-361,
//    432-440; one ring with beta={1} is exactly its Lenia kernel);
//  - kernel spectrum for the FFT path: the kernel is even in x and in y, so
//    its DFT is real and separable into two cosine sums, done in FP64 and
//    stored in FP32 already divided by W*H (which the reference applies after
//    the inverse transform instead, spectral.rs:215,227-234);
//  - seeds: the Orbium lattice (spectral.rs:183-197,313-335), SmoothLife's
//    discs (:381-391), MNCA's noise (:456-460) and the patch/blob inits, all
//    from the PCG hash of program.rs:140-147;
//  - statistics: FP64 mass, min, max and non-finite count per channel, with
//    wavefront shuffles and a fixed-order second pass (bitwise reproducible).
#include <mutex>

#include "ah_internal.h"

namespace ah {

namespace {

__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ uint32_t rand3(uint32_t a, uint32_t b, uint32_t c) { return pcg(pcg(pcg(a) ^ b) ^ c); }
__device__ __forceinline__ float unit_f(uint32_t r) { return (float)r * (1.0f / 4294967296.0f); }

__constant__ uint8_t c_orbium[400] = {
    0, 0, 0, 0, 0, 0, 0, 13, 4, 0, 0, 0, 0, 0, 0, 60, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 35, 53, 54, 21, 18, 25, 26, 18, 1, 65, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 22, 68, 91, 98, 87, 48, 47, 44, 47, 45, 27, 89, 0, 0, 0,
    0, 0, 0, 0, 3, 17, 89, 116, 119, 97, 33, 20, 14, 16, 31, 55, 180, 0, 0, 0,
    0, 0, 0, 9, 33, 47, 87, 103, 98, 72, 34, 0, 0, 0, 0, 12, 102, 84, 0, 0,
    1, 0, 4, 35, 43, 34, 28, 63, 69, 67, 53, 0, 0, 0, 0, 0, 5, 220, 0, 0,
    60, 0, 26, 44, 20, 0, 0, 51, 79, 94, 95, 61, 0, 0, 0, 0, 0, 116, 40, 0,
    0, 31, 47, 28, 0, 0, 0, 69, 109, 129, 137, 130, 0, 0, 0, 0, 0, 0, 132, 0,
    0, 150, 55, 8, 0, 0, 0, 48, 135, 162, 174, 174, 107, 0, 0, 0, 0, 0, 109, 0,
    0, 141, 60, 0, 0, 0, 0, 7, 158, 193, 214, 218, 206, 27, 0, 0, 0, 0, 72, 25,
    0, 0, 152, 0, 0, 0, 0, 0, 170, 223, 245, 253, 248, 143, 0, 0, 0, 0, 57, 36,
    0, 0, 214, 0, 0, 0, 0, 0, 129, 247, 255, 255, 255, 233, 71, 0, 0, 6, 56, 34,
    0, 0, 141, 19, 0, 0, 0, 0, 85, 255, 255, 250, 255, 248, 142, 40, 13, 30, 62, 22,
    0, 0, 8, 114, 0, 0, 0, 0, 45, 217, 255, 228, 220, 221, 166, 85, 50, 55, 59, 10,
    0, 0, 0, 108, 29, 0, 0, 0, 29, 158, 224, 210, 199, 189, 156, 104, 75, 68, 42, 0,
    0, 0, 0, 20, 85, 19, 0, 0, 30, 108, 172, 184, 173, 160, 134, 103, 79, 57, 16, 0,
    0, 0, 0, 0, 42, 66, 38, 30, 44, 86, 127, 142, 139, 127, 109, 86, 62, 30, 0, 0,
    0, 0, 0, 0, 0, 37, 59, 60, 66, 81, 99, 108, 105, 92, 78, 58, 32, 5, 0, 0,
    0, 0, 0, 0, 0, 0, 18, 43, 58, 64, 70, 71, 66, 59, 42, 24, 5, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 15, 26, 33, 34, 30, 20, 11, 0, 0, 0, 0, 0};

__global__ void k_cos_table64(double* tab, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) tab[t] = cos(2.0 * M_PI * (double)t / (double)n);
}

__global__ void k_twiddles(float2* tw, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        double sn, cs;
        sincos(2.0 * M_PI * (double)t / (double)n, &sn, &cs);
        tw[t] = make_float2((float)cs, (float)-sn);
    }
}

struct ShellArgs {
    double radius;
    int rings;
    double beta[AH_MAX_RINGS];
    int reach;
    int shape;      // ah_shape
    double inner;   // SMOOTH_RING, BAND
    double aa;      // SMOOTH_DISC, SMOOTH_RING
    float radius_f, inner_f;  // BAND compares a float32 distance, as the reference does
};

__device__ double shell64(double r, const ShellArgs& a) {
    if (r <= 0.0 || r >= a.radius) return 0.0;
    double b = (double)a.rings * r / a.radius;
    int k = (int)floor(b);
    if (k >= a.rings) return 0.0;
    double q = b - (double)k;
    if (q <= 0.0 || q >= 1.0) return 0.0;
    return a.beta[k] * exp(4.0 - 1.0 / (q * (1.0 - q)));
}

__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// Radial profile at offset (dx, dy).  The smooth SmoothLife shapes are
// spectral.rs:357-361 in FP64; the MNCA band keeps the reference's float32
// distance and comparisons (spectral.rs:432-440) so its support is the same
// set of cells.
__device__ double profile64(int dx, int dy, const ShellArgs& a) {
    const double r = sqrt((double)(dx * dx + dy * dy));
    switch (a.shape) {
        case AH_SHAPE_SMOOTH_DISC: return clamp01((a.radius + a.aa - r) / (2.0 * a.aa));
        case AH_SHAPE_SMOOTH_RING:
            return clamp01((a.radius + a.aa - r) / (2.0 * a.aa)) * clamp01((r - a.inner + a.aa) / (2.0 * a.aa));
        case AH_SHAPE_BAND: {
            const float rf = sqrtf((float)(dx * dx + dy * dy));
            return (rf >= a.inner_f && rf < a.radius_f) ? 1.0 : 0.0;
        }
        default: return shell64(r, a);
    }
}

// One workgroup: taps of the (2*reach+1)^2 box, FP64 sum (fixed order), normalise.
__global__ void __launch_bounds__(256) k_kernel_taps64(ShellArgs a, double* taps, double* sum_out) {
    __shared__ double part[256];
    const int side = 2 * a.reach + 1, n = side * side;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        int dy = i / side - a.reach, dx = i % side - a.reach;
        double v = profile64(dx, dy, a);
        taps[i] = v;
        acc += v;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    const double sum = part[0];
    for (int i = threadIdx.x; i < n; i += 256) taps[i] /= sum;
    if (threadIdx.x == 0) *sum_out = sum;
}

// T[dy][kx] = sum_dx K[dy][dx] cos(2 pi kx dx / W), dy in [0, 2*reach], kx in [0, W/2].
__global__ void k_khat_rows64(int W, int reach, const double* taps, const double* cosW, double* T) {
    const int nk = W / 2 + 1, side = 2 * reach + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk * side) return;
    int dyi = i / nk, kx = i - dyi * nk;
    const double* row = taps + (size_t)dyi * side + reach;
    double acc = row[0];
    for (int dx = 1; dx <= reach; ++dx) acc += 2.0 * row[dx] * cosW[(kx * dx) & (W - 1)];
    T[i] = acc;
}

// khat[kb][p][c] = (sum_dy T[dy][kx] cos(2 pi p dy / H)) / (W*H), p in [0, H/2];
// kx = 0 and kx = W/2 go to kpack[p] = ((K0+KN)/2, (K0-KN)/2).  `layout`
// KHAT_DIGITS stores khat[kx][digits(p)] instead (the block column kernels'
// digit order: H = 4096 and H = 8192).
__global__ void k_khat64(int W, int H, int reach, const double* T, const double* cosH, float* khat, float2* kpack,
                         int layout) {
    const int nk = W / 2 + 1, np = H / 2 + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk * np) return;
    int p = i / nk, kx = i - p * nk;
    double acc = T[(size_t)reach * nk + kx];
    for (int dy = 1; dy <= reach; ++dy)
        acc += (T[(size_t)(reach + dy) * nk + kx] + T[(size_t)(reach - dy) * nk + kx]) * cosH[(p * dy) & (H - 1)];
    const double scale = 1.0 / ((double)W * (double)H);
    acc *= scale;
    if (kx == 0 || kx == W / 2) {
        // kx = 0 owns the pair; recompute the partner bin so one thread writes kpack[p].
        if (kx != 0) return;
        double accN = T[(size_t)reach * nk + W / 2];
        for (int dy = 1; dy <= reach; ++dy)
            accN += (T[(size_t)(reach + dy) * nk + W / 2] + T[(size_t)(reach - dy) * nk + W / 2]) *
                    cosH[(p * dy) & (H - 1)];
        accN *= scale;
        kpack[p] = make_float2((float)(0.5 * (acc + accN)), (float)(0.5 * (acc - accN)));
        return;
    }
    if (layout == KHAT_DIGITS) {
        const int d = H == 8192 ? (p >> 9) * 512 + (p & 31) * 16 + ((p >> 5) & 15)
                                : (p >> 8) * 256 + (p & 15) * 16 + ((p >> 4) & 15);
        khat[(size_t)kx * np + d] = (float)acc;
        return;
    }
    int kb = kx >> 2, c = kx & 3;
    khat[((size_t)kb * np + p) * BC + c] = (float)acc;
}

struct SeedArgs {
    int W, H, C, init;
    uint32_t seed, size, count;
    const uint32_t* corners;  // BLOBS: (x, y) per blob
};

__global__ void k_seed(SeedArgs a, float* state) {
    const size_t cells = (size_t)a.W * a.H;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * a.C) return;
    int c = (int)(i / cells);
    size_t ci = i - (size_t)c * cells;
    uint32_t x = (uint32_t)(ci % a.W), y = (uint32_t)(ci / a.W);
    float v = 0.0f;
    if (a.init == AH_INIT_ORBIUM) {
        if (c == 0) {
            uint32_t r = rand3(x / 64u, y / 64u, a.seed);
            uint32_t px = x % 64u, py = y % 64u;
            if ((r & 1u) == 0u && px < 20u && py < 20u) {
                uint32_t ix = px, iy = py;
                switch ((r >> 8u) & 3u) {
                    case 1u: ix = py; iy = 19u - px; break;
                    case 2u: ix = 19u - px; iy = 19u - py; break;
                    case 3u: ix = 19u - py; iy = px; break;
                    default: break;
                }
                v = (float)c_orbium[iy * 20u + ix] / 255.0f;
            }
        }
    } else if (a.init == AH_INIT_PATCH) {
        uint32_t x0 = (a.W - a.size) / 2, y0 = (a.H - a.size) / 2;
        if (x >= x0 && x < x0 + a.size && y >= y0 && y < y0 + a.size) v = unit_f(rand3(x, y, a.seed + c));
    } else if (a.init == AH_INIT_DISCS) {
        // spectral.rs:381-391: a disc in every 96-cell tile whose hash is odd.
        const uint32_t tile = 96u;
        if (rand3(x / tile, y / tile, a.seed) & 1u) {
            float dx = (float)(x % tile) - (float)tile / 2.0f, dy = (float)(y % tile) - (float)tile / 2.0f;
            v = (dx * dx + dy * dy < 24.0f * 24.0f) ? 1.0f : 0.0f;
        }
    } else if (a.init == AH_INIT_NOISE) {
        // spectral.rs:456-460: 3x3 blocks alive with probability 0.32.
        v = unit_f(rand3(x / 3u, y / 3u, a.seed)) < 0.32f ? 1.0f : 0.0f;
    } else if (a.init == AH_INIT_BLOBS) {
        for (uint32_t b = 0; b < a.count; ++b) {
            uint32_t dx = (x + a.W - a.corners[2 * b]) % a.W, dy = (y + a.H - a.corners[2 * b + 1]) % a.H;
            if (dx < a.size && dy < a.size) {
                v = unit_f(rand3(x, y, a.seed + c));
                break;
            }
        }
    }
    state[i] = v;
}

// ---------------------------------------------------------------- stats
constexpr int STATS_THREADS = 256;

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ unsigned long long wave_sumu(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Grid (nblocks, channels): each block reduces a strided slice of one channel.
__global__ void __launch_bounds__(STATS_THREADS) k_stats_partial(const float* state, size_t cells, double* pm,
                                                                 float* pmin, float* pmax, unsigned long long* pnf) {
    const int c = blockIdx.y;
    const float4* s4 = reinterpret_cast<const float4*>(state + (size_t)c * cells);
    const size_t n4 = cells / 4;
    double m = 0.0;
    float lo = INFINITY, hi = -INFINITY;
    unsigned long long nf = 0;
    for (size_t i = (size_t)blockIdx.x * STATS_THREADS + threadIdx.x; i < n4; i += (size_t)gridDim.x * STATS_THREADS) {
        float4 v = s4[i];
        float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!isfinite(e[j])) {
                ++nf;
                continue;
            }
            m += (double)e[j];
            lo = fminf(lo, e[j]);
            hi = fmaxf(hi, e[j]);
        }
    }
    m = wave_sum(m);
    lo = wave_min(lo);
    hi = wave_max(hi);
    nf = wave_sumu(nf);
    __shared__ double sm[STATS_THREADS / 64];
    __shared__ float slo[STATS_THREADS / 64], shi[STATS_THREADS / 64];
    __shared__ unsigned long long snf[STATS_THREADS / 64];
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        sm[w] = m;
        slo[w] = lo;
        shi[w] = hi;
        snf[w] = nf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 1; j < STATS_THREADS / 64; ++j) {
            m += sm[j];
            lo = fminf(lo, slo[j]);
            hi = fmaxf(hi, shi[j]);
            nf += snf[j];
        }
        size_t o = (size_t)c * gridDim.x + blockIdx.x;
        pm[o] = m;
        pmin[o] = lo;
        pmax[o] = hi;
        pnf[o] = nf;
    }
}

// One block per channel folds the partials in a fixed order.
__global__ void __launch_bounds__(64) k_stats_final(int nblocks, double* pm, float* pmin, float* pmax,
                                                   unsigned long long* pnf) {
    const int c = blockIdx.x;
    double m = 0.0;
    float lo = INFINITY, hi = -INFINITY;
    unsigned long long nf = 0;
    for (int i = threadIdx.x; i < nblocks; i += 64) {
        size_t o = (size_t)c * nblocks + i;
        m += pm[o];
        lo = fminf(lo, pmin[o]);
        hi = fmaxf(hi, pmax[o]);
        nf += pnf[o];
    }
    m = wave_sum(m);
    lo = wave_min(lo);
    hi = wave_max(hi);
    nf = wave_sumu(nf);
    if (threadIdx.x == 0) {
        // Results land in slot 0 of each channel's partial range.
        size_t o = (size_t)c * nblocks;
        pm[o] = m;
        pmin[o] = lo;
        pmax[o] = hi;
        pnf[o] = nf;
    }
}

}  // namespace

void launch_cos_table64(hipStream_t s, Profiler* p, double* tab, int n) {
    Launch L(p, s, "cos_table64");
    hipLaunchKernelGGL(k_cos_table64, dim3((n + 255) / 256), dim3(256), 0, s, tab, n);
    AH_HIP(hipGetLastError());
}

void launch_twiddles(hipStream_t s, Profiler* p, float2* tw, int n) {
    Launch L(p, s, "twiddles64");
    hipLaunchKernelGGL(k_twiddles, dim3((n + 255) / 256), dim3(256), 0, s, tw, n);
    AH_HIP(hipGetLastError());
}

void launch_kernel_taps64(hipStream_t s, Profiler* p, const ah_kernel& k, int reach, double* taps, double* sum) {
    ShellArgs a;
    a.radius = (double)k.radius;
    a.rings = (int)k.rings;
    for (int i = 0; i < AH_MAX_RINGS; ++i) a.beta[i] = (double)k.beta[i];
    a.reach = reach;
    a.shape = (int)k.shape;
    a.inner = (double)k.inner;
    a.aa = (double)k.aa;
    a.radius_f = k.radius;
    a.inner_f = k.inner;
    Launch L(p, s, "kernel_taps64");
    hipLaunchKernelGGL(k_kernel_taps64, dim3(1), dim3(256), 0, s, a, taps, sum);
    AH_HIP(hipGetLastError());
}

void launch_khat64(hipStream_t s, Profiler* p, int W, int H, int reach, const double* taps, const double* cosW,
                   const double* cosH, double* T, float* khat, float2* kpack, int layout) {
    const int side = 2 * reach + 1, nk = W / 2 + 1, np = H / 2 + 1;
    {
        Launch L(p, s, "khat_rows64");
        int n = nk * side;
        hipLaunchKernelGGL(k_khat_rows64, dim3((n + 255) / 256), dim3(256), 0, s, W, reach, taps, cosW, T);
        AH_HIP(hipGetLastError());
    }
    {
        Launch L(p, s, "khat64");
        int n = nk * np;
        hipLaunchKernelGGL(k_khat64, dim3((n + 255) / 256), dim3(256), 0, s, W, H, reach, T, cosH, khat, kpack,
                           layout);
        AH_HIP(hipGetLastError());
    }
}

void launch_seed(hipStream_t s, Profiler* p, const ah_lenia_desc& d, uint32_t seed, const uint32_t* corners,
                 float* state) {
    SeedArgs a;
    a.W = (int)d.width;
    a.H = (int)d.height;
    a.C = (int)d.channels;
    a.init = (int)d.init;
    a.seed = seed;
    a.size = d.init_size;
    a.count = d.init_count;
    a.corners = corners;
    size_t n = (size_t)d.width * d.height * d.channels;
    Launch L(p, s, "seed");
    hipLaunchKernelGGL(k_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, state);
    AH_HIP(hipGetLastError());
}

int stats_blocks(size_t cells) {
    size_t b = (cells / 4 + STATS_THREADS * 16 - 1) / (STATS_THREADS * 16);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

void launch_stats(hipStream_t s, Profiler* p, const float* state, size_t cells, int channels, double* pm,
                  float* pmin, float* pmax, unsigned long long* pnf, int nblocks) {
    {
        Launch L(p, s, "stats_partial");
        hipLaunchKernelGGL(k_stats_partial, dim3(nblocks, channels), dim3(STATS_THREADS), 0, s, state, cells, pm, pmin,
                           pmax, pnf);
        AH_HIP(hipGetLastError());
    }
    {
        Launch L(p, s, "stats_final");
        hipLaunchKernelGGL(k_stats_final, dim3(channels), dim3(64), 0, s, nblocks, pm, pmin, pmax, pnf);
        AH_HIP(hipGetLastError());
    }
}

}  // namespace ah

namespace ah {
namespace {
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace

void launch_copy_probe(hipStream_t s, const float4* src, float4* dst, size_t n4) {
    hipLaunchKernelGGL(k_copy_probe, dim3(8192), dim3(256), 0, s, src, dst, n4);
    AH_HIP(hipGetLastError());
}
}  // namespace ah

// ------------------------------------------------------------ stats without allocation
// ah_state_stats_noalloc: the reduction of ah_sim_stats over a state the
// caller holds in device memory, for callers that cannot allocate.  Its
// partial results live in `scratch`, a fixed device array (one per GPU,
// held until the process ends): per channel NOALLOC_BLOCKS FP64 masses,
// float minima and maxima and u64 non-finite counts.  Every partial the
// final pass reads is written by the first pass of the same call, so
// nothing needs clearing between calls; calls that overlap on one GPU share
// the array and race.
namespace ah {
namespace {

constexpr int NOALLOC_BLOCKS = 1024;  // = the largest block count stats_blocks() picks
constexpr size_t NOALLOC_SLOTS = (size_t)AH_MAX_CHANNELS * NOALLOC_BLOCKS;
constexpr size_t NOALLOC_BYTES = NOALLOC_SLOTS * (sizeof(double) + 2 * sizeof(float) + sizeof(unsigned long long));

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Device address of this GPU's `scratch`, looked up once per device with
// that device current.  The array's name in host code is only a stand-in:
// kernels and copies must use this address.
unsigned char* scratch_address() {
    static std::mutex mu;
    static unsigned char* addr[64] = {};
    int dev = 0;
    AH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error{AH_ERR_DEVICE, "device index out of range"};
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        AH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error{AH_ERR_DEVICE, "scratch array has no device address"};
        addr[dev] = static_cast<unsigned char*>(p);
    }
    return addr[dev];
}

}  // namespace
}  // namespace ah

extern "C" ah_status ah_state_stats_noalloc(const float* state, uint64_t cells, uint32_t channels, void* stream,
                                            ah_stats* out) {
    using namespace ah;
    return guard([&] {
        if (!state || !out) throw Error{AH_ERR_INVALID, "null argument"};
        if (channels < 1 || channels > AH_MAX_CHANNELS)
            throw Error{AH_ERR_INVALID, "channels must be 1..3 (the scratch array holds partials for at most 3)"};
        if (cells < 4 || cells % 4) throw Error{AH_ERR_INVALID, "cells per channel must be a positive multiple of 4"};
        if (reinterpret_cast<uintptr_t>(state) % 16) throw Error{AH_ERR_INVALID, "state must be 16-byte aligned"};
        const int nb = stats_blocks((size_t)cells);
        if ((size_t)nb * channels > NOALLOC_SLOTS) throw Error{AH_ERR_INVALID, "partials exceed the scratch array"};
        hipStream_t st = static_cast<hipStream_t>(stream);
        unsigned char* base = scratch_address();
        double* pm = reinterpret_cast<double*>(base);
        unsigned long long* pnf = reinterpret_cast<unsigned long long*>(base + NOALLOC_SLOTS * sizeof(double));
        float* pmin = reinterpret_cast<float*>(base + NOALLOC_SLOTS * (sizeof(double) + sizeof(unsigned long long)));
        float* pmax = pmin + NOALLOC_SLOTS;
        launch_stats(st, nullptr, state, (size_t)cells, (int)channels, pm, pmin, pmax, pnf, nb);
        ah_stats r{};
        double mass[AH_MAX_CHANNELS];
        float lo[AH_MAX_CHANNELS], hi[AH_MAX_CHANNELS];
        unsigned long long nf[AH_MAX_CHANNELS];
        for (uint32_t c = 0; c < channels; ++c) {
            const size_t o = (size_t)c * nb;  // the final pass leaves each channel's result in slot 0
            AH_HIP(hipMemcpyAsync(&mass[c], pm + o, sizeof(double), hipMemcpyDeviceToHost, st));
            AH_HIP(hipMemcpyAsync(&lo[c], pmin + o, sizeof(float), hipMemcpyDeviceToHost, st));
            AH_HIP(hipMemcpyAsync(&hi[c], pmax + o, sizeof(float), hipMemcpyDeviceToHost, st));
            AH_HIP(hipMemcpyAsync(&nf[c], pnf + o, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        }
        AH_HIP(hipStreamSynchronize(st));
        for (uint32_t c = 0; c < channels; ++c) {
            r.mass[c] = mass[c];
            r.min[c] = lo[c];
            r.max[c] = hi[c];
            r.nonfinite += nf[c];
        }
        *out = r;
    });
}
