/*
 * lenia_oracle.c -- CPU oracle for the Lenia world step.  TEST INFRASTRUCTURE.
 *
 * Test infrastructure, never the product: only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load it.  The HIP engine in auto_amd/
 * does not link, call or fall back to anything here.
 *
 * Contents (each function cites the reference code it restates):
 *   1. Integer generators: PCG hash, rand3, unit_f (program.rs:140-147) and
 *      SplitMix64 (crates/auto-life-gpu/src/world.rs:153-159).
 *   2. Seeds: the Orbium lattice (spectral.rs:183-197,287,313-335),
 *      SmoothLife's discs (:381-391), MNCA's noise (:456-460) and the
 *      patch / blob inits of the benchmark workloads.
 *   3. Kernels: the Lenia core (spectral.rs:288-303) generalised to Chan's
 *      multi-ring shell, SmoothLife's disc and ring (:357-361) and MNCA's
 *      bands (:432-440), normalised in FP64; the reference's own FP32 kernel
 *      pass and box sum (spectral.rs:75-83,150-171,296).
 *   4. The reference's FP32 step, pass by pass: radix-2 Stockham stages with
 *      per-butterfly FP32 twiddles (spectral.rs:14-64), load / fft x / fft y /
 *      multiply / ifft y / ifft x / scale / update (spectral.rs:199-264,
 *      337-345), with the SmoothLife and MNCA transitions (:393-405,462-473)
 *      in f32.  This is the "auto-cpu" Lenia step the reference lacks
 *      (crates/auto-cpu/src/lib.rs:1-5) and the CPU baseline.
 *   5. FP64 direct periodic convolution and growth/update, as the reference's
 *      known-answer test computes it (crates/auto-core/tests/automata.rs:317-351),
 *      generalised to any kernel, several channels and mixing weights, and the
 *      SmoothLife / MNCA transitions in FP64.
 *   6. FP64 statistics (mass, min, max, non-finite count).
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - Stockham FFT: brute-force FP64 DFT at small sizes and numpy.fft.
 *   - FP32 step: the reference's own tolerance, max-abs < 0.002 after one
 *     step at 128x64, R=13, Orbium seed, vs the FP64 direct convolution
 *     (automata.rs:300-362); golden fixtures under tests/golden/.
 *   - FP64 direct convolution: brute force agrees with numpy's FP64 FFT
 *     convolution (two independent methods), and a uniform field convolves to
 *     itself (kernel sums to 1).
 *   - Kernel: closed-form peaks (core(1/2)=1; ring k peaks at beta_k),
 *     support r<R, 8-fold symmetry, FP64 normalisation; SmoothLife shapes:
 *     the box sum tends to the disc / annulus area; MNCA band: lattice-point
 *     counts by brute force.
 *   - SmoothLife / MNCA transitions: closed-form cases (birth interval,
 *     survival interval, band overrides) and f32 vs FP64 agreement.
 *   - PCG/SplitMix64: SplitMix64's published first outputs for seed 0;
 *     PCG: bijectivity on samples and the Orbium KAT (a seed in 1..63 places
 *     a creature at 128x64).  PCG's bit values vs the reference's WGSL are
 *     parity unpinned beyond the formula (no stored vectors exist).
 *
 * Build: oracle/Makefile -> oracle/liblenia_oracle.so (gcc, OpenMP).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/auto_hip.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* 1. generators                                                       */
/* ------------------------------------------------------------------ */

/* PCG hash (Jarzynski & Olano 2020), program.rs:140-144. */
uint32_t or_pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

/* program.rs:145 */
uint32_t or_rand3(uint32_t a, uint32_t b, uint32_t c) { return or_pcg(or_pcg(or_pcg(a) ^ b) ^ c); }

/* program.rs:146: f32(r) * (1/2^32), the product in f32. */
float or_unit_f(uint32_t r) { return (float)r * (1.0f / 4294967296.0f); }

/* SplitMix64, crates/auto-life-gpu/src/world.rs:153-159. */
uint64_t or_splitmix64(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* ------------------------------------------------------------------ */
/* 2. seeds                                                            */
/* ------------------------------------------------------------------ */

/* Orbium unicaudatus, 20x20, scaled to 0..255 (spectral.rs:287). */
static const uint8_t ORBIUM[400] = {
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

/* seed() of spectral.rs:313-335: one Orbium per occupied 64-cell slot. */
float or_seed_orbium(uint32_t x, uint32_t y, uint32_t seed) {
    const uint32_t SPACING = 64u;
    uint32_t r = or_rand3(x / SPACING, y / SPACING, seed);
    if ((r & 1u) != 0u) return 0.0f;
    uint32_t px = x % SPACING, py = y % SPACING;
    if (px >= 20u || py >= 20u) return 0.0f;
    uint32_t ix = px, iy = py;
    switch ((r >> 8u) & 3u) {
        case 1u: ix = py; iy = 19u - px; break;
        case 2u: ix = 19u - px; iy = 19u - py; break;
        case 3u: ix = 19u - py; iy = px; break;
        default: break;
    }
    return (float)ORBIUM[iy * 20u + ix] / 255.0f;
}

/*
 * Fill `out` ([channels][h][w]) with the init the descriptor names.
 * PATCH: centred init_size^2 square, channel c seeded with seed+c.
 * BLOBS: init_count squares of init_size^2; corner i = (next % w, next % h)
 *        from one SplitMix64 stream started at `seed`; wrapped on the torus.
 * Cell values depend only on (x, y, seed), so overlaps are order-free.
 */
int or_init_state(const ah_lenia_desc* d, uint32_t seed, float* out) {
    const uint32_t w = d->width, h = d->height, C = d->channels;
    const size_t cells = (size_t)w * h;
    memset(out, 0, sizeof(float) * cells * C);
    if (d->init == AH_INIT_ORBIUM) {
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) out[(size_t)y * w + x] = or_seed_orbium(x, y, seed);
    } else if (d->init == AH_INIT_PATCH) {
        uint32_t p = d->init_size;
        if (p > w || p > h) return -1;
        uint32_t x0 = (w - p) / 2, y0 = (h - p) / 2;
        for (uint32_t c = 0; c < C; ++c)
            for (uint32_t y = y0; y < y0 + p; ++y)
                for (uint32_t x = x0; x < x0 + p; ++x)
                    out[c * cells + (size_t)y * w + x] = or_unit_f(or_rand3(x, y, seed + c));
    } else if (d->init == AH_INIT_BLOBS) {
        uint64_t st = seed;
        uint32_t b = d->init_size;
        for (uint32_t i = 0; i < d->init_count; ++i) {
            uint32_t cx = (uint32_t)(or_splitmix64(&st) % w);
            uint32_t cy = (uint32_t)(or_splitmix64(&st) % h);
            for (uint32_t c = 0; c < C; ++c)
                for (uint32_t j = 0; j < b; ++j)
                    for (uint32_t k = 0; k < b; ++k) {
                        uint32_t x = (cx + k) % w, y = (cy + j) % h;
                        out[c * cells + (size_t)y * w + x] = or_unit_f(or_rand3(x, y, seed + c));
                    }
        }
    } else if (d->init == AH_INIT_DISCS) {
        /* SmoothLife seed(), spectral.rs:381-391 (every channel alike). */
        const uint32_t tile = 96u;
        for (uint32_t c = 0; c < C; ++c)
            for (uint32_t y = 0; y < h; ++y)
                for (uint32_t x = 0; x < w; ++x) {
                    if ((or_rand3(x / tile, y / tile, seed) & 1u) == 0u) continue;
                    float dx = (float)(x % tile) - (float)tile / 2.0f, dy = (float)(y % tile) - (float)tile / 2.0f;
                    out[c * cells + (size_t)y * w + x] = (dx * dx + dy * dy < 24.0f * 24.0f) ? 1.0f : 0.0f;
                }
    } else if (d->init == AH_INIT_NOISE) {
        /* MNCA seed(), spectral.rs:456-460. */
        for (uint32_t c = 0; c < C; ++c)
            for (uint32_t y = 0; y < h; ++y)
                for (uint32_t x = 0; x < w; ++x)
                    out[c * cells + (size_t)y * w + x] = or_unit_f(or_rand3(x / 3u, y / 3u, seed)) < 0.32f ? 1.0f : 0.0f;
    } else if (d->init != AH_INIT_ZERO) {
        return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* 3. kernels                                                          */
/* ------------------------------------------------------------------ */

/* Lenia core, spectral.rs:290-297: exp(4 - 1/(q(1-q))) on (0,1), else 0. */
double or_core64(double q) {
    if (q <= 0.0 || q >= 1.0) return 0.0;
    return exp(4.0 - 1.0 / (q * (1.0 - q)));
}
float or_core32(float q) {
    if (q <= 0.0f || q >= 1.0f) return 0.0f;
    return expf(4.0f - 1.0f / (q * (1.0f - q)));
}

/*
 * Chan's multi-ring shell at distance r: with b = rings*r/R, ring k=floor(b)
 * gets beta[k]*core(b-k); zero for r >= R.  rings=1, beta={1} is the
 * reference shape core(r/R) (spectral.rs:288-303).
 */
double or_shell64(double r, double R, uint32_t rings, const float* beta) {
    if (r >= R || r <= 0.0) return 0.0;
    double b = (double)rings * r / R;
    uint32_t k = (uint32_t)floor(b);
    if (k >= rings) return 0.0;
    return (double)beta[k] * or_core64(b - (double)k);
}

/* Half-width of the box that holds every tap of a Lenia shell: ceil(R). */
int or_reach(float R) { return (int)ceil((double)R); }

/* The same for any shape: the smooth SmoothLife shapes reach radius + aa. */
int or_kernel_reach(const ah_kernel* k) {
    if (k->shape == AH_SHAPE_SMOOTH_DISC || k->shape == AH_SHAPE_SMOOTH_RING)
        return (int)ceil((double)k->radius + (double)k->aa);
    return or_reach(k->radius);
}

static double clamp01d(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

/*
 * Radial profile of any kernel shape at offset (dx, dy).  SmoothLife's inner
 * disc and outer ring, spectral.rs:357-361, in FP64; the MNCA band of
 * spectral.rs:432-440 with the reference's float32 distance and comparisons.
 */
double or_profile64(const ah_kernel* k, int dx, int dy) {
    const double r = sqrt((double)(dx * dx + dy * dy));
    const double R = k->radius, ri = k->inner, aa = k->aa;
    switch (k->shape) {
        case AH_SHAPE_SMOOTH_DISC: return clamp01d((R + aa - r) / (2.0 * aa));
        case AH_SHAPE_SMOOTH_RING: return clamp01d((R + aa - r) / (2.0 * aa)) * clamp01d((r - ri + aa) / (2.0 * aa));
        case AH_SHAPE_BAND: {
            float rf = sqrtf((float)(dx * dx + dy * dy));
            return (rf >= k->inner && rf < k->radius) ? 1.0 : 0.0;
        }
        default: return or_shell64(r, R, k->rings, k->beta);
    }
}

/*
 * FP64 kernel taps in a (2*reach+1)^2 box, normalised to sum 1.
 * taps[(dy+reach)*(2*reach+1) + (dx+reach)].  Returns the raw sum.
 */
double or_kernel_box64(const ah_kernel* k, double* taps) {
    int reach = or_kernel_reach(k), side = 2 * reach + 1;
    double sum = 0.0;
    for (int dy = -reach; dy <= reach; ++dy)
        for (int dx = -reach; dx <= reach; ++dx) {
            double v = or_profile64(k, dx, dy);
            taps[(dy + reach) * side + (dx + reach)] = v;
            sum += v;
        }
    for (int i = 0; i < side * side; ++i) taps[i] /= sum;
    return sum;
}

/* kernel_sum of spectral.rs:75-83 with reach ceil(R)+1 (:296): f32 shape, f64 sum, f32 result. */
float or_kernel_sum_ref32(float R) {
    int reach = (int)ceilf(R) + 1;
    double sum = 0.0;
    for (int y = -reach; y <= reach; ++y)
        for (int x = -reach; x <= reach; ++x) sum += (double)or_core32(sqrtf((float)(x * x + y * y)) / R);
    return (float)sum;
}

/* The reference's "kernel" init pass (spectral.rs:150-171), as interleaved complex f32. */
void or_kernel_ref32(uint32_t w, uint32_t h, float R, float sum, float* work) {
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float dx = (float)(x < w - x ? x : w - x);
            float dy = (float)(y < h - y ? y : h - y);
            float v = or_core32(sqrtf(dx * dx + dy * dy) / R) / sum;
            size_t i = (size_t)y * w + x;
            work[2 * i] = v;
            work[2 * i + 1] = 0.0f;
        }
}

/* ------------------------------------------------------------------ */
/* 4. the reference FP32 step                                          */
/* ------------------------------------------------------------------ */

/*
 * All log2(n) stages of fft_pass (spectral.rs:14-64) along one axis, each
 * stage a full ping-pong as the reference's repeated dispatch does.
 * Stage `phase`: m = 2^phase, k = p & (m-1), a = cur[p], b = cur[p+n/2],
 * t = b * (cos, sin)(sign*pi*k/m), next[2p-k] = a+t, next[2p-k+m] = a-t.
 * The angle is formed in f32 exactly as the WGSL: ((sign*pi)*f32(k))/f32(m)
 * with pi = 3.14159265358979 rounded to f32.  Result ends in `buf`.
 */
void or_fft_axis_ref32(uint32_t w, uint32_t h, float* buf, float* tmp, int axis_x, int inverse) {
    const uint32_t n = axis_x ? w : h, rows = axis_x ? h : w, half = n / 2;
    const float sign = inverse ? 1.0f : -1.0f;
    const float PI32 = (float)3.14159265358979;
    uint32_t log2n = 0;
    while ((1u << log2n) < n) ++log2n;
    float* cur = buf;
    float* nxt = tmp;
    for (uint32_t phase = 0; phase < log2n; ++phase) {
        const uint32_t m = 1u << phase;
#pragma omp parallel for schedule(static)
        for (uint32_t r = 0; r < rows; ++r) {
            for (uint32_t p = 0; p < half; ++p) {
                uint32_t k = p & (m - 1u);
                size_t i0, i1, o0, o1;
                if (axis_x) {
                    i0 = (size_t)r * w + p; i1 = (size_t)r * w + p + half;
                    o0 = (size_t)r * w + (2 * p - k); o1 = (size_t)r * w + (2 * p - k + m);
                } else {
                    i0 = (size_t)p * w + r; i1 = (size_t)(p + half) * w + r;
                    o0 = (size_t)(2 * p - k) * w + r; o1 = (size_t)(2 * p - k + m) * w + r;
                }
                float ax = cur[2 * i0], ay = cur[2 * i0 + 1];
                float bx = cur[2 * i1], by = cur[2 * i1 + 1];
                float angle = sign * PI32 * (float)k / (float)m;
                float twx = cosf(angle), twy = sinf(angle);
                float tx = bx * twx - by * twy;
                float ty = bx * twy + by * twx;
                nxt[2 * o0] = ax + tx; nxt[2 * o0 + 1] = ay + ty;
                nxt[2 * o1] = ax - tx; nxt[2 * o1 + 1] = ay - ty;
            }
        }
        float* t = cur; cur = nxt; nxt = t;
    }
    if (cur != buf) memcpy(buf, cur, sizeof(float) * 2 * (size_t)w * h);
}

/* Kernel spectrum as the reference init builds it: kernel pass, fft x, fft y (spectral.rs:148-181). */
void or_kernel_spectrum_ref32(uint32_t w, uint32_t h, float R, float* kern) {
    float* tmp = (float*)malloc(sizeof(float) * 2 * (size_t)w * h);
    or_kernel_ref32(w, h, R, or_kernel_sum_ref32(R), kern);
    or_fft_axis_ref32(w, h, kern, tmp, 1, 0);
    or_fft_axis_ref32(w, h, kern, tmp, 0, 0);
    free(tmp);
}

/*
 * Kernel spectrum for any ah_kernel, through the same FP32 passes: the FP64
 * normalised shell rounded to f32 on the torus (min-image), then fft x, fft y.
 */
void or_kernel_spectrum32(uint32_t w, uint32_t h, const ah_kernel* kd, float* kern) {
    int reach = or_kernel_reach(kd), side = 2 * reach + 1;
    double* taps = (double*)malloc(sizeof(double) * side * side);
    or_kernel_box64(kd, taps);
    memset(kern, 0, sizeof(float) * 2 * (size_t)w * h);
    for (int dy = -reach; dy <= reach; ++dy)
        for (int dx = -reach; dx <= reach; ++dx) {
            uint32_t x = (uint32_t)((dx + (int)w) % (int)w), y = (uint32_t)((dy + (int)h) % (int)h);
            kern[2 * ((size_t)y * w + x)] += (float)taps[(dy + reach) * side + (dx + reach)];
        }
    free(taps);
    float* tmp = (float*)malloc(sizeof(float) * 2 * (size_t)w * h);
    or_fft_axis_ref32(w, h, kern, tmp, 1, 0);
    or_fft_axis_ref32(w, h, kern, tmp, 0, 0);
    free(tmp);
}

/*
 * Convolution of one f32 channel with a kernel spectrum through the
 * reference passes: load, fft x, fft y, multiply, ifft y, ifft x, scale
 * (spectral.rs:199-234).  `u` gets re(work) * f32(1/cells).
 */
void or_conv_ref32(uint32_t w, uint32_t h, const float* kern, const float* s, float* u) {
    const size_t cells = (size_t)w * h;
    float* work = (float*)malloc(sizeof(float) * 2 * cells);
    float* tmp = (float*)malloc(sizeof(float) * 2 * cells);
    for (size_t i = 0; i < cells; ++i) { work[2 * i] = s[i]; work[2 * i + 1] = 0.0f; }
    or_fft_axis_ref32(w, h, work, tmp, 1, 0);
    or_fft_axis_ref32(w, h, work, tmp, 0, 0);
    for (size_t i = 0; i < cells; ++i) {
        float ax = work[2 * i], ay = work[2 * i + 1], bx = kern[2 * i], by = kern[2 * i + 1];
        work[2 * i] = ax * bx - ay * by;
        work[2 * i + 1] = ax * by + ay * bx;
    }
    or_fft_axis_ref32(w, h, work, tmp, 0, 1);
    or_fft_axis_ref32(w, h, work, tmp, 1, 1);
    const float scale = 1.0f / (float)cells;
    for (size_t i = 0; i < cells; ++i) u[i] = work[2 * i] * scale;
    free(work);
    free(tmp);
}

/* Growth and update of spectral.rs:337-345 in f32. */
static inline float growth32(float u, float mu, float sigma) {
    float d = (u - mu) / sigma;
    return 2.0f * expf(-0.5f * d * d) - 1.0f;
}
static inline float clamp01f(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

/*
 * One FP32 step for any descriptor, through the reference passes: per kernel
 * k, u_k = conv(s[src_k], K_k); then for each channel c,
 * s'_c = clamp(s_c + dt * sum_{k: dst_k=c} h_k * g_k(u_k), 0, 1) with h_k
 * normalised per destination.  For one kernel with h=1 this is exactly the
 * reference update (spectral.rs:236-264,337-345).  `kerns` holds n_kernels
 * spectra of w*h complex f32 (or_kernel_spectrum32).
 */
/* SmoothLife transition, spectral.rs:393-405, in f32 as the WGSL does. */
static float sig32(float x, float a, float alpha) { return 1.0f / (1.0f + expf((a - x) * 4.0f / alpha)); }
float or_smoothlife32(float s, float c0, float c1, float dt) {
    const float B1 = 0.278f, B2 = 0.365f, D1 = 0.267f, D2 = 0.445f, AN = 0.028f, AM = 0.147f;
    float m = sig32(c0, 0.5f, AM);
    float lo = B1 * (1.0f - m) + D1 * m, hi = B2 * (1.0f - m) + D2 * m;
    float target = sig32(c1, lo, AN) * (1.0f - sig32(c1, hi, AN));
    float v = s + dt * (2.0f * target - 1.0f);
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

/* MNCA bands, spectral.rs:462-473: later bands overrule earlier ones. */
static int band32(float v, float lo, float hi) { return v >= lo && v <= hi; }
float or_mnca32(float s, float c0, float c1) {
    float o = s;
    if (band32(c0, 0.210f, 0.220f)) o = 1.0f;
    if (band32(c0, 0.350f, 0.500f)) o = 0.0f;
    if (band32(c0, 0.750f, 0.850f)) o = 0.0f;
    if (band32(c0, 0.120f, 0.150f)) o = 0.0f;
    if (band32(c1, 0.100f, 0.280f)) o = 0.0f;
    if (band32(c1, 0.430f, 0.550f)) o = 1.0f;
    return o;
}

void or_step_ref32(const ah_lenia_desc* d, const float* kerns, const float* s, float* out) {
    const uint32_t w = d->width, h = d->height, C = d->channels, K = d->n_kernels;
    const size_t cells = (size_t)w * h;
    if (d->rule != AH_RULE_LENIA) {
        /* SmoothLife / MNCA: one channel, u0 and u1 through the FP32 passes. */
        float* u = (float*)malloc(sizeof(float) * cells * 2);
        for (uint32_t k = 0; k < 2; ++k) or_conv_ref32(w, h, kerns + 2 * cells * k, s, u + cells * k);
        for (size_t i = 0; i < cells; ++i)
            out[i] = d->rule == AH_RULE_SMOOTHLIFE ? or_smoothlife32(s[i], u[i], u[cells + i], d->dt)
                                                   : or_mnca32(s[i], u[i], u[cells + i]);
        free(u);
        return;
    }
    float hsum[AH_MAX_CHANNELS] = {0};
    for (uint32_t k = 0; k < K; ++k) hsum[d->kernels[k].dst] += d->kernels[k].weight;
    float* acc = (float*)calloc(cells * C, sizeof(float));
    float* u = (float*)malloc(sizeof(float) * cells);
    for (uint32_t k = 0; k < K; ++k) {
        const ah_kernel* kd = &d->kernels[k];
        or_conv_ref32(w, h, kerns + 2 * cells * k, s + cells * kd->src, u);
        float hk = kd->weight / hsum[kd->dst];
        float* a = acc + cells * kd->dst;
        for (size_t i = 0; i < cells; ++i) a[i] += hk * growth32(u[i], kd->mu, kd->sigma);
    }
    for (size_t i = 0; i < cells * C; ++i) out[i] = clamp01f(s[i] + d->dt * acc[i]);
    free(acc);
    free(u);
}

/* ------------------------------------------------------------------ */
/* 5. FP64 direct periodic convolution (automata.rs:317-351)            */
/* ------------------------------------------------------------------ */

/* u[y][x] = sum over the box of K(dx,dy) * s[(y+dy) mod h][(x+dx) mod w], FP64. */
void or_conv_direct64(uint32_t w, uint32_t h, const ah_kernel* kd, const float* s, double* u) {
    int reach = or_kernel_reach(kd), side = 2 * reach + 1;
    double* taps = (double*)malloc(sizeof(double) * side * side);
    or_kernel_box64(kd, taps);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < (int)h; ++y)
        for (int x = 0; x < (int)w; ++x) {
            double acc = 0.0;
            for (int dy = -reach; dy <= reach; ++dy) {
                const float* row = s + (size_t)(((y + dy) % (int)h + (int)h) % (int)h) * w;
                const double* trow = taps + (dy + reach) * side + reach;
                for (int dx = -reach; dx <= reach; ++dx) {
                    double t = trow[dx];
                    if (t != 0.0) acc += t * (double)row[((x + dx) % (int)w + (int)w) % (int)w];
                }
            }
            u[(size_t)y * w + x] = acc;
        }
    free(taps);
}

/* SmoothLife transition in FP64 (constants as the WGSL's f32 literals). */
static double sig64(double x, double a, double alpha) { return 1.0 / (1.0 + exp((a - x) * 4.0 / alpha)); }
double or_smoothlife64(double s, double c0, double c1, double dt) {
    const double B1 = 0.278f, B2 = 0.365f, D1 = 0.267f, D2 = 0.445f, AN = 0.028f, AM = 0.147f;
    double m = sig64(c0, 0.5, AM);
    double lo = B1 * (1.0 - m) + D1 * m, hi = B2 * (1.0 - m) + D2 * m;
    double target = sig64(c1, lo, AN) * (1.0 - sig64(c1, hi, AN));
    return clamp01d(s + dt * (2.0 * target - 1.0));
}

/* MNCA bands in FP64 against the WGSL's f32 thresholds. */
static int band64(double v, float lo, float hi) { return v >= (double)lo && v <= (double)hi; }
double or_mnca64(double s, double c0, double c1) {
    double o = s;
    if (band64(c0, 0.210f, 0.220f)) o = 1.0;
    if (band64(c0, 0.350f, 0.500f)) o = 0.0;
    if (band64(c0, 0.750f, 0.850f)) o = 0.0;
    if (band64(c0, 0.120f, 0.150f)) o = 0.0;
    if (band64(c1, 0.100f, 0.280f)) o = 0.0;
    if (band64(c1, 0.430f, 0.550f)) o = 1.0;
    return o;
}

/* Growth/update in FP64 from given potentials, rounding the result to f32. */
void or_update64(const ah_lenia_desc* d, const double* u, const float* s, float* out) {
    const size_t cells = (size_t)d->width * d->height;
    const uint32_t C = d->channels, K = d->n_kernels;
    if (d->rule != AH_RULE_LENIA) {
        for (size_t i = 0; i < cells; ++i)
            out[i] = (float)(d->rule == AH_RULE_SMOOTHLIFE ? or_smoothlife64(s[i], u[i], u[cells + i], d->dt)
                                                           : or_mnca64(s[i], u[i], u[cells + i]));
        return;
    }
    double hsum[AH_MAX_CHANNELS] = {0};
    for (uint32_t k = 0; k < K; ++k) hsum[d->kernels[k].dst] += d->kernels[k].weight;
    for (uint32_t c = 0; c < C; ++c)
        for (size_t i = 0; i < cells; ++i) {
            double acc = 0.0;
            for (uint32_t k = 0; k < K; ++k) {
                const ah_kernel* kd = &d->kernels[k];
                if (kd->dst != c) continue;
                double dd = (u[cells * k + i] - (double)kd->mu) / (double)kd->sigma;
                acc += ((double)kd->weight / hsum[c]) * (2.0 * exp(-0.5 * dd * dd) - 1.0);
            }
            double v = (double)s[cells * c + i] + (double)d->dt * acc;
            out[cells * c + i] = (float)(v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v));
        }
}

/* One FP64 direct step for any descriptor (the reference KAT generalised). */
void or_step_direct64(const ah_lenia_desc* d, const float* s, float* out) {
    const size_t cells = (size_t)d->width * d->height;
    double* u = (double*)malloc(sizeof(double) * cells * d->n_kernels);
    for (uint32_t k = 0; k < d->n_kernels; ++k)
        or_conv_direct64(d->width, d->height, &d->kernels[k], s + cells * d->kernels[k].src, u + cells * k);
    or_update64(d, u, s, out);
    free(u);
}

/* ------------------------------------------------------------------ */
/* 6. statistics                                                       */
/* ------------------------------------------------------------------ */

void or_stats(const float* s, size_t cells, uint32_t channels, ah_stats* st) {
    memset(st, 0, sizeof(*st));
    for (uint32_t c = 0; c < channels; ++c) {
        double m = 0.0;
        float lo = INFINITY, hi = -INFINITY;
        for (size_t i = 0; i < cells; ++i) {
            float v = s[c * cells + i];
            if (!isfinite(v)) { st->nonfinite++; continue; }
            m += (double)v;
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
        st->mass[c] = m;
        st->min[c] = lo;
        st->max[c] = hi;
    }
}

int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
