// Plain FP32 FFT step for gfx950: three fused passes per step.
//
// The reference runs Lenia as 4*log2(N)+5 dispatches per step, one radix-2
// Stockham stage each, ping-ponging a full complex field through memory
// (crates/auto-core/src/automata/spectral.rs:14-64 fft_pass, :199-264 step).
// Here a step is two passes (a third, row_r2c, runs only after the state was
// written from outside), each holding whole transforms on chip:
//
//   1. row_r2c   state rows -> real FFT along x (length-W/2 complex FFT plus
//                the even/odd split) -> half spectrum, 4 rows per workgroup so
//                every 32-byte spectrum block gets a 128-byte contiguous write.
//   2. col_conv  columns of the half spectrum, all H rows: forward FFT along
//                y, multiply by the real kernel spectrum (scaled by 1/(W*H)),
//                inverse FFT along y, written in place.
//   3. row_c2r   inverse real FFT along x, then the growth map, dt and clip
//                (spectral.rs:236-264,337-345) written to the state, and (fused
//                variants) the next step's forward row FFT.
//
// Inside LDS the transforms are Stockham autosort stages of radix 8 (with a
// final radix 2 or 4 stage), the same stage recurrence as fft_pass
// generalised to radix R: thread j, k = j mod Ls, loads v[r] = x[j + r*M/R],
// twiddles by w^(r*k) with w = exp(sign*2*pi*i/(Ls*R)), runs a radix-R DFT in
// registers and stores at (j-k)*R + k + r*Ls.  The column kernels at
// H = 256..4096 run radix-16 stages in registers instead (k_col_blk,
// k_col_block4096), with LDS only to change hands between them.
//
// Twiddle precision: the LDS stages (twiddles()) take the hardware's f32
// sine and cosine of the exact dyadic angle k*step/n turns; the register-
// resident column kernels and the real-FFT twist factors read FP64-built
// tables (init.hip) and form powers by products.  One-step error vs FP64 at 4096^2
// stays near 2e-6 either way (bar: 1e-3).
//
// Half-spectrum layout (complex64): spec[kb][y][c], kx = 4*kb + c.  Bin kx=0
// packs the two real bins X[0] and X[W/2] as (re, im); the column pass treats
// that column as two real sequences sharing one complex transform.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>

#include "ah_internal.h"

namespace ah {

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conj2(float2 a) { return make_float2(a.x, -a.y); }
// a * (SIGN * i)
template <int SIGN>
__device__ __forceinline__ float2 mul_si(float2 a) {
    return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// lds_sync() (ah_internal.h) keeps global stores and global->LDS prefetches
// running across FFT stages.  Nothing in this file exchanges global memory
// between the threads of a workgroup.

// log2 of a power of two.  Every size here (row and column lengths, rows and
// columns per workgroup, stage spans) is one, so index math is shifts and
// masks rather than integer division.
__device__ __forceinline__ int ilog2(int v) { return __builtin_ctz(v); }

// LDS index padding: one float2 every 32 breaks the power-of-two strides.
__device__ __forceinline__ int pad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr size_t padded(size_t n) { return n + (n >> 5) + 1; }

template <int SIGN>
__device__ __forceinline__ void dft2(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int SIGN>
__device__ __forceinline__ void dft4(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <int SIGN>
__device__ __forceinline__ void dft8(float2* v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SIGN>(e);
    dft4<SIGN>(o);
    const float c = 0.70710678118654752f;
    // w8^k = exp(SIGN*2*pi*i*k/8)
    float2 o1 = cmul(o[1], make_float2(c, SIGN * c));
    float2 o2 = mul_si<SIGN>(o[2]);
    float2 o3 = cmul(o[3], make_float2(-c, SIGN * c));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

template <int R, int SIGN>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else dft8<SIGN>(v);
}

// Twiddles w^r, r = 1..R-1, w = exp(SIGN 2 pi i k tstep / nt): w from the
// hardware sine and cosine (which take turns) of the exact angle
// k*tstep/nt, powers by products (w^2 = w w, w^4 = w^2 w^2, ...; a few ulp).
// A table lookup here would be a scattered L2 request per butterfly, and
// these passes are bound by L2 requests.  `tw` is unused: the callers keep
// one signature for table-based and computed twiddles.
template <int R, int SIGN>
__device__ __forceinline__ void twiddles(float2* w, const float2* __restrict__ tw, int k, int tstep, int nt) {
    (void)tw;
    float x = (float)(k * tstep) / (float)nt;  // exact: a dyadic fraction of a turn
    // Opaque, so the compiler never folds a constant angle with its own
    // cosine: every kernel then gets the hardware's bits for every angle.
    asm volatile("" : "+v"(x));
    w[1] = make_float2(__builtin_amdgcn_cosf(x), -__builtin_amdgcn_sinf(x));
    if (SIGN > 0) w[1].y = -w[1].y;
    if constexpr (R >= 4) {
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[1], w[2]);
    }
    if constexpr (R >= 8) {
        w[4] = cmul(w[2], w[2]);
        w[5] = cmul(w[1], w[4]);
        w[6] = cmul(w[2], w[4]);
        w[7] = cmul(w[3], w[4]);
    }
}

// One radix-R Stockham stage over `nseq` sequences of length M stored back
// to back in LDS.  Every thread owns E/R butterflies; all loads happen before
// the first barrier, all stores after it.  tw[t] = exp(-2*pi*i*t/NT);
// the twiddle for r*k over Ls*R points is tw[r*k*(NT/(Ls*R))].
template <int R, int E, int SIGN>
__device__ __forceinline__ void stage(float2* lds, int M, int nseq, int Ls, const float2* __restrict__ tw, int NT) {
    constexpr int NB = E / R;
    const int per_seq = M / R, lg_seq = ilog2(per_seq);
    const int nthreads = blockDim.x, tid = threadIdx.x;
    const int tstep = NT >> ilog2(Ls * R);
    float2 v[NB][R];
    int base[NB], kk[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        int b = tid + i * nthreads;
        int seq = b >> lg_seq, j = b & (per_seq - 1);
        int off = seq * M;
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = lds[pad(off + j + r * per_seq)];
        int k = j & (Ls - 1);
        kk[i] = k;
        base[i] = off + (j - k) * R + k;
    }
    lds_sync();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if (Ls > 1) {
            float2 w[R];
            twiddles<R, SIGN>(w, tw, kk[i], tstep, NT);
#pragma unroll
            for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], w[r]);
        }
        dft<R, SIGN>(v[i]);
#pragma unroll
        for (int r = 0; r < R; ++r) lds[pad(base[i] + r * Ls)] = v[i][r];
    }
    lds_sync();
}

// Complex FFT of length M (power of two, >= 8) on nseq sequences in LDS.
// E = elements per thread = nseq*M/blockDim.x.
template <int E, int SIGN>
__device__ void fft_lds(float2* lds, int M, int nseq, const float2* __restrict__ tw, int NT) {
    constexpr int R0 = E >= 8 ? 8 : 4;  // radix of the main stages (at most E)
    int Ls = 1;
    while (Ls * R0 <= M) {
        stage<R0, E, SIGN>(lds, M, nseq, Ls, tw, NT);
        Ls *= R0;
    }
    if (R0 == 8 && Ls * 4 == M) stage<4, E, SIGN>(lds, M, nseq, Ls, tw, NT);
    else if (Ls * 2 == M) stage<2, E, SIGN>(lds, M, nseq, Ls, tw, NT);
}

// gfx950 hands consecutive workgroup ids to its 8 XCDs in turn.  Remapping
// makes consecutive logical ids run on one XCD at about the same time, so the
// rows (or columns) that share a 128-byte spectrum line meet in one L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    if (G & 7) return b;
    return (b & 7) * (G >> 3) + (b >> 3);
}

constexpr int FFT_THREADS = 512;  // upper bound on the workgroup size

// ------------------------------------------------------------ pass 1
// Even/odd split of `rows` forward-transformed rows in LDS (row r at r*M,
// the length-M complex FFT of z[n] = x[2n] + i x[2n+1]) into the half
// spectrum rows y0..y0+rows-1: X[k] = E[k] + w^k O[k], E = (Z[k] + conj
// Z[M-k])/2, O = -i/2 (Z[k] - conj Z[M-k]), w = exp(-2*pi*i/W).  Output
// float4 q holds bins (kx, kx+1) of one row: kb = q/(2*rows), row, pair.
__device__ __forceinline__ void store_half_spectrum(const float2* lds, int M, int rows, int H, int y0,
                                                    const float2* __restrict__ twW, float2* __restrict__ spec,
                                                    int nt) {
    float4* dst = reinterpret_cast<float4*>(spec);
    const int n4 = rows * M / 2, lg2r = ilog2(2 * rows);
    for (int q = threadIdx.x; q < n4; q += nt) {
        int kb = q >> lg2r, rem = q & (2 * rows - 1);
        int row = rem >> 1, kx = 4 * kb + 2 * (rem & 1);
        int off = row * M;
        float2 out[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            int k = kx + t;
            if (k == 0) {
                float2 z0 = lds[pad(off)];
                out[t] = make_float2(z0.x + z0.y, z0.x - z0.y);
            } else {
                float2 a = lds[pad(off + k)], b = conj2(lds[pad(off + M - k)]);
                float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
                float2 d = csub(a, b);
                float2 o = make_float2(0.5f * d.y, -0.5f * d.x);  // -i/2 * d
                out[t] = cadd(e, cmul(twW[k], o));
            }
        }
        size_t o4 = ((size_t)kb * H + y0 + row) * 2 + (rem & 1);
        dst[o4] = make_float4(out[0].x, out[0].y, out[1].x, out[1].y);
    }
}

// Grid: (H/rows, channels); `rows` consecutive rows per workgroup (one row
// when W >= 4096), rows*M/8 threads.
template <int E>
__global__ void __launch_bounds__(FFT_THREADS) k_row_r2c(int W, int H, int rows, const float* __restrict__ state,
                                                         float2* __restrict__ spec, const float2* __restrict__ twW) {
    extern __shared__ float2 lds[];
    const int M = W >> 1;
    const int y0 = xcd_remap(blockIdx.x, gridDim.x) * rows;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int ch = blockIdx.y;
    state += (size_t)ch * W * H;
    spec += (size_t)ch * (W >> 1) * H;
    // `rows` consecutive rows are one contiguous run of rows*M float2.
    const float4* src = reinterpret_cast<const float4*>(state + (size_t)y0 * W);
    {
        // Every load is issued before the first LDS store: E/2 in flight per thread.
        float4 v[E / 2];
#pragma unroll
        for (int j = 0; j < E / 2; ++j) v[j] = src[tid + j * nt];
#pragma unroll
        for (int j = 0; j < E / 2; ++j) {
            const int i = tid + j * nt;
            lds[pad(2 * i)] = make_float2(v[j].x, v[j].y);
            lds[pad(2 * i + 1)] = make_float2(v[j].z, v[j].w);
        }
    }
    lds_sync();
    fft_lds<E, -1>(lds, M, rows, twW, W);  // table stride W/M = 2 folds into NT = W
    store_half_spectrum(lds, M, rows, H, y0, twW, spec, blockDim.x);
}

// ------------------------------------------------------------ pass 2
// Grid: (W/2/cw, source channels); each workgroup owns `cw` (1, 2 or 4)
// adjacent columns of one 4-column block, all H rows; cw*H/8 threads.
// spec_in: [channels][W/8][H][4]; spec_out: [n_kernels][W/8][H][4] (may alias
// spec_in when one kernel reads one channel).  khat: [n_kernels][W/8][H/2+1][4],
// kpack: [n_kernels][H/2+1].
struct ColArgs {
    int W, H, cw;
    const float2* spec_in;
    float2* spec_out;
    const float* khat;
    const float2* kpack;
    const float2* twH;
    int n_kernels;
    int ksrc[AH_MAX_KERNELS];
    int diag;  // kernel k reads and writes channel k only (batched same-size worlds)
};

// Element (y, column c of this workgroup) of a block in global memory.
__device__ __forceinline__ size_t col_index(int kb, int c, int y, int H) { return ((size_t)kb * H + y) * BC + c; }

__device__ __forceinline__ void load_cols(float2* lds, const float2* __restrict__ src, int kb, int c0, int cw, int H) {
    const int lg = ilog2(cw);
    for (int t = threadIdx.x; t < cw * H; t += blockDim.x) {
        int y = t >> lg, c = t & (cw - 1);
        lds[pad(c * H + y)] = src[col_index(kb, c0 + c, y, H)];
    }
}

__device__ __forceinline__ void store_cols(const float2* lds, float2* __restrict__ dst, int kb, int c0, int cw, int H) {
    const int lg = ilog2(cw);
    for (int t = threadIdx.x; t < cw * H; t += blockDim.x) {
        int y = t >> lg, c = t & (cw - 1);
        dst[col_index(kb, c0 + c, y, H)] = lds[pad(c * H + y)];
    }
}

// Multiply the transformed columns in LDS by kernel k's spectrum, in place.
// Pairs (ky, H-ky) are owned by one thread, which the packed kx=0 column
// needs: D[ky] = A C[ky] + B conj C[-ky] with A=(K0+KN)/2, B=(K0-KN)/2.
__device__ __forceinline__ void multiply_cols(float2* lds, const ColArgs& a, int kb, int c0, int k) {
    const int H = a.H, half = H >> 1, cw = a.cw;
    const int npairs = half + 1, lg = ilog2(cw);
    const float* kh = a.khat + ((size_t)k * (a.W >> 3) + kb) * (size_t)npairs * BC;
    for (int t = threadIdx.x; t < npairs * cw; t += blockDim.x) {
        int p = t >> lg, c = t & (cw - 1);
        int y1 = p, y2 = (H - p) & (H - 1);
        float2 c1 = lds[pad(c * H + y1)], c2 = lds[pad(c * H + y2)];
        float2 d1, d2;
        if (kb == 0 && c0 + c == 0) {
            float2 ab = a.kpack[(size_t)k * npairs + p];
            d1 = cadd(make_float2(ab.x * c1.x, ab.x * c1.y), make_float2(ab.y * c2.x, -ab.y * c2.y));
            d2 = cadd(make_float2(ab.x * c2.x, ab.x * c2.y), make_float2(ab.y * c1.x, -ab.y * c1.y));
        } else {
            float kv = kh[(size_t)p * BC + c0 + c];
            d1 = make_float2(kv * c1.x, kv * c1.y);
            d2 = make_float2(kv * c2.x, kv * c2.y);
        }
        lds[pad(c * H + y1)] = d1;
        if (y2 != y1) lds[pad(c * H + y2)] = d2;
    }
    lds_sync();
}

template <int E>
__global__ void __launch_bounds__(FFT_THREADS) k_col_conv(ColArgs a) {
    extern __shared__ float2 lds[];
    const int H = a.H, cw = a.cw;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int kb = (L * cw) >> 2, c0 = (L * cw) & 3;
    const int ch = blockIdx.y;
    const size_t chan = (size_t)(a.W >> 1) * H;  // float2 per channel / kernel spectrum
    load_cols(lds, a.spec_in + ch * chan, kb, c0, cw, H);
    lds_sync();
    fft_lds<E, -1>(lds, H, cw, a.twH, H);
    int users = 0, last = -1;
    if (a.diag) {
        users = 1;
        last = ch;
    } else {
        for (int k = 0; k < a.n_kernels; ++k)
            if (a.ksrc[k] == ch) {
                ++users;
                last = k;
            }
    }
    if (users == 1) {
        multiply_cols(lds, a, kb, c0, last);
        fft_lds<E, 1>(lds, H, cw, a.twH, H);
        store_cols(lds, a.spec_out + last * chan, kb, c0, cw, H);
        return;
    }
    // Several kernels read this channel: keep its spectrum in registers and
    // rebuild the LDS copy before each kernel's multiply.
    float2 keep[E];
#pragma unroll
    for (int i = 0; i < E; ++i) keep[i] = lds[pad(threadIdx.x + i * blockDim.x)];
    bool first = true;
    for (int k = 0; k < a.n_kernels; ++k) {
        if (a.ksrc[k] != ch) continue;
        if (!first) {
            lds_sync();
#pragma unroll
            for (int i = 0; i < E; ++i) lds[pad(threadIdx.x + i * blockDim.x)] = keep[i];
            lds_sync();
        }
        first = false;
        multiply_cols(lds, a, kb, c0, k);
        fft_lds<E, 1>(lds, H, cw, a.twH, H);
        store_cols(lds, a.spec_out + k * chan, kb, c0, cw, H);
    }
}

// Column pass for H = 4096 with one channel read by one kernel, one
// workgroup per 4-column block of the half spectrum.  The block is one
// contiguous run of 4096 rows x 32 bytes; its 16384 complex values stay in
// registers, 16 per thread over 1024 threads, and pass through LDS only to
// change hands between radix-16 stages.  Every global access of a wave is
// 16 whole rows of the block, 512 contiguous bytes: each L2 request carries
// only bytes this workgroup uses (a column-per-workgroup kernel uses 8 of
// each 32-byte row and leaves the rest of the line to three other
// workgroups).
//
// The length-4096 transform as 16 x 16 x 16: n = n2 + 256 n1 (n2 < 256),
// k = k1 + 16 k2; X[k] = sum_n2 W256^(n2 k2) W4096^(n2 k1) sum_n1 x[n] W16^(n1 k1),
// and the 256-point inner transforms again as n2 = m + 16 n1', k2 = k2a + 16 k2b.
//   phase 1, thread (c, t = n2): loads x[t + 256 j], j < 16; DFT-16 over j -> k1;
//            twiddle W4096^(t k1).  Exchange 1 through LDS (workgroup barrier).
//   phase 2, thread (c, k1, m): n1' < 16; DFT-16 -> k2a; twiddle W256^(m k2a).
//            Exchange 2: a 16 x 16 transpose among the 16 lanes that share
//            (c, k1), all in one wave: wave-local LDS, no barrier.
//   phase 3, thread (c, k1, k2a): DFT-16 over m -> k2b: the thread holds
//            frequencies k = k1 + 16 k2a + 256 k2b, multiplies by K-hat and
//            runs the first inverse DFT-16 in registers.
// The inverse retraces these steps with conjugate twiddles (the transform
// is unnormalised; 1/(W*H) is in K-hat) and stores phase 1's rows back.
namespace c4k {
constexpr int H = 4096;
constexpr int KS = 272;        // LDS float2 per (c, k1): 256 + 16, so two k1 land on distinct bank halves
constexpr int NP = H / 2 + 1;  // K-hat entries per column (real, even in ky)
// NCOL columns of a block per workgroup, 256 threads per column.  The LDS
// stride of a column keeps the NCOL columns of one row on distinct banks.
template <int NCOL>
struct Plan {
    static constexpr int THREADS = 256 * NCOL;
    static constexpr int CS = 16 * KS + (NCOL == 4 ? 8 : 16);
    static constexpr size_t LDS_BYTES = (size_t)NCOL * CS * sizeof(float2);
};
}  // namespace c4k

// X[c + 4d] = sum_a W4^(ad) W16^(ac) sum_b x[a + 4b] W4^(bc), W_n = exp(SIGN 2 pi i / n).
template <int SIGN>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
    const float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
    float2 y[16];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        float2 q[4] = {v[a], v[a + 4], v[a + 8], v[a + 12]};
        dft4<SIGN>(q);
#pragma unroll
        for (int c = 0; c < 4; ++c) y[4 * a + c] = q[c];
    }
    // W16^(ac) for a, c in 1..3: ac = 1, 2, 3, 2, 4, 6, 3, 6, 9.
    y[5] = cmul(y[5], make_float2(c1, SIGN * s1));     // a=1 c=1
    y[6] = cmul(y[6], make_float2(h, SIGN * h));       // a=1 c=2
    y[7] = cmul(y[7], make_float2(s1, SIGN * c1));     // a=1 c=3
    y[9] = cmul(y[9], make_float2(h, SIGN * h));       // a=2 c=1
    y[10] = mul_si<SIGN>(y[10]);                       // a=2 c=2
    y[11] = cmul(y[11], make_float2(-h, SIGN * h));    // a=2 c=3
    y[13] = cmul(y[13], make_float2(s1, SIGN * c1));   // a=3 c=1
    y[14] = cmul(y[14], make_float2(-h, SIGN * h));    // a=3 c=2
    y[15] = cmul(y[15], make_float2(-c1, -SIGN * s1)); // a=3 c=3
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float2 q[4] = {y[c], y[4 + c], y[8 + c], y[12 + c]};
        dft4<SIGN>(q);
#pragma unroll
        for (int d = 0; d < 4; ++d) v[c + 4 * d] = q[d];
    }
}

// v[r] *= w^r, r = 1..15, w = w1 (conjugated for SIGN > 0); each power is at
// most four products deep from w1.
template <int SIGN>
__device__ __forceinline__ void twiddle16(float2 (&v)[16], float2 w1) {
    if (SIGN > 0) w1.y = -w1.y;
    float2 w[16];
    w[1] = w1;
    w[2] = cmul(w1, w1);
    w[3] = cmul(w[2], w1);
    w[4] = cmul(w[2], w[2]);
    w[5] = cmul(w[4], w1);
    w[6] = cmul(w[4], w[2]);
    w[7] = cmul(w[4], w[3]);
    w[8] = cmul(w[4], w[4]);
#pragma unroll
    for (int r = 9; r < 16; ++r) w[r] = cmul(w[8], w[r - 8]);
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], w[r]);
}

__device__ __forceinline__ void wave_lds_order() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// A thread's roles: phase 1, column c1 of the workgroup's columns and row
// t1 (with t1 + 256 j); phases 2 and 3, column c2, k1 and m (phase 2) or
// k2a (phase 3).  `region` is (c2, k1)'s LDS, owned by 16 lanes of one wave.
template <int NCOL>
struct ColRoles {
    static constexpr int LGC = NCOL == 4 ? 2 : NCOL == 2 ? 1 : 0;
    static_assert(NCOL == 1 || NCOL == 2 || NCOL == 4, "columns per workgroup");
    int c1, t1, c2, k1, m;
    float2* region;
    __device__ ColRoles(float2* lds) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        c1 = lane & (NCOL - 1);
        t1 = (lane >> LGC) + (64 >> LGC) * wave;
        c2 = wave >> 2;
        k1 = ((lane >> 4) & 3) + 4 * (wave & 3);
        m = lane & 15;
        region = lds + c2 * c4k::Plan<NCOL>::CS + k1 * c4k::KS;
    }
};

// Logical task L: columns cbase.. of block kb (the NCOL-column tasks of one
// block are consecutive).
template <int NCOL>
__device__ __forceinline__ void col_task_id(int L, int& kb, int& cbase) {
    kb = L >> (2 - ColRoles<NCOL>::LGC);
    cbase = (L << ColRoles<NCOL>::LGC) & 3;
}

// A task's K-hat factors for its phase-3 frequencies k = k1 + 16 m + 256 r
// (none for the packed column 0), and its 16 phase-1 values.
template <int NCOL>
__device__ __forceinline__ void col_task_load_khat(const ColArgs& a, int kb, int cbase, const ColRoles<NCOL>& q,
                                                   float (&kf)[16]) {
    using namespace c4k;
    const int col = kb * BC + cbase + q.c2;
    if (col != 0) {
        // K-hat in digit order (KHAT_DIGITS), real and even in ky: row k
        // takes p = min(k, H - k).  For r < 8, p = k sits at 256 r + 16 k1 + m
        // (a wave reads 256 contiguous bytes); for r >= 8, p = H - k has the
        // digits (16 - k1, 15 - m, 15 - r) (with borrows when k1 or m is 0),
        // which sit at digits(H - k at r = 8) - 256 (r - 8).
        const float* kh = a.khat + (size_t)col * NP;
        const int lo = 16 * q.k1 + q.m;
        const int p8 = H / 2 - q.k1 - 16 * q.m;
        const int hi = (p8 >> 8) * 256 + (p8 & 15) * 16 + ((p8 >> 4) & 15);
#pragma unroll
        for (int r = 0; r < 16; ++r) kf[r] = kh[r < 8 ? 256 * r + lo : hi - 256 * (r - 8)];
    }
}

template <int NCOL>
__device__ __forceinline__ void col_task_load_rows(const ColArgs& a, int kb, int cbase, const ColRoles<NCOL>& q,
                                                   float2 (&v)[16]) {
    const float2* src = a.spec_in + (size_t)kb * c4k::H * BC + cbase;
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = src[(size_t)(q.t1 + 256 * j) * BC + q.c1];
}

// Forward transform, K-hat multiply, inverse transform and store of one task
// whose phase-1 values are in v.  w_t1 = W4096^t1, w_m = W256^m.  `between`
// runs once phase 1's values have gone to LDS (their registers are free).
template <int NCOL, typename F>
__device__ __forceinline__ void col_task_run(const ColArgs& a, float2* lds, int kb, int cbase, const ColRoles<NCOL>& q,
                                             float2 (&v)[16], const float (&kf)[16], float2 w_t1, float2 w_m,
                                             F&& between) {
    using namespace c4k;
    constexpr int CS = Plan<NCOL>::CS;
    const int c1 = q.c1, t1 = q.t1, c2 = q.c2, k1 = q.k1, m = q.m;
    float2* region = q.region;
    // Phase 1.
    dft16<-1>(v);
    twiddle16<-1>(v, w_t1);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[c1 * CS + r * KS + t1] = v[r];
    between();
    lds_sync();
    // Phase 2.
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = region[m + 16 * r];
    dft16<-1>(v);
    twiddle16<-1>(v, w_m);
    wave_lds_order();  // this wave's phase-2 reads of `region` are done
#pragma unroll
    for (int r = 0; r < 16; ++r) region[17 * r + m] = v[r];
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = region[17 * m + r];
    // Phase 3: k = k1 + 16 m + 256 r.
    dft16<-1>(v);
    const int col = kb * BC + cbase + c2;
    if (kb == 0 && cbase == 0) {
        // Workgroup-uniform: column 0 packs the real bins kx = 0 and W/2.
        // D[k] = A C[k] + B conj C[-k] with (A, B) = kpack[min(k, H - k)];
        // C[-k] is held by another thread, so column 0 goes through LDS in
        // natural order (its four waves' regions hold 4352 >= 4096 entries).
        lds_sync();  // column 0's waves have read their exchange-2 data
        if (c2 == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) lds[k1 + 16 * m + 256 * r] = v[r];
        }
        lds_sync();
        if (c2 == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int k = k1 + 16 * m + 256 * r, kn = (H - k) & (H - 1);
                const float2 cn = lds[kn], ab = a.kpack[min(k, kn)];
                v[r] = cadd(make_float2(ab.x * v[r].x, ab.x * v[r].y), make_float2(ab.y * cn.x, -ab.y * cn.y));
            }
        }
        lds_sync();  // partner reads are done before the region is reused
    }
    if (col != 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = make_float2(kf[r] * v[r].x, kf[r] * v[r].y);
    }
    // Inverse: phase 3, exchange 2, phase 2, exchange 1, phase 1.
    dft16<1>(v);
#pragma unroll
    for (int r = 0; r < 16; ++r) region[17 * m + r] = v[r];
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = region[17 * r + m];
    twiddle16<1>(v, w_m);
    dft16<1>(v);
    wave_lds_order();  // the transpose reads are done before the region is rewritten
#pragma unroll
    for (int r = 0; r < 16; ++r) region[m + 16 * r] = v[r];
    lds_sync();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = lds[c1 * CS + r * KS + t1];
    twiddle16<1>(v, w_t1);
    dft16<1>(v);
}

// A task's results (rows t1 + 256 j of column c1, in v) back to the spectrum.
template <int NCOL>
__device__ __forceinline__ void col_task_store(const ColArgs& a, int kb, int cbase, const ColRoles<NCOL>& q,
                                               const float2 (&v)[16]) {
    float2* dst = a.spec_out + (size_t)kb * c4k::H * BC + cbase;
#pragma unroll
    for (int j = 0; j < 16; ++j) dst[(size_t)(q.t1 + 256 * j) * BC + q.c1] = v[j];
}

// One task per workgroup (grid W/2/NCOL).  Workgroups of one block are
// consecutive logical ids, kept on one XCD.
template <int NCOL>
__global__ void __launch_bounds__(c4k::Plan<NCOL>::THREADS) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_col_block4096(ColArgs in) {
    extern __shared__ float2 lds[];
    // Grid y: the kernel, which reads channel ksrc[kernel].
    const int kern = blockIdx.y;
    const size_t chan = (size_t)(in.W >> 1) * c4k::H;
    ColArgs a = in;
    a.spec_in = in.spec_in + in.ksrc[kern] * chan;
    a.spec_out = in.spec_out + kern * chan;
    a.khat = in.khat + kern * (size_t)(in.W >> 1) * c4k::NP;
    a.kpack = in.kpack + kern * c4k::NP;
    int kb, cbase;
    col_task_id<NCOL>(xcd_remap(blockIdx.x, gridDim.x), kb, cbase);
    const ColRoles<NCOL> q(lds);
    float2 v[16];
    float kf[16];
    col_task_load_rows<NCOL>(a, kb, cbase, q, v);
    // The K-hat loads go out once phase 1's values are in LDS, behind the
    // rows rather than beside them (about 1 us less at 4096^2).
    col_task_run<NCOL>(a, lds, kb, cbase, q, v, kf, a.twH[q.t1], a.twH[16 * q.m],
                       [&] { col_task_load_khat<NCOL>(a, kb, cbase, q, kf); });
    col_task_store<NCOL>(a, kb, cbase, q, v);
}

// Column pass for H = 8192: the H = 4096 kernel's form with a radix-32
// first phase.  n = t + 256 j (t < 256, j < 32), k = k1 + 32 k2;
//   phase 1, thread (c, t): rows t + 256 j; DFT-32 over j -> k1; twiddle
//            W_8192^(t k1).  Exchange 1 through LDS (barrier).
//   phases 2 and 3, thread (c, k1lo, m) for the two values k1 = k1lo and
//            k1lo + 16: the 256-point transform over t exactly as at
//            H = 4096 (DFT-16, twiddle W_256^(m k2a), 16 x 16 transpose in
//            one wave, DFT-16), holding k = k1 + 32 m + 512 r.
// 256 threads and 70 KiB of LDS per column (NCOL columns per workgroup);
// K-hat in digit order
// (KHAT_DIGITS at H = 8192: digits(p) = (p >> 9) 512 + (p & 31) 16 +
// ((p >> 5) & 15), so a wave's K-hat loads are contiguous).
namespace c8k {
constexpr int H = 8192;
constexpr int KS = 272;        // LDS float2 per (c, k1) region
constexpr int NP = H / 2 + 1;
template <int NCOL>
struct Plan {
    static constexpr int THREADS = 256 * NCOL;
    static constexpr int CS = 32 * KS + 16;
    static constexpr size_t LDS_BYTES = (size_t)NCOL * CS * sizeof(float2);
};
}  // namespace c8k

// X[k] = E[k mod 16] + W32^k O[k mod 16] with E, O the DFT-16 of the even
// and odd inputs, W32 = exp(SIGN 2 pi i / 32).
template <int SIGN>
__device__ __forceinline__ void dft32(float2 (&v)[32]) {
    constexpr float C[16] = {1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
                             0.70710678118654752f, 0.55557023301960218f, 0.38268343236508977f, 0.19509032201612827f,
                             0.0f, -0.19509032201612827f, -0.38268343236508977f, -0.55557023301960218f,
                             -0.70710678118654752f, -0.83146961230254524f, -0.92387953251128674f,
                             -0.98078528040323043f};
    float2 e[16], o[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        e[b] = v[2 * b];
        o[b] = v[2 * b + 1];
    }
    dft16<SIGN>(e);
    dft16<SIGN>(o);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        // W32^k = (cos(2 pi k / 32), SIGN sin(2 pi k / 32)); sin(2 pi k / 32) = C[8 - k] for k <= 8,
        // C[k - 8] for k > 8.
        float2 t;
        if (k == 0) t = o[0];
        else if (k == 8) t = mul_si<SIGN>(o[8]);
        else t = cmul(o[k], make_float2(C[k], SIGN * (k < 8 ? C[8 - k] : C[k - 8])));
        v[k] = cadd(e[k], t);
        v[k + 16] = csub(e[k], t);
    }
}

// v[r] *= w^r, r = 1..31, w = w1 (conjugated for SIGN > 0).
template <int SIGN>
__device__ __forceinline__ void twiddle32(float2 (&v)[32], float2 w1) {
    if (SIGN > 0) w1.y = -w1.y;
    float2 w[16];
    w[1] = w1;
    w[2] = cmul(w1, w1);
    w[3] = cmul(w[2], w1);
    w[4] = cmul(w[2], w[2]);
    w[5] = cmul(w[4], w1);
    w[6] = cmul(w[4], w[2]);
    w[7] = cmul(w[4], w[3]);
    w[8] = cmul(w[4], w[4]);
#pragma unroll
    for (int r = 9; r < 16; ++r) w[r] = cmul(w[8], w[r - 8]);
    const float2 w16 = cmul(w[8], w[8]);
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], w[r]);
    v[16] = cmul(v[16], w16);
#pragma unroll
    for (int r = 17; r < 32; ++r) v[r] = cmul(v[r], cmul(w16, w[r - 16]));
}

template <int NCOL>
__global__ void __launch_bounds__(256 * NCOL) k_col_block8192(ColArgs in) {
    using namespace c8k;
    constexpr int CS = Plan<NCOL>::CS, LGC = NCOL == 2 ? 1 : 0;
    static_assert(NCOL == 1 || NCOL == 2, "columns per workgroup");
    extern __shared__ float2 lds[];
    // Grid y: the kernel, which reads channel ksrc[kernel].
    const int kern = blockIdx.y;
    const size_t chan = (size_t)(in.W >> 1) * H;
    const float2* spec_in = in.spec_in + in.ksrc[kern] * chan;
    float2* spec_out = in.spec_out + kern * chan;
    const float* khat = in.khat + kern * (size_t)(in.W >> 1) * NP;
    const float2* kpack = in.kpack + kern * NP;
    const int col0 = xcd_remap(blockIdx.x, gridDim.x) * NCOL, kb = col0 >> 2, cbase = col0 & 3;
    const int tid = threadIdx.x;
    const int c1 = tid & (NCOL - 1), t1 = tid >> LGC;                       // phase 1
    const int c2 = tid >> 8, k1lo = (tid & 255) >> 4, m = tid & 15;         // phases 2, 3
    float2* ra = lds + c2 * CS + k1lo * KS;                                 // region of k1 = k1lo
    float2* rb = ra + 16 * KS;                                              // and of k1lo + 16
    const float2* src = spec_in + (size_t)kb * H * BC + cbase;
    {
        float2 v[32];
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = src[(size_t)(t1 + 256 * j) * BC + c1];
        dft32<-1>(v);
        twiddle32<-1>(v, in.twH[t1]);
#pragma unroll
        for (int r = 0; r < 32; ++r) lds[c1 * CS + r * KS + t1] = v[r];
    }
    // K-hat of the phase-3 frequencies k = k1 + 32 m + 512 r, p = min(k, H - k):
    // p = k for r < 8; for r >= 8, p = H - k = p8 - 512 (r - 8), which keeps
    // the lower digits of p8 = H/2 - k1 - 32 m.
    const int col = kb * BC + cbase + c2;
    float kf[2][16];
    if (col != 0) {
        const float* kh = khat + (size_t)col * NP;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int k1 = k1lo + 16 * q;
            const int lo = 16 * k1 + m, p8 = H / 2 - k1 - 32 * m;
            const int hi = (p8 >> 9) * 512 + (p8 & 31) * 16 + ((p8 >> 5) & 15);
#pragma unroll
            for (int r = 0; r < 16; ++r) kf[q][r] = kh[r < 8 ? 512 * r + lo : hi - 512 * (r - 8)];
        }
    }
    lds_sync();
    float2 va[16], vb[16];
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        va[n] = ra[m + 16 * n];
        vb[n] = rb[m + 16 * n];
    }
    const float2 wm = in.twH[32 * m];
    dft16<-1>(va);
    dft16<-1>(vb);
    twiddle16<-1>(va, wm);
    twiddle16<-1>(vb, wm);
    wave_lds_order();  // this wave's phase-2 reads of its regions are done
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        ra[17 * r + m] = va[r];
        rb[17 * r + m] = vb[r];
    }
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        va[r] = ra[17 * m + r];
        vb[r] = rb[17 * m + r];
    }
    dft16<-1>(va);
    dft16<-1>(vb);
    if (kb == 0 && cbase == 0) {
        // Workgroup-uniform: column 0 packs the real bins kx = 0 and W/2;
        // D[k] = A C[k] + B conj C[-k], (A, B) = kpack[min(k, H - k)], with
        // C[-k] from LDS in natural order (column 0's regions hold
        // 32 KS >= H entries).
        lds_sync();
        if (c2 == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                lds[k1lo + 32 * m + 512 * r] = va[r];
                lds[k1lo + 16 + 32 * m + 512 * r] = vb[r];
            }
        }
        lds_sync();
        if (c2 == 0) {
            auto packed = [&](float2 x, int k) {
                const int kn = (H - k) & (H - 1);
                const float2 cn = lds[kn], ab = kpack[min(k, kn)];
                return cadd(make_float2(ab.x * x.x, ab.x * x.y), make_float2(ab.y * cn.x, -ab.y * cn.y));
            };
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                va[r] = packed(va[r], k1lo + 32 * m + 512 * r);
                vb[r] = packed(vb[r], k1lo + 16 + 32 * m + 512 * r);
            }
        }
        lds_sync();  // partner reads are done before the regions are reused
    }
    if (col != 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            va[r] = make_float2(kf[0][r] * va[r].x, kf[0][r] * va[r].y);
            vb[r] = make_float2(kf[1][r] * vb[r].x, kf[1][r] * vb[r].y);
        }
    }
    // Inverse: phase 3, exchange 2, phase 2, exchange 1, phase 1.
    dft16<1>(va);
    dft16<1>(vb);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        ra[17 * m + r] = va[r];
        rb[17 * m + r] = vb[r];
    }
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        va[r] = ra[17 * r + m];
        vb[r] = rb[17 * r + m];
    }
    twiddle16<1>(va, wm);
    twiddle16<1>(vb, wm);
    dft16<1>(va);
    dft16<1>(vb);
    wave_lds_order();  // the transpose reads are done before the regions are rewritten
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        ra[m + 16 * n] = va[n];
        rb[m + 16 * n] = vb[n];
    }
    lds_sync();
    float2 v[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r] = lds[c1 * CS + r * KS + t1];
    twiddle32<1>(v, in.twH[t1]);
    dft32<1>(v);
    float2* dst = spec_out + (size_t)kb * H * BC + cbase;
#pragma unroll
    for (int j = 0; j < 32; ++j) dst[(size_t)(t1 + 256 * j) * BC + c1] = v[j];
}

// Column pass for H = 256..2048 in the same register-resident form as the
// H = 4096 kernel above, for one kernel per channel (batched same-size
// worlds in diagonal mode, or a single world with one kernel).  H = 16 S,
// S = 16 Q (Q = 1, 2, 4, 8): n = t + S j, k = k1 + 16 k2;
//   phase 1, thread (c, t): rows t + S j, j < 16; DFT-16 over j -> k1;
//            twiddle W_H^(t k1).  Exchange 1 through LDS (barrier).
//   phase 2, thread (c, k1, m), m < Q: t = m + Q n, n < 16; DFT-16 over n
//            -> k2a; twiddle W_S^(m k2a).  Exchange 2 among the Q lanes of
//            (c, k1), one wave, no barrier: thread m' takes k2a = m' 16/Q + i.
//   phase 3: DFT-Q over m -> k2b for each i: the thread holds
//            k = k1 + 16 (m' 16/Q + i) + 256 k2b, multiplies by K-hat and
//            runs the inverse back through the same steps.
// One workgroup takes NCOL columns of a 4-column block: its loads and
// stores are whole 32-byte rows of the block when NCOL = 4.  Grid y is the
// world (diagonal mode) or the kernel, which reads channel ksrc[kernel]
// (several kernels of one channel each transform it: no workgroup holds
// more than one column's values).
namespace cblk {
template <int H>
struct Dims {
    static constexpr int S = H / 16, Q = S / 16, P16 = 16 / Q;
    // LDS float2 per (c, k1) region and the exchange-2 index a Q + m + a/P16:
    // both chosen so every phase's LDS access is free of bank conflicts.
    static constexpr int KS = H == 256 ? 17 : H == 512 ? 34 : H == 1024 ? 76 : 152;
    static constexpr int NP = H / 2 + 1;
    static_assert(H >= 256 && H <= 2048 && (H & (H - 1)) == 0, "column length");
};
template <int H, int NCOL>
struct Plan {
    static constexpr int THREADS = Dims<H>::S * NCOL;
    static constexpr int CS = 16 * Dims<H>::KS + (NCOL == 4 ? 4 : 8);
    static constexpr size_t LDS_BYTES = (size_t)NCOL * CS * sizeof(float2);
};
__device__ __forceinline__ constexpr int x2(int a, int m, int Q, int P16) { return a * Q + m + a / P16; }
}  // namespace cblk

template <int Q, int SIGN>
__device__ __forceinline__ void dft_q(float2* v) {
    if constexpr (Q == 2) dft2<SIGN>(v);
    else if constexpr (Q == 4) dft4<SIGN>(v);
    else if constexpr (Q == 8) dft8<SIGN>(v);
}

template <int H, int NCOL>
__global__ void __launch_bounds__((cblk::Plan<H, NCOL>::THREADS)) k_col_blk(ColArgs a) {
    using D = cblk::Dims<H>;
    using P = cblk::Plan<H, NCOL>;
    constexpr int S = D::S, Q = D::Q, P16 = D::P16, KS = D::KS, NP = D::NP, CS = P::CS;
    constexpr int LGC = NCOL == 4 ? 2 : NCOL == 2 ? 1 : 0;
    extern __shared__ float2 lds[];
    int kb, cbase;
    col_task_id<NCOL>(xcd_remap(blockIdx.x, gridDim.x), kb, cbase);
    const int world = blockIdx.y;  // world, or kernel: K-hat, kpack and output
    const int src_ch = a.diag ? world : a.ksrc[world];
    const size_t chan = (size_t)(a.W >> 1) * H;
    const int tid = threadIdx.x;
    const int c1 = tid & (NCOL - 1), t1 = tid >> LGC;            // phase 1
    const int c2 = tid / S, u = tid & (S - 1), k1 = u / Q, m = u & (Q - 1);  // phases 2, 3
    float2* region = lds + c2 * CS + k1 * KS;
    const float2* src = a.spec_in + src_ch * chan + (size_t)kb * H * BC + cbase;
    float2 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = src[(size_t)(t1 + S * j) * BC + c1];
    // Phase 1.
    dft16<-1>(v);
    twiddle16<-1>(v, a.twH[t1]);
#pragma unroll
    for (int r = 0; r < 16; ++r) lds[c1 * CS + r * KS + t1] = v[r];
    // K-hat of this thread's phase-3 frequencies (real, even in ky: row k
    // takes p = min(k, H - k)), behind the rows.
    const int col = kb * BC + cbase + c2;
    float kf[16];
    if (col != 0) {
        const float* kh = a.khat + ((size_t)world * (a.W >> 3) + kb) * (size_t)NP * BC + cbase + c2;
#pragma unroll
        for (int i = 0; i < P16; ++i)
#pragma unroll
            for (int b = 0; b < Q; ++b) {
                const int k = k1 + 16 * (m * P16 + i) + 256 * b;
                kf[i * Q + b] = kh[(size_t)min(k, H - k) * BC];
            }
    }
    lds_sync();
    // Phase 2.
#pragma unroll
    for (int n = 0; n < 16; ++n) v[n] = region[m + Q * n];
    dft16<-1>(v);
    if constexpr (Q > 1) {
        twiddle16<-1>(v, a.twH[16 * m]);
        wave_lds_order();  // this wave's phase-2 reads of `region` are done
#pragma unroll
        for (int r = 0; r < 16; ++r) region[cblk::x2(r, m, Q, P16)] = v[r];
        wave_lds_order();
#pragma unroll
        for (int i = 0; i < P16; ++i)
#pragma unroll
            for (int b = 0; b < Q; ++b) v[i * Q + b] = region[cblk::x2(m * P16 + i, b, Q, P16)];
        // Phase 3.
#pragma unroll
        for (int i = 0; i < P16; ++i) dft_q<Q, -1>(&v[i * Q]);
    }
    if (kb == 0 && cbase == 0) {
        // Workgroup-uniform: column 0 packs the real bins kx = 0 and W/2.
        // D[k] = A C[k] + B conj C[-k] with (A, B) = kpack[min(k, H - k)];
        // C[-k] is held by another thread: column 0 goes through LDS in
        // natural order (its regions hold 16 KS >= H entries).
        lds_sync();  // column 0's exchange-2 reads are done
        if (c2 == 0) {
#pragma unroll
            for (int i = 0; i < P16; ++i)
#pragma unroll
                for (int b = 0; b < Q; ++b) lds[k1 + 16 * (m * P16 + i) + 256 * b] = v[i * Q + b];
        }
        lds_sync();
        if (c2 == 0) {
            const float2* kp = a.kpack + (size_t)world * NP;
#pragma unroll
            for (int i = 0; i < P16; ++i)
#pragma unroll
                for (int b = 0; b < Q; ++b) {
                    const int k = k1 + 16 * (m * P16 + i) + 256 * b, kn = (H - k) & (H - 1);
                    const float2 cn = lds[kn], ab = kp[min(k, kn)], x = v[i * Q + b];
                    v[i * Q + b] = cadd(make_float2(ab.x * x.x, ab.x * x.y), make_float2(ab.y * cn.x, -ab.y * cn.y));
                }
        }
        lds_sync();  // partner reads are done before the region is reused
    }
    if (col != 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = make_float2(kf[r] * v[r].x, kf[r] * v[r].y);
    }
    // Inverse: phase 3, exchange 2, phase 2, exchange 1, phase 1.
    if constexpr (Q > 1) {
#pragma unroll
        for (int i = 0; i < P16; ++i) dft_q<Q, 1>(&v[i * Q]);
#pragma unroll
        for (int i = 0; i < P16; ++i)
#pragma unroll
            for (int b = 0; b < Q; ++b) region[cblk::x2(m * P16 + i, b, Q, P16)] = v[i * Q + b];
        wave_lds_order();
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = region[cblk::x2(r, m, Q, P16)];
        twiddle16<1>(v, a.twH[16 * m]);
    }
    dft16<1>(v);
    wave_lds_order();  // the exchange reads are done before the region is rewritten
#pragma unroll
    for (int n = 0; n < 16; ++n) region[m + Q * n] = v[n];
    lds_sync();
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = lds[c1 * CS + r * KS + t1];
    twiddle16<1>(v, a.twH[t1]);
    dft16<1>(v);
    float2* dst = a.spec_out + world * chan + (size_t)kb * H * BC + cbase;
#pragma unroll
    for (int j = 0; j < 16; ++j) dst[(size_t)(t1 + S * j) * BC + c1] = v[j];
}

// Fused column pass for batched same-size worlds (diagonal mode: world c's
// kernel reads and writes channel c), H = 64..2048.  A workgroup takes CPW
// columns of one block (all four for H <= 1024) with TPC = H/8 threads per
// column.  Its large workgroups pay off only with many of them: measured on
// single worlds (128-512 workgroups) the generic kernel is faster, on a
// batch group of ~100 worlds this one is.  The forward transform's last stage (radix RL, span L = H/RL)
// leaves thread t of a column holding rows t + TPC m, m = 0..7, exactly the
// rows the inverse transform's first stage (radix 8, span 1) reads: the
// K-hat multiply and that stage run in registers, saving two LDS round trips.
template <int H>
struct ColPlan {
    static constexpr int TPC = H / 8, CPW = H <= 1024 ? 4 : 2, THREADS = TPC * CPW;
    static constexpr int N8 = H >= 512 ? 3 : 2;                   // radix-8 stages with 8^N8 <= H
    static constexpr int RLAST = H / (N8 == 3 ? 512 : 64);        // 1, 2 or 4
    static constexpr int RL = RLAST == 1 ? 8 : RLAST, L = H / RL, NBL = 8 / RL;
    static constexpr int NLDS = RLAST == 1 ? N8 - 1 : N8;         // forward stages through LDS
    static_assert(H >= 64 && H <= 2048 && (H & (H - 1)) == 0, "column length");
};

template <int H>
__global__ void __launch_bounds__(ColPlan<H>::THREADS) k_col_fused(ColArgs a) {
    using P = ColPlan<H>;
    extern __shared__ float2 lds[];
    constexpr int GPB = 4 / P::CPW;  // workgroups per block
    // The grid size (all blocks) stays a multiple of 8, so xcd_remap keeps
    // the GPB workgroups of a block on one XCD (one L2).
    const int L0 = xcd_remap(blockIdx.x, gridDim.x);
    const int kb = L0 / GPB, c0 = (L0 % GPB) * P::CPW;
    const int ch = blockIdx.y;  // world ch: channel ch, kernel ch
    const size_t chan = (size_t)(a.W >> 1) * H;
    const float2* src = a.spec_in + ch * chan;
    float2* dst = a.spec_out + ch * chan;
    // Load: thread t takes (row, column) = (t / CPW + THREADS / CPW * i, t % CPW);
    // the CPW columns of a row are adjacent in memory.
    {
        float2 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = threadIdx.x + P::THREADS * i;
            v[i] = src[col_index(kb, c0 + e % P::CPW, e / P::CPW, H)];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = threadIdx.x + P::THREADS * i;
            lds[pad((e % P::CPW) * H + e / P::CPW)] = v[i];
        }
    }
    lds_sync();
    int Ls = 1;
#pragma unroll
    for (int s = 0; s < P::NLDS; ++s, Ls *= 8) stage<8, 8, -1>(lds, H, P::CPW, Ls, a.twH, H);
    const int col = threadIdx.x / P::TPC, t = threadIdx.x % P::TPC, off = col * H;
    // K-hat factors of this thread's rows y = t + TPC m: real and even in ky,
    // khat[p = min(y, H - y)] of column (kb, c0 + col).
    constexpr int NP = H / 2 + 1;
    const float* kh = a.khat + ((size_t)ch * (a.W >> 3) + kb) * (size_t)NP * BC + c0 + col;
    float kf[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int y = t + P::TPC * m;
        kf[m] = kh[(size_t)min(y, H - y) * BC];
    }
    // Last forward stage (radix RL, span L) into registers: output r of
    // butterfly i = t + TPC i lands on row t + TPC (i + NBL r).
    float2 v[8];
#pragma unroll
    for (int i = 0; i < P::NBL; ++i) {
        float2 q[P::RL];
        const int b = t + P::TPC * i;
#pragma unroll
        for (int r = 0; r < P::RL; ++r) q[r] = lds[pad(off + b + r * P::L)];
        float2 w[P::RL];
        twiddles<P::RL, -1>(w, a.twH, b, 1, H);  // table stride H / (L * RL)
#pragma unroll
        for (int r = 1; r < P::RL; ++r) q[r] = cmul(q[r], w[r]);
        dft<P::RL, -1>(q);
#pragma unroll
        for (int r = 0; r < P::RL; ++r) v[i + P::NBL * r] = q[r];
    }
    lds_sync();  // every thread has read its last-stage inputs
    if (kb == 0 && c0 == 0) {
        // This workgroup holds the packed kx = 0 column (col 0): D[y] =
        // A C[y] + B conj C[H - y] with (A, B) = kpack[min(y, H - y)] needs
        // row H - y, which thread TPC - t holds (at m' = 7 - m; thread 0's
        // own rows pair among themselves).  One exchange through LDS; the
        // barriers are workgroup-uniform.
        const bool packed = col == 0;
        if (packed) {
#pragma unroll
            for (int m = 0; m < 8; ++m) lds[pad(t + P::TPC * m)] = v[m];
        }
        lds_sync();
        if (packed) {
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int y = t + P::TPC * m, y2 = (H - y) & (H - 1);
                const float2 c2 = lds[pad(y2)], ab = a.kpack[(size_t)ch * NP + min(y, y2)];
                v[m] = cadd(make_float2(ab.x * v[m].x, ab.x * v[m].y), make_float2(ab.y * c2.x, -ab.y * c2.y));
            }
        } else {
#pragma unroll
            for (int m = 0; m < 8; ++m) v[m] = make_float2(kf[m] * v[m].x, kf[m] * v[m].y);
        }
        lds_sync();  // the partner reads are done
    } else {
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = make_float2(kf[m] * v[m].x, kf[m] * v[m].y);
    }
    // First inverse stage (radix 8, span 1) on the registers.
    dft8<1>(v);
#pragma unroll
    for (int r = 0; r < 8; ++r) lds[pad(off + 8 * t + r)] = v[r];
    lds_sync();
    Ls = 8;
#pragma unroll
    for (int s = 1; s < P::N8; ++s, Ls *= 8) stage<8, 8, 1>(lds, H, P::CPW, Ls, a.twH, H);
    if constexpr (P::RLAST == 4) stage<4, 8, 1>(lds, H, P::CPW, Ls, a.twH, H);
    if constexpr (P::RLAST == 2) stage<2, 8, 1>(lds, H, P::CPW, Ls, a.twH, H);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int e = threadIdx.x + P::THREADS * i;
        dst[col_index(kb, c0 + e % P::CPW, e / P::CPW, H)] = lds[pad((e % P::CPW) * H + e / P::CPW)];
    }
}

// ------------------------------------------------------------ pass 3
struct RowArgs {
    int W, H, rows;
    const float2* spec;  // [n_kernels][W/8][H][4]
    const float* state_in;
    float* state_out;
    const float2* twW;
    float dt;
    int n_kernels;
    KernelParams kp[AH_MAX_KERNELS];
    float* conv_out;  // when set: write kernel conv_k's potential instead of updating
    int conv_k;
    float2* spec_next;  // NEXT: half spectra [channels][W/8][H][4] of the updated state
    // Batched same-size worlds: channel c is world c with one kernel (k = c)
    // whose growth and dt come from these device arrays.
    const KernelParams* kp_dev;
    const float* dt_dev;
};

// Load kernel k's spectrum rows into LDS and turn them into the length-M
// complex sequence whose inverse FFT interleaves the real output:
// Z[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) conj(w^k).
__device__ void load_c2r(float2* lds, const RowArgs& a, int k, int y0) {
    const int W = a.W, H = a.H, M = W >> 1, rows = a.rows;
    const float4* s4 = reinterpret_cast<const float4*>(a.spec + (size_t)k * M * H);
    const int n4 = rows * M / 2, lg2r = ilog2(2 * rows);
    for (int q = threadIdx.x; q < n4; q += blockDim.x) {
        int kb = q >> lg2r, rem = q & (2 * rows - 1);
        int row = rem >> 1, kx = 4 * kb + 2 * (rem & 1);
        float4 v = s4[((size_t)kb * H + y0 + row) * 2 + (rem & 1)];
        lds[pad(row * M + kx)] = make_float2(v.x, v.y);
        lds[pad(row * M + kx + 1)] = make_float2(v.z, v.w);
    }
    lds_sync();
    // Thread t owns bins kk and M-kk of one row (kk in 1..M/2-1); kk = 0 also
    // takes the DC/Nyquist pair and the self-paired middle bin M/2.
    const int half = M >> 1, lgh = ilog2(half);
    auto z_of = [&](float2 x1, float2 x2, float2 cw) {
        float2 s1 = cadd(x1, conj2(x2)), d1 = csub(x1, conj2(x2));
        return cadd(s1, cmul(make_float2(-d1.y, d1.x), cw));
    };
    for (int t = threadIdx.x; t < rows * half; t += blockDim.x) {
        int row = t >> lgh, kk = t & (half - 1);
        int off = row * M;
        if (kk == 0) {
            float2 p = lds[pad(off)];
            lds[pad(off)] = make_float2(p.x + p.y, p.x - p.y);
            float2 xm = lds[pad(off + half)];
            lds[pad(off + half)] = z_of(xm, xm, conj2(a.twW[half]));
            continue;
        }
        int k2 = M - kk;
        float2 x1 = lds[pad(off + kk)], x2 = lds[pad(off + k2)];
        // Z[kk] from (x1, x2); Z[k2] from (x2, x1) with conj(w^k2) = -w^kk.
        float2 w1 = a.twW[kk];
        lds[pad(off + kk)] = z_of(x1, x2, conj2(w1));
        lds[pad(off + k2)] = z_of(x2, x1, make_float2(-w1.x, -w1.y));
    }
    lds_sync();
}

// Rafler's SmoothLife transition (spectral.rs:393-405): c0 is the inner-disc
// average, c1 the outer-ring average; mix(a, b, t) = a (1 - t) + b t.
__device__ __forceinline__ float smoothlife_next(float s, float c0, float c1, float dt) {
    constexpr float B1 = 0.278f, B2 = 0.365f, D1 = 0.267f, D2 = 0.445f, AN = 0.028f, AM = 0.147f;
    auto sig = [](float x, float a, float alpha) { return 1.0f / (1.0f + __expf((a - x) * 4.0f / alpha)); };
    const float m = sig(c0, 0.5f, AM);
    const float lo = B1 * (1.0f - m) + D1 * m, hi = B2 * (1.0f - m) + D2 * m;
    const float target = sig(c1, lo, AN) * (1.0f - sig(c1, hi, AN));
    return fminf(fmaxf(s + dt * (2.0f * target - 1.0f), 0.0f), 1.0f);
}

// MNCA bands (spectral.rs:462-473): each band that holds overrules the ones
// before it; the state is kept, set to 1 or set to 0.
__device__ __forceinline__ float mnca_next(float s, float c0, float c1) {
    auto band = [](float v, float lo, float hi) { return v >= lo && v <= hi; };
    float o = s;
    if (band(c0, 0.210f, 0.220f)) o = 1.0f;
    if (band(c0, 0.350f, 0.500f)) o = 0.0f;
    if (band(c0, 0.750f, 0.850f)) o = 0.0f;
    if (band(c0, 0.120f, 0.150f)) o = 0.0f;
    if (band(c1, 0.100f, 0.280f)) o = 0.0f;
    if (band(c1, 0.430f, 0.550f)) o = 1.0f;
    return o;
}

// RULE: AH_RULE_LENIA sums each kernel's growth into its destination;
// SmoothLife and MNCA keep the two potentials of their single channel and
// apply their transition to them.
// NEXT: the updated rows are also the next step's pass-1 input, so they go
// back into LDS and through the forward real FFT here, and the state is read
// once per step instead of twice.  In place when spec_next aliases spec: a
// workgroup reads and writes only its own rows, and every read of them is
// behind the barrier at the end of the kernel loop.
template <int E, bool NEXT, int RULE>
__global__ void __launch_bounds__(FFT_THREADS) k_row_c2r(RowArgs a) {
    extern __shared__ float2 lds[];
    const int W = a.W, H = a.H, M = W >> 1, rows = a.rows;
    const int y0 = xcd_remap(blockIdx.x, gridDim.x) * rows;
    const int ch = blockIdx.y;  // destination channel
    const int nt = blockDim.x, tid = threadIdx.x;
    constexpr int NQ = E / 2;  // float4 (4 output cells) per thread
    float acc[NQ][4];                                  // growth sum, or potential 0
    float pot1[RULE == AH_RULE_LENIA ? 1 : NQ][4];     // potential 1 (SmoothLife, MNCA)
#pragma unroll
    for (int i = 0; i < NQ; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0f;
    const bool diag = a.kp_dev != nullptr;
    const int k_begin = diag ? ch : 0, k_end = diag ? ch + 1 : a.n_kernels;
    const float dt = diag ? a.dt_dev[ch] : a.dt;
    for (int k = k_begin; k < k_end; ++k) {
        const KernelParams kp = diag ? a.kp_dev[k] : a.kp[k];
        if (!diag && (a.conv_out ? (k != a.conv_k) : ((int)kp.dst != ch))) continue;
        load_c2r(lds, a, k, y0);
        fft_lds<E, 1>(lds, M, rows, a.twW, W);
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            int q = tid + i * nt;  // float4 index within the rows*W/4 cells
            float2 u0 = lds[pad(2 * q)], u1 = lds[pad(2 * q + 1)];
            float u[4] = {u0.x, u0.y, u1.x, u1.y};
            if (a.conv_out) {
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = u[j];
            } else if constexpr (RULE == AH_RULE_LENIA) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float d = (u[j] - kp.mu) * kp.inv_sigma;
                    acc[i][j] += kp.weight * (2.0f * __expf(-0.5f * d * d) - 1.0f);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (k == 0) acc[i][j] = u[j];
                    else pot1[i][j] = u[j];
                }
            }
        }
        lds_sync();
    }
    const size_t cell0 = (size_t)y0 * W;
    if (a.conv_out) {
        float4* o4 = reinterpret_cast<float4*>(a.conv_out + cell0);
#pragma unroll
        for (int i = 0; i < NQ; ++i) o4[tid + i * nt] = make_float4(acc[i][0], acc[i][1], acc[i][2], acc[i][3]);
        return;
    }
    const size_t plane = (size_t)W * H * ch;
    const float4* s4 = reinterpret_cast<const float4*>(a.state_in + plane + cell0);
    float4* o4 = reinterpret_cast<float4*>(a.state_out + plane + cell0);
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        int q = tid + i * nt;
        float4 s = s4[q];
        float r[4] = {s.x, s.y, s.z, s.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if constexpr (RULE == AH_RULE_SMOOTHLIFE) r[j] = smoothlife_next(r[j], acc[i][j], pot1[i][j], dt);
            else if constexpr (RULE == AH_RULE_MNCA) r[j] = mnca_next(r[j], acc[i][j], pot1[i][j]);
            else r[j] = fminf(fmaxf(r[j] + dt * acc[i][j], 0.0f), 1.0f);
        }
        o4[q] = make_float4(r[0], r[1], r[2], r[3]);
        if constexpr (NEXT) {
            lds[pad(2 * q)] = make_float2(r[0], r[1]);
            lds[pad(2 * q + 1)] = make_float2(r[2], r[3]);
        }
    }
    if constexpr (NEXT) {
        lds_sync();
        fft_lds<E, -1>(lds, M, rows, a.twW, W);
        store_half_spectrum(lds, M, rows, H, y0, a.twW, a.spec_next + (size_t)ch * M * H, blockDim.x);
    }
}

// Pass 3 fused with the next pass 1 for the Lenia rule with one kernel per
// destination channel (a single world, or batched same-size worlds in
// diagonal mode): k_row_c2r<8, true, LENIA> with the middle in registers.
// Rows of M complex (W = 2M, 128 <= W <= 8192), 2048/M rows per workgroup
// of 256 threads (one row of 512 threads at W = 8192), TPR = M/8 threads
// per row.  The inverse FFT's last stage
// (radix RL = 8, 4 or 2, span L = M/RL) leaves thread t of a row holding
// positions t + TPR m, m = 0..7; complex position p is cells 2p and 2p+1,
// so the thread applies the growth to exactly those cells; and the forward
// FFT's first stage (radix 8, span 1) reads positions t + TPR m, the same
// set.  The LDS round trips of the epilogue and of that stage disappear.
template <int M>
struct RowPlan {
    static constexpr int TPR = M / 8, ROWS = M >= 2048 ? 1 : 2048 / M, THREADS = TPR * ROWS;
    static constexpr int N8 = M >= 4096 ? 4 : M >= 512 ? 3 : 2;   // radix-8 stages with 8^N8 <= M
    static constexpr int RLAST = M / (N8 == 4 ? 4096 : N8 == 3 ? 512 : 64);  // 1, 2 or 4
    static constexpr int RL = RLAST == 1 ? 8 : RLAST, L = M / RL, NBL = 8 / RL;
    static constexpr int NLDS = RLAST == 1 ? N8 - 1 : N8;         // inverse stages through LDS
    static_assert(M >= 64 && M <= 4096 && (M & (M - 1)) == 0, "row length");
};

// Pass-3 input of the workgroup's rows, twisted on the way in: thread t of
// a row loads bins kk = t + TPR i and M - kk of the product spectrum straight
// from memory and writes Z[kk], Z[M-kk] of load_c2r (kk = 0 also takes the
// packed DC / Nyquist bin and the self-paired middle bin M/2), saving an
// LDS pass.
template <int M>
__device__ __forceinline__ void load_twist_rows(float2* lds, const RowArgs& a, const float2* __restrict__ spec,
                                                int y0) {
    using P = RowPlan<M>;
    constexpr int half = M / 2;
    const int row = P::ROWS == 1 ? 0 : threadIdx.x / P::TPR, t = P::ROWS == 1 ? threadIdx.x : threadIdx.x % P::TPR;
    const int off = row * M, H = a.H, y = y0 + row;
    auto bin = [&](int k) { return spec[((size_t)(k >> 2) * H + y) * BC + (k & 3)]; };
    auto z_of = [&](float2 x1, float2 x2, float2 cw) {
        float2 s1 = cadd(x1, conj2(x2)), d1 = csub(x1, conj2(x2));
        return cadd(s1, cmul(make_float2(-d1.y, d1.x), cw));
    };
    float2 x1[half / P::TPR], x2[half / P::TPR];
#pragma unroll
    for (int i = 0; i < half / P::TPR; ++i) {
        const int kk = t + P::TPR * i;
        x1[i] = bin(kk);
        x2[i] = bin(kk == 0 ? half : M - kk);
    }
#pragma unroll
    for (int i = 0; i < half / P::TPR; ++i) {
        const int kk = t + P::TPR * i;
        if (kk == 0) {
            lds[pad(off)] = make_float2(x1[i].x + x1[i].y, x1[i].x - x1[i].y);
            lds[pad(off + half)] = z_of(x2[i], x2[i], conj2(a.twW[half]));
            continue;
        }
        const float2 w1 = a.twW[kk];
        lds[pad(off + kk)] = z_of(x1[i], x2[i], conj2(w1));
        lds[pad(off + M - kk)] = z_of(x2[i], x1[i], make_float2(-w1.x, -w1.y));
    }
    lds_sync();
}

// Kernel k's row through the inverse transform into registers: the
// twisted load, the stages through LDS (spans 1, 8, ...), and the last stage
// (radix RL, span L) whose output r of butterfly i = t + TPR i lands on
// position t + TPR m, m = i + NBL r, in out[m].
template <int M>
__device__ __forceinline__ void inverse_row(float2* lds, const RowArgs& a, const float2* __restrict__ spec, int y0,
                                            int off, int t, float2 (&out)[8]) {
    using P = RowPlan<M>;
    constexpr int W = 2 * M;
    load_twist_rows<M>(lds, a, spec, y0);
    int Ls = 1;
#pragma unroll
    for (int s = 0; s < P::NLDS; ++s, Ls *= 8) stage<8, 8, 1>(lds, M, P::ROWS, Ls, a.twW, W);
#pragma unroll
    for (int i = 0; i < P::NBL; ++i) {
        float2 q[P::RL];
        const int b = t + P::TPR * i;
#pragma unroll
        for (int r = 0; r < P::RL; ++r) q[r] = lds[pad(off + b + r * P::L)];
        float2 w[P::RL];
        twiddles<P::RL, 1>(w, a.twW, b, W / M, W);  // table stride W / (L * RL)
#pragma unroll
        for (int r = 1; r < P::RL; ++r) q[r] = cmul(q[r], w[r]);
        dft<P::RL, 1>(q);
#pragma unroll
        for (int r = 0; r < P::RL; ++r) out[i + P::NBL * r] = q[r];
    }
}

// The rule, by mode: LENIA with one kernel (a world, or a batched world in
// diagonal mode, DIAG); LENIA with several kernels per destination channel
// (MULTI: each kernel ending in this channel adds its weighted growth, in
// kernel order, as the generic kernel does); SmoothLife and MNCA (two kernels
// of the one channel, whose potentials the transition combines).  Kernel
// k's values stay in registers while the next kernel's row goes through LDS.
template <int M, bool DIAG, int RULE, bool MULTI>
__global__ void __launch_bounds__(RowPlan<M>::THREADS) k_row_fused(RowArgs a) {
    using P = RowPlan<M>;
    extern __shared__ float2 lds[];
    constexpr int W = 2 * M;
    const int H = a.H;
    const int y0 = xcd_remap(blockIdx.x, gridDim.x) * P::ROWS;
    const int ch = blockIdx.y;  // destination channel (world in diagonal mode)
    const int row = P::ROWS == 1 ? 0 : threadIdx.x / P::TPR, t = P::ROWS == 1 ? threadIdx.x : threadIdx.x % P::TPR;
    const int off = row * M;
    const size_t plane = (size_t)W * H * ch, cell0 = plane + (size_t)(y0 + row) * W;
    const float2* srow = reinterpret_cast<const float2*>(a.state_in + cell0);
    float2* orow = reinterpret_cast<float2*>(a.state_out + cell0);
    const size_t kspec = (size_t)M * H;
    float2 sv[8];  // state cells (2p, 2p+1) of positions p = t + TPR m
    float2 ns[8];  // the new state of those cells
    if constexpr (MULTI) {
        float acc[8][2];
#pragma unroll
        for (int m = 0; m < 8; ++m) acc[m][0] = acc[m][1] = 0.0f;
        bool first = true;
        for (int k = 0; k < a.n_kernels; ++k) {
            const KernelParams kp = a.kp[k];
            if ((int)kp.dst != ch) continue;  // uniform over the workgroup
            if (!first) lds_sync();           // kernel k-1's last-stage inputs are read
            first = false;
            float2 u[8];
            inverse_row<M>(lds, a, a.spec + k * kspec, y0, off, t, u);
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const float uu[2] = {u[m].x, u[m].y};
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    float d = (uu[j] - kp.mu) * kp.inv_sigma;
                    acc[m][j] += kp.weight * (2.0f * __expf(-0.5f * d * d) - 1.0f);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) sv[m] = srow[t + P::TPR * m];
#pragma unroll
        for (int m = 0; m < 8; ++m)
            ns[m] = make_float2(fminf(fmaxf(sv[m].x + a.dt * acc[m][0], 0.0f), 1.0f),
                                fminf(fmaxf(sv[m].y + a.dt * acc[m][1], 0.0f), 1.0f));
    } else {
        constexpr int NK = RULE == AH_RULE_LENIA ? 1 : 2;
        float2 v[NK][8];  // potentials of the thread's cells, per kernel
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            if (k > 0) lds_sync();  // every thread has read kernel k-1's last-stage inputs
            load_twist_rows<M>(lds, a, a.spec + (NK == 1 ? ch : k) * kspec, y0);
            int Ls = 1;
#pragma unroll
            for (int s = 0; s < P::NLDS; ++s, Ls *= 8) stage<8, 8, 1>(lds, M, P::ROWS, Ls, a.twW, W);
            if (k == NK - 1) {
#pragma unroll
                for (int m = 0; m < 8; ++m) sv[m] = srow[t + P::TPR * m];
            }
#pragma unroll
            for (int i = 0; i < P::NBL; ++i) {
                float2 q[P::RL];
                const int b = t + P::TPR * i;
#pragma unroll
                for (int r = 0; r < P::RL; ++r) q[r] = lds[pad(off + b + r * P::L)];
                float2 w[P::RL];
                twiddles<P::RL, 1>(w, a.twW, b, W / M, W);
#pragma unroll
                for (int r = 1; r < P::RL; ++r) q[r] = cmul(q[r], w[r]);
                dft<P::RL, 1>(q);
#pragma unroll
                for (int r = 0; r < P::RL; ++r) v[k][i + P::NBL * r] = q[r];
            }
        }
        const KernelParams kp = DIAG ? a.kp_dev[ch] : a.kp[0];
        const float dt = DIAG ? a.dt_dev[ch] : a.dt;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            float u[2] = {v[0][m].x, v[0][m].y}, s2[2] = {sv[m].x, sv[m].y};
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if constexpr (RULE == AH_RULE_LENIA) {
                    float d = (u[j] - kp.mu) * kp.inv_sigma;
                    float acc = 0.0f;
                    acc += kp.weight * (2.0f * __expf(-0.5f * d * d) - 1.0f);
                    s2[j] = fminf(fmaxf(s2[j] + dt * acc, 0.0f), 1.0f);
                } else {
                    const float u1 = j == 0 ? v[NK - 1][m].x : v[NK - 1][m].y;
                    if constexpr (RULE == AH_RULE_SMOOTHLIFE) s2[j] = smoothlife_next(s2[j], u[j], u1, dt);
                    else s2[j] = mnca_next(s2[j], u[j], u1);
                }
            }
            ns[m] = make_float2(s2[0], s2[1]);
        }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) orow[t + P::TPR * m] = ns[m];
    lds_sync();  // every thread has read its last-stage inputs
    // First forward stage (radix 8, span 1) on the registers: ns[m] is
    // position t + TPR m, its input r = m.
    dft8<-1>(ns);
#pragma unroll
    for (int r = 0; r < 8; ++r) lds[pad(off + 8 * t + r)] = ns[r];
    lds_sync();
    int Ls = 8;
#pragma unroll
    for (int s = 1; s < P::N8; ++s, Ls *= 8) stage<8, 8, -1>(lds, M, P::ROWS, Ls, a.twW, W);
    if constexpr (P::RLAST == 4) stage<4, 8, -1>(lds, M, P::ROWS, Ls, a.twW, W);
    if constexpr (P::RLAST == 2) stage<2, 8, -1>(lds, M, P::ROWS, Ls, a.twW, W);
    store_half_spectrum(lds, M, P::ROWS, H, y0, a.twW, a.spec_next + (size_t)ch * M * H, P::THREADS);
}

template <typename K>
void set_lds(K kernel, size_t bytes) {
    allow_dynamic_lds(reinterpret_cast<const void*>(kernel), bytes);
}

// Rows per row-pass workgroup: one row of a wide world, more rows of a
// narrow one so a workgroup has at least 256 threads (8 elements each).
int row_rows(int W, int H) {
    int M = W / 2;
    int rows = 1;
    while (rows * M < 2048 && rows < H) rows *= 2;
    return rows;
}

}  // namespace

// H = 4096 or 8192 outside batches: the block column kernels, any channels
// and kernels.
bool col_block_ok(int W, int H) { return (H == c4k::H || H == c8k::H) && W >= 64; }

int khat_layout(int W, int H, int channels, int n_kernels, const int* ksrc) {
    (void)channels;
    (void)n_kernels;
    (void)ksrc;
    return col_block_ok(W, H) ? KHAT_DIGITS : KHAT_BLOCK;
}

size_t khat_floats(int W, int H, int layout) {
    (void)layout;  // every layout holds p in [0, H/2] for each of the W/2 columns
    return (size_t)(W / 2) * (H / 2 + 1);
}

// Compute units of the current device (cached per device; sims on several
// host threads may ask at once).
int cu_count() {
    static std::atomic<int> cached[64] = {};
    int dev = 0;
    AH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error{AH_ERR_DEVICE, "device index out of range"};
    int n = cached[dev].load(std::memory_order_relaxed);
    if (!n) {
        AH_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        cached[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Dynamic LDS above 64 KiB must be allowed per kernel: once per (device,
// kernel symbol), up to the 160 KiB a gfx950 workgroup can hold, before the
// kernel's first launch on that device (so never first inside a stream
// capture).  Guarded by a mutex: sims on different host threads launch
// concurrently.
void allow_dynamic_lds(const void* kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    static std::mutex mu;
    static std::vector<std::pair<int, const void*>> done;
    int dev = 0;
    AH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (std::find(done.begin(), done.end(), std::make_pair(dev, kernel)) != done.end()) return;
    AH_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    done.emplace_back(dev, kernel);
}

namespace {

// Columns per column-pass workgroup (1, 2 or 4 of a 4-column block).
int col_cols(int H) {
    int cw = 1;
    while (cw * H < 2048 && cw < BC) cw *= 2;
    return cw;
}

}  // namespace

size_t fft_lds_limit() { return 160 * 1024; }

void launch_row_r2c(hipStream_t s, Profiler* p, int W, int H, int channels, const float* state, float2* spec,
                    const float2* twW) {
    const int M = W / 2, rows = row_rows(W, H);
    const int elems = rows * M;
    const size_t lds = padded(elems) * sizeof(float2);
    Launch L(p, s, "row_r2c");
    set_lds(k_row_r2c<8>, lds);
    hipLaunchKernelGGL(k_row_r2c<8>, dim3(H / rows, channels), dim3(elems / 8), lds, s, W, H, rows, state, spec, twW);
    AH_HIP(hipGetLastError());
}

bool batch_col_fused_ok(int W, int H) { return H >= 64 && H <= 2048 && W >= 16; }
bool col_blk_ok(int W, int H) { return H >= 256 && H <= 2048 && W >= 16; }
bool batch_row_fused_ok(int W, int H) {
    return W >= 128 && W <= 8192 && row_rows(W, H) * W == std::max(4096, W);
}

void launch_col_conv(hipStream_t s, Profiler* p, int W, int H, int channels, const float2* spec_in,
                     float2* spec_out, const float* khat, const float2* kpack, const float2* twH,
                     const int* ksrc, int n_kernels, bool diag, int variant) {
    ColArgs a;
    a.W = W;
    a.H = H;
    a.cw = col_cols(H);
    a.spec_in = spec_in;
    a.spec_out = spec_out;
    a.khat = khat;
    a.kpack = kpack;
    a.twH = twH;
    a.diag = diag ? 1 : 0;
    a.n_kernels = diag ? channels : n_kernels;
    for (int k = 0; k < AH_MAX_KERNELS; ++k) a.ksrc[k] = (!diag && k < n_kernels) ? ksrc[k] : -1;
    if (!diag && col_block_ok(W, H) && H == c8k::H) {
        // Two columns per workgroup (139 KiB of LDS, one workgroup per CU):
        // at 8192^2, 183 us against 213 us with one column (70 KiB, two
        // workgroups per CU) and 346 us for the LDS-staged k_col_conv<16>.
        using P = c8k::Plan<2>;
        Launch L(p, s, "col_conv");
        set_lds(k_col_block8192<2>, P::LDS_BYTES);
        hipLaunchKernelGGL(k_col_block8192<2>, dim3(W / 4, n_kernels), dim3(P::THREADS), P::LDS_BYTES, s, a);
        AH_HIP(hipGetLastError());
        return;
    }
    if (!diag && col_block_ok(W, H)) {
        // Two columns per workgroup (68 KiB of LDS): two workgroups share a
        // CU, so one's loads overlap the other's arithmetic.  Grid y is the
        // kernel: with several kernels per channel (SmoothLife, MNCA) each
        // transforms the channel itself (measured against one workgroup
        // running both kernels from one transform in LDS: SmoothLife 4096^2
        // column pass 77.3 -> 64.1 us).
        using P = c4k::Plan<2>;
        Launch L(p, s, "col_conv");
        set_lds(k_col_block4096<2>, P::LDS_BYTES);
        hipLaunchKernelGGL(k_col_block4096<2>, dim3(W / 4, n_kernels), dim3(P::THREADS), P::LDS_BYTES, s, a);
        AH_HIP(hipGetLastError());
        return;
    }
    const int elems = a.cw * H;
    const size_t lds = padded(elems) * sizeof(float2);
    Launch L(p, s, "col_conv");
    set_lds(k_col_conv<8>, lds);
    if (col_blk_ok(W, H) && (!diag || variant == COL_BLOCK)) {
        // Four columns (one block) per workgroup; grid y is the world in
        // diagonal mode, else the kernel.  Measured per column pass against
        // the kernels below: multichannel 1024^2 (9 kernels) 28.3 -> 18.4 us,
        // one kernel at 256^2 7.6 -> 5.4, 1024^2 10.8 -> 8.4, 2048^2
        // 19.3 -> 13.5 us, SmoothLife 2048^2 28.5 -> 18.7 us; two columns per
        // workgroup came within 0.3 us, one column per workgroup behind.
        const int gy = diag ? channels : n_kernels;
        auto block = [&](auto kernel, int threads, size_t bytes) {
            set_lds(kernel, bytes);
            hipLaunchKernelGGL(kernel, dim3(W / 8, gy), dim3(threads), bytes, s, a);
        };
        using namespace cblk;
        switch (H) {
            case 256: block(k_col_blk<256, 4>, Plan<256, 4>::THREADS, Plan<256, 4>::LDS_BYTES); break;
            case 512: block(k_col_blk<512, 4>, Plan<512, 4>::THREADS, Plan<512, 4>::LDS_BYTES); break;
            case 1024: block(k_col_blk<1024, 4>, Plan<1024, 4>::THREADS, Plan<1024, 4>::LDS_BYTES); break;
            default: block(k_col_blk<2048, 4>, Plan<2048, 4>::THREADS, Plan<2048, 4>::LDS_BYTES); break;
        }
        AH_HIP(hipGetLastError());
        return;
    }
    if (diag && variant == COL_FUSED && batch_col_fused_ok(W, H)) {
        auto fused = [&](auto kernel, int threads, int cpw) {
            const size_t flds = padded((size_t)cpw * H) * sizeof(float2);
            set_lds(kernel, flds);
            hipLaunchKernelGGL(kernel, dim3((W / 8) * (4 / cpw), channels), dim3(threads), flds, s, a);
        };
        switch (H) {
            case 64: fused(k_col_fused<64>, ColPlan<64>::THREADS, ColPlan<64>::CPW); break;
            case 128: fused(k_col_fused<128>, ColPlan<128>::THREADS, ColPlan<128>::CPW); break;
            case 256: fused(k_col_fused<256>, ColPlan<256>::THREADS, ColPlan<256>::CPW); break;
            case 512: fused(k_col_fused<512>, ColPlan<512>::THREADS, ColPlan<512>::CPW); break;
            case 1024: fused(k_col_fused<1024>, ColPlan<1024>::THREADS, ColPlan<1024>::CPW); break;
            default: fused(k_col_fused<2048>, ColPlan<2048>::THREADS, ColPlan<2048>::CPW); break;
        }
        AH_HIP(hipGetLastError());
        return;
    }
    if (!diag && elems / 4 <= FFT_THREADS && (W / 2 / a.cw) * channels < 4 * cu_count()) {
        // A grid of under four workgroups per CU: four elements per thread
        // (radix-4 stages), so twice the waves hide the latency of each
        // workgroup's chain.  Measured per column pass: multichannel 1024^2
        // 37.0 -> 33.6 us, 512^2 25.4 -> 23.8 us, SmoothLife 1024^2 23.0 ->
        // 20.0 us, one kernel at 1024^2 16.1 -> 13.9 us and 256^2 13.5 ->
        // 11.4 us; SmoothLife 2048^2 (1024 workgroups) 32.6 -> 38.1 us, so
        // larger grids keep radix 8.
        set_lds(k_col_conv<4>, lds);
        hipLaunchKernelGGL(k_col_conv<4>, dim3(W / 2 / a.cw, channels), dim3(elems / 4), lds, s, a);
    } else if (elems / 8 > FFT_THREADS) {
        // H = 8192: one 66 KiB column per workgroup, sixteen elements per
        // thread (two radix-8 butterflies per thread and stage).
        set_lds(k_col_conv<16>, lds);
        hipLaunchKernelGGL(k_col_conv<16>, dim3(W / 2 / a.cw, channels), dim3(elems / 16), lds, s, a);
    } else {
        hipLaunchKernelGGL(k_col_conv<8>, dim3(W / 2 / a.cw, channels), dim3(elems / 8), lds, s, a);
    }
    AH_HIP(hipGetLastError());
}

void launch_row_c2r(hipStream_t s, Profiler* p, int W, int H, int channels, int n_kernels, const float2* spec,
                    const float* state_in, float* state_out, const KernelParams* kp, float dt, const float2* twW,
                    float* conv_out, int conv_k, float2* spec_next, int rule,
                    const KernelParams* kp_dev, const float* dt_dev, bool generic) {
    RowArgs a;
    a.W = W;
    a.H = H;
    a.rows = row_rows(W, H);
    a.spec = spec;
    a.state_in = state_in;
    a.state_out = state_out;
    a.twW = twW;
    a.dt = dt;
    a.kp_dev = kp_dev;
    a.dt_dev = dt_dev;
    a.n_kernels = kp_dev ? channels : n_kernels;
    if (!kp_dev)
        for (int k = 0; k < n_kernels; ++k) a.kp[k] = kp[k];
    a.conv_out = conv_out;
    a.conv_k = conv_k;
    a.spec_next = conv_out ? nullptr : spec_next;
    const int elems = a.rows * (W / 2);
    const size_t lds = padded(elems) * sizeof(float2);
    const int gy = conv_out ? 1 : channels;
    const dim3 grid(H / a.rows, gy), block(elems / 8);
    if (!conv_out && rule != AH_RULE_LENIA && !a.spec_next)
        throw Error{AH_ERR_INVALID, "SmoothLife and MNCA steps always write the next spectra"};
    auto go = [&](auto kernel, const char* name) {
        Launch L(p, s, name);
        set_lds(kernel, lds);
        hipLaunchKernelGGL(kernel, grid, block, lds, s, a);
    };
    const bool fusable = !generic && batch_row_fused_ok(W, H) && !conv_out && a.spec_next;
    auto fused = [&](auto diag_tag, auto rule_tag, auto multi_tag) {
        constexpr bool D = decltype(diag_tag)::value;
        constexpr int R = decltype(rule_tag)::value;
        constexpr bool MU = decltype(multi_tag)::value;
        Launch L(p, s, "row_c2r_update_r2c");
        auto go_fused = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, dim3(elems / 8), lds, s, a); };
        switch (W) {
            case 128: go_fused(k_row_fused<64, D, R, MU>); break;
            case 256: go_fused(k_row_fused<128, D, R, MU>); break;
            case 512: go_fused(k_row_fused<256, D, R, MU>); break;
            case 1024: go_fused(k_row_fused<512, D, R, MU>); break;
            case 2048: go_fused(k_row_fused<1024, D, R, MU>); break;
            case 4096: go_fused(k_row_fused<2048, D, R, MU>); break;
            default: go_fused(k_row_fused<4096, D, R, MU>); break;
        }
    };
    const std::false_type no;
    using Lenia = std::integral_constant<int, AH_RULE_LENIA>;
    using Smooth = std::integral_constant<int, AH_RULE_SMOOTHLIFE>;
    using Mnca = std::integral_constant<int, AH_RULE_MNCA>;
    const bool two_kernels = !kp_dev && channels == 1 && n_kernels == 2;
    if (conv_out) go(k_row_c2r<8, false, AH_RULE_LENIA>, "row_c2r_conv");
    else if (!a.spec_next) go(k_row_c2r<8, false, AH_RULE_LENIA>, "row_c2r_update");
    else if (rule == AH_RULE_SMOOTHLIFE) {
        if (fusable && two_kernels) fused(no, Smooth{}, no);
        else go(k_row_c2r<8, true, AH_RULE_SMOOTHLIFE>, "row_c2r_update_r2c");
    } else if (rule == AH_RULE_MNCA) {
        if (fusable && two_kernels) fused(no, Mnca{}, no);
        else go(k_row_c2r<8, true, AH_RULE_MNCA>, "row_c2r_update_r2c");
    } else if (fusable && kp_dev) fused(std::true_type{}, Lenia{}, no);
    else if (fusable && channels == 1 && n_kernels == 1 && kp[0].dst == 0) fused(no, Lenia{}, no);
    else if (fusable) fused(no, Lenia{}, std::true_type{});
    else go(k_row_c2r<8, true, AH_RULE_LENIA>, "row_c2r_update_r2c");
    AH_HIP(hipGetLastError());
}

}  // namespace ah
