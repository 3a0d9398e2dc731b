// Direct periodic convolution on BF16 MFMA (FP32 accumulate), fused with the
// growth map, dt and clip.  This replaces the reference's whole FFT pass chain
// (crates/auto-core/src/automata/spectral.rs:199-264) for kernels of radius
// <= 31 with one launch per step.
//
// Formulation.  For one kernel row dy the x-convolution of a strip of 16
// output columns is a banded (Toeplitz) matrix product:
//     u[y][x0+i] += sum_j A_dy[i][j] * S[y+dy][xw+j],  A_dy[i][j] = K[dy][xw+j-x0-i]
// v_mfma_f32_16x16x32_bf16 computes D(16x16) += A(16x32) B(32x16) with
//   A rows = the 16 output columns i, A cols = 32 consecutive input columns j,
//   B rows = the same 32 input columns, B cols = 16 output rows y,
// so every lane's B operand is 8 consecutive bf16 of one state row (one
// aligned 16-byte LDS read) and its A operand is a precomputed fragment of the
// kernel row.  Each kernel row only spans its chord of the disc, |dx| <= dxm(dy),
// so the number of 32-column chunks varies per row (3 near the centre of an
// R=26 kernel, 2 further out, 1 at the rim).  K[dy] = K[-dy] halves the
// fragment table.
//
// Operand reuse.  The window of strip s+2 is the window of strip s moved by
// exactly one 32-column chunk, so a wave owns the S strips {p, p+2, ...} of
// one parity p and NB 16-row blocks: for a kernel row with c chunks it loads
// c A fragments and (c+S-1)*NB B fragments for S*c*NB MFMAs (S=4, NB=2,
// c=3: 15 KiB of LDS reads for 24 MFMAs, against 2 KiB per MFMA without
// any reuse).
//
// Lane layouts assumed for the 16x16x32 bf16 MFMA (wave64, lane l, q = l/16):
//   A: row l%16, k = 8q..8q+7     B: col l%16, k = 8q..8q+7
//   D: col l%16, rows 4q..4q+3
//
// The state tile is staged in LDS as bf16; kernel taps are rounded to bf16
// and the accumulated sum is rescaled by 1/sum(bf16 taps) (an FP64 sum done
// on the host), so the rounded kernel still has unit mass.
#include <algorithm>

#include "ah_internal.h"

namespace ah {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int DTHREADS = 512;   // 8 waves: strip parity = wave & 1, row group = wave >> 1
// Tile: DX = 32*S columns (2S strips of 16), DY = 64*NB rows (4 groups of NB blocks).

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// Two floats as packed bf16 (a in the low half), rounded to nearest even by
// one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
// Four floats as bf16 hi terms, or (lo) the bf16 of what the hi terms leave.
__device__ __forceinline__ uint2 pack4(const float (&e)[4], bool lo) {
    uint2 h = make_uint2(pack2(e[0], e[1]), pack2(e[2], e[3]));
    if (lo)
        h = make_uint2(pack2(e[0] - __uint_as_float(h.x << 16), e[1] - __uint_as_float(h.x & 0xffff0000u)),
                       pack2(e[2] - __uint_as_float(h.y << 16), e[3] - __uint_as_float(h.y & 0xffff0000u)));
    return h;
}

// Stage rows [y0-hy, y0+DY+hy) x cols [x0-hx, x0-hx+tc) of one channel as
// bf16, wrapped on the torus: the value's bf16 rounding (hi), or with lo set
// the bf16 rounding of what hi leaves over, x - hi.  Columns outside
// [x0-hy, x0+DX+hy) are zero: no nonzero tap reaches them, and zeros keep
// non-finite data out of range.
//
// Wave w stages rows w, w+8, ...; lane l covers columns 4l..4l+3 (and +256,
// ... for wider tiles).  A lane issues the loads of STAGE_BATCH rows before
// converting any of them, so a tile's halo arrives with up to STAGE_BATCH
// loads in flight per lane instead of one.  Row wrap is a single +-H: the
// halo (<= 63 rows) never exceeds the world height (>= 64).
constexpr int STAGE_BATCH = 16;

__device__ __forceinline__ void stage_tile(uint16_t* tile, const float* src, int W, int H, int x0, int y0, int DX,
                                           int halo_y, int halo_x, int tr, int tc, int tstride, bool lo) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 4 * lane; j < tc; j += 256) {
        const int cx = j - halo_x;  // column relative to x0
        int gx = (x0 + cx) % W;
        if (gx < 0) gx += W;
        for (int r0 = wave; r0 < tr; r0 += 8 * STAGE_BATCH) {
            float4 v[STAGE_BATCH];
#pragma unroll
            for (int u = 0; u < STAGE_BATCH; ++u) {
                const int r = r0 + 8 * u;
                if (r < tr) {
                    int gy = y0 - halo_y + r;
                    if (gy < 0) gy += H;
                    else if (gy >= H) gy -= H;
                    v[u] = *reinterpret_cast<const float4*>(src + (size_t)gy * W + gx);
                }
            }
#pragma unroll
            for (int u = 0; u < STAGE_BATCH; ++u) {
                const int r = r0 + 8 * u;
                if (r < tr) {
                    float e[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int c = cx + t;
                        if (!(c >= -halo_y && c < DX + halo_y)) e[t] = 0.0f;  // bf16 0 for hi and lo
                    }
                    *reinterpret_cast<uint2*>(tile + (size_t)r * tstride + j) = pack4(e, lo);
                }
            }
        }
    }
}

// One kernel row's operands: NCH A fragments and, for every chunk u of the
// S strips' joint window, NB B fragments.
template <int S, int NB, int NCH>
struct RowFrags {
    bf16x8 A[NCH];
    bf16x8 B[NCH + S - 1][NB];
};

template <int S, int NB, int NCH>
__device__ __forceinline__ void load_row(RowFrags<S, NB, NCH>& f, const uint16_t* brow, const uint4* afrag, int lane,
                                         int tstride) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) f.A[c] = __builtin_bit_cast(bf16x8, afrag[c * 64 + lane]);
#pragma unroll
    for (int u = 0; u < NCH + S - 1; ++u)
#pragma unroll
        for (int b = 0; b < NB; ++b)
            f.B[u][b] =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(brow + (size_t)16 * b * tstride + 32 * u));
}

// The MFMAs of one row: B fragment u feeds the strips k whose own chunk
// u - k exists.  Per accumulator the chunks run in increasing order.
template <int S, int NB, int NCH>
__device__ __forceinline__ void mfma_row(f32x4 (&acc)[S][NB], const RowFrags<S, NB, NCH>& f) {
#pragma unroll
    for (int u = 0; u < NCH + S - 1; ++u)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const int c = u - k;
                if (c >= 0 && c < NCH)
                    acc[k][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(f.A[c], f.B[u][b], acc[k][b], 0, 0, 0);
            }
}

// The rows [i0, i1) of the row list, which all have NCH chunks.  Software
// pipelined over two register sets: the LDS reads of row i+1 are issued
// before the MFMAs of row i.  Lane i of `ent` holds row i's entry
// (dy, e, first fragment), read out with readlane, so no row waits on LDS
// for its own entry.
template <int S, int NB, int NCH>
__device__ __forceinline__ void conv_rows(f32x4 (&acc)[S][NB], const uint16_t* tile, const uint4* frag_lds,
                                          const int4& ent, int i0, int i1, int base, int ring_rows, int colP,
                                          int tstride, int lane) {
    if (i0 >= i1) return;
    auto ptrs = [&](int i, const uint16_t*& brow, const uint4*& afrag) {
        const int dy = __builtin_amdgcn_readlane(ent.x, i);
        const int e = __builtin_amdgcn_readlane(ent.y, i);
        const int fr = __builtin_amdgcn_readlane(ent.z, i);
        int slot = base + dy;
        if (slot < 0) slot += ring_rows;
        else if (slot >= ring_rows) slot -= ring_rows;
        brow = tile + (size_t)slot * tstride + colP + e;
        afrag = frag_lds + fr * 64;
    };
    RowFrags<S, NB, NCH> f0, f1;
    const uint16_t* brow;
    const uint4* afrag;
    ptrs(i0, brow, afrag);
    load_row(f0, brow, afrag, lane, tstride);
    for (int i = i0; i < i1; i += 2) {
        // f0 holds row i.
        if (i + 1 < i1) {
            ptrs(i + 1, brow, afrag);
            load_row(f1, brow, afrag, lane, tstride);
        }
        mfma_row(acc, f0);
        if (i + 1 < i1) {
            if (i + 2 < i1) {
                ptrs(i + 2, brow, afrag);
                load_row(f0, brow, afrag, lane, tstride);
            }
            mfma_row(acc, f1);
        }
    }
}

// The rows [i0, i1) of the row list (NCH chunks each) for NB = 1, where the
// list holds kernel rows dy and -dy next to each other (row 0 first when the
// group has it): the two rows of a pair share their A fragments, so a pair
// reads NCH A fragments instead of 2 NCH and runs 2 S NCH MFMAs between
// loads.  The MFMAs per accumulator run in the order of conv_rows (row by
// row, chunks in increasing order), so both give the same bits.
template <int S, int NCH>
struct PairFrags {
    bf16x8 A[NCH];
    bf16x8 B[2][NCH + S - 1];
};

template <int S, int NCH>
__device__ __forceinline__ void mfma_pair(f32x4 (&acc)[S][1], const PairFrags<S, NCH>& f) {
    // Raised issue priority while this wave's MFMAs run: with two waves per
    // SIMD, the other wave's LDS reads and VALU work fill the gaps between
    // them instead of delaying them.
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int u = 0; u < NCH + S - 1; ++u)
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const int c = u - k;
                if (c >= 0 && c < NCH)
                    acc[k][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(f.A[c], f.B[h][u], acc[k][0], 0, 0, 0);
            }
    __builtin_amdgcn_s_setprio(0);
}

template <int S, int NCH>
__device__ __forceinline__ void conv_row_pairs(f32x4 (&acc)[S][1], const uint16_t* tile, const uint4* frag_lds,
                                               const int4& ent, int i0, int i1, int base, int ring_rows, int colP,
                                               int tstride, int lane) {
    if ((i1 - i0) & 1) {  // kernel row 0, unpaired
        conv_rows<S, 1, NCH>(acc, tile, frag_lds, ent, i0, i0 + 1, base, ring_rows, colP, tstride, lane);
        ++i0;
    }
    if (i0 >= i1) return;
    auto brow_of = [&](int i) {
        const int dy = __builtin_amdgcn_readlane(ent.x, i);
        int slot = base + dy;
        if (slot < 0) slot += ring_rows;
        else if (slot >= ring_rows) slot -= ring_rows;
        return tile + (size_t)slot * tstride + colP;
    };
    auto load = [&](PairFrags<S, NCH>& f, int i) {
        const int e = __builtin_amdgcn_readlane(ent.y, i);
        const int fr = __builtin_amdgcn_readlane(ent.z, i);
        const uint16_t* b0 = brow_of(i) + e;
        const uint16_t* b1 = brow_of(i + 1) + e;
        const uint4* afrag = frag_lds + fr * 64;
#pragma unroll
        for (int c = 0; c < NCH; ++c) f.A[c] = __builtin_bit_cast(bf16x8, afrag[c * 64 + lane]);
#pragma unroll
        for (int u = 0; u < NCH + S - 1; ++u) {
            f.B[0][u] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(b0 + 32 * u));
            f.B[1][u] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(b1 + 32 * u));
        }
    };
    PairFrags<S, NCH> f0, f1;
    load(f0, i0);
    for (int i = i0; i < i1; i += 4) {
        // f0 holds the pair at i.
        if (i + 2 < i1) load(f1, i + 2);
        mfma_pair(acc, f0);
        if (i + 2 < i1) {
            if (i + 4 < i1) load(f0, i + 4);
            mfma_pair(acc, f1);
        }
    }
}

// All MFMAs of one kernel term over this wave's S strips x NB blocks,
// accumulated into acc, row group by row group (3, 2, then 1 chunks).
// Tile row r of the lane sits in LDS row (vbase + r) mod ring_rows: the
// plain tile passes vbase 0 and a ring_rows no row reaches; the ring kernel
// (NB = 1) passes its ring.
template <int S, int NB, bool PAIRS = false>
__device__ __forceinline__ void conv_tile(f32x4 (&acc)[S][NB], const uint16_t* tile, const uint4* frag_lds,
                                          const int4* rows, const int (&nrow)[3], int halo_y, int halo_x,
                                          int tstride, int ring_rows, int vbase) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int par = wave & 1, grp = wave >> 1;
    const int q = lane >> 4, n = lane & 15;
    const int base = (vbase + grp * 16 * NB + n + halo_y) % ring_rows;  // LDS row of dy = 0
    const int colP = 16 * par + halo_x + 8 * q;
    const int a3 = nrow[0], a2 = a3 + nrow[1], a1 = a2 + nrow[2];
    const int4 ent = lane < a1 ? rows[lane] : int4{0, 0, 0, 0};  // at most 63 rows
    if constexpr (PAIRS && NB == 1) {
        conv_row_pairs<S, 3>(acc, tile, frag_lds, ent, 0, a3, base, ring_rows, colP, tstride, lane);
        conv_row_pairs<S, 2>(acc, tile, frag_lds, ent, a3, a2, base, ring_rows, colP, tstride, lane);
        conv_row_pairs<S, 1>(acc, tile, frag_lds, ent, a2, a1, base, ring_rows, colP, tstride, lane);
    } else {
        conv_rows<S, NB, 3>(acc, tile, frag_lds, ent, 0, a3, base, ring_rows, colP, tstride, lane);
        conv_rows<S, NB, 2>(acc, tile, frag_lds, ent, a3, a2, base, ring_rows, colP, tstride, lane);
        conv_rows<S, NB, 1>(acc, tile, frag_lds, ent, a2, a1, base, ring_rows, colP, tstride, lane);
    }
}

// One kernel's potential over the tile: K_hi s_hi, plus K_lo s_hi and
// K_hi s_lo for three terms.  `staged` tracks which (source, hi/lo) tile sits
// in LDS (source*2 + lo) so that consecutive kernels on one source skip the
// restaging; `loaded` tracks the fragment block in LDS the same way.
template <int S, int NB>
__device__ __forceinline__ void kernel_potential(f32x4 (&acc)[S][NB], uint16_t* tile, uint4* frag_lds,
                                                 int4* rows_lds, int& staged, int& loaded,
                                                 const float* src_plane, int src, int W, int H, int x0, int y0,
                                                 int halo_y, int halo_x, int tr, int tc, int tstride,
                                                 const uint4* __restrict__ frags, int frag0, int nfrag, int terms,
                                                 const int4* rows, const int (&nrow)[3]) {
#pragma unroll
    for (int k = 0; k < S; ++k)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[k][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // (tile lo?, fragment block) per term
    const int nterm = terms == 3 ? 3 : 1;
    for (int term = 0; term < nterm; ++term) {
        const bool lo_tile = term == 2;
        const int fbase = frag0 + (term == 1 ? nfrag : 0);
        __syncthreads();  // earlier reads of tile / fragments are done
        const int want = src * 2 + (lo_tile ? 1 : 0);
        const bool new_frags = loaded != fbase;
        if (new_frags) {
            // Fragment table by DMA (16 bytes per lane, one wave instruction
            // per 64 entries), in flight with the tile's loads; the barrier
            // below waits for it.
            const uint4* fsrc = frags + (size_t)fbase * 64;
            const int w0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
            for (int i0 = w0 * 64; i0 < nfrag * 64; i0 += DTHREADS)
                __builtin_amdgcn_global_load_lds(fsrc + i0 + (threadIdx.x & 63),
                                                 (__attribute__((address_space(3))) void*)(frag_lds + i0), 16, 0, 0);
        }
        if (staged != want) {
            stage_tile(tile, src_plane, W, H, x0, y0, 32 * S, halo_y, halo_x, tr, tc, tstride, lo_tile);
            staged = want;
        }
        if (new_frags) {
            const int nr = nrow[0] + nrow[1] + nrow[2];
            for (int i = threadIdx.x; i < nr; i += DTHREADS) rows_lds[i] = rows[i];
            loaded = fbase;
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the fragment DMA
        __syncthreads();
        conv_tile<S, NB>(acc, tile, frag_lds, rows_lds, nrow, halo_y, halo_x, tstride, tr + 1, 0);
    }
}

// Epilogue arithmetic shared by the tile, ring and batch kernels.  Each
// rounding is spelled out (no contraction into FMAs), so the compiler cannot
// fuse differently in different kernels: the same potential gives the same
// bits in every variant.
__device__ __forceinline__ float potential(float acc, float scale) {
#pragma clang fp contract(off)
    return acc * scale;
}

__device__ __forceinline__ float growth(float u, float mu, float inv_sigma) {
#pragma clang fp contract(off)
    const float d = (u - mu) * inv_sigma;
    return 2.0f * __expf(-0.5f * d * d) - 1.0f;
}

__device__ __forceinline__ float add_weighted(float g, float weight, float gr) {
#pragma clang fp contract(off)
    return g + weight * gr;
}

__device__ __forceinline__ float lenia_update(float r, float dt, float g) {
#pragma clang fp contract(off)
    return fminf(fmaxf(r + dt * g, 0.0f), 1.0f);
}

// Output (strip k, block b, element r) of this lane: column
// x0 + 16*(par + 2k) + 4q + r, row y0 + grp*16*NB + 16b + n.
template <int S, int NB>
__global__ void __launch_bounds__(DTHREADS) k_direct(DirectArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_u16[];
    constexpr int DX = 32 * S, DY = 64 * NB;
    const int W = a.W, H = a.H;
    const int x0 = blockIdx.x * DX, y0 = blockIdx.y * DY;
    const int dst = blockIdx.z;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int par = wave & 1, grp = wave >> 1;
    const int q = lane >> 4, n = lane & 15;
    uint16_t* tile = lds_u16;
    uint4* frag_lds = reinterpret_cast<uint4*>(lds_u16 + (size_t)a.tr * a.tstride);
    int4* rows_lds = reinterpret_cast<int4*>(frag_lds + (size_t)a.nfrag_max * 64);

    float g[S][NB][4];
#pragma unroll
    for (int k = 0; k < S; ++k)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) g[k][b][r] = 0.0f;

    int staged = -1, loaded = -1;
    for (int kn = 0; kn < a.n_kernels; ++kn) {
        const DirectKernel kk = a.k[kn];
        if (a.conv_out ? (kn != a.conv_k) : (kk.dst != dst)) continue;
        f32x4 acc[S][NB];
        kernel_potential<S, NB>(acc, tile, frag_lds, rows_lds, staged, loaded, a.state_in + (size_t)kk.src * W * H,
                                kk.src, W, H, x0, y0, a.halo_y, a.halo_x, a.tr, a.tc, a.tstride, a.frags, kk.frag0,
                                kk.nfrag, kk.terms, a.sched + a.sched_base[kn], kk.nrow);
#pragma unroll
        for (int k = 0; k < S; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float u = potential(acc[k][b][r], kk.scale);
                    if (a.conv_out) g[k][b][r] = u;
                    else g[k][b][r] = add_weighted(g[k][b][r], kk.weight, growth(u, kk.mu, kk.inv_sigma));
                }
    }

#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int xo = x0 + 16 * (par + 2 * k) + 4 * q;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int yo = y0 + grp * 16 * NB + 16 * b + n;
            const size_t o = (size_t)yo * W + xo;
            if (a.conv_out) {
                *reinterpret_cast<float4*>(a.conv_out + o) = make_float4(g[k][b][0], g[k][b][1], g[k][b][2], g[k][b][3]);
            } else {
                const size_t plane = (size_t)dst * W * H;
                float4 sv = *reinterpret_cast<const float4*>(a.state_in + plane + o);
                float r[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
                for (int t = 0; t < 4; ++t) r[t] = lenia_update(r[t], a.dt, g[k][b][t]);
                *reinterpret_cast<float4*>(a.state_out + plane + o) = make_float4(r[0], r[1], r[2], r[3]);
            }
        }
    }
}

// Variant 2: workgroup (tile, kernel kn) writes kernel kn's weighted growth,
// add_weighted(0, weight, growth(u)), to plane kn of a.gsum; k_direct_sum
// then adds the planes of each channel's kernels in kernel order, which is
// the sum k_direct forms in registers, rounding for rounding.  Several
// kernels per channel then fill the chip with n_kernels times the
// workgroups, each staging one source.
template <int S, int NB>
__global__ void __launch_bounds__(DTHREADS) k_direct_split(DirectArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_u16[];
    constexpr int DX = 32 * S, DY = 64 * NB;
    const int W = a.W, H = a.H;
    const int x0 = blockIdx.x * DX, y0 = blockIdx.y * DY;
    const int kn = a.split_order[blockIdx.z];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int par = wave & 1, grp = wave >> 1;
    const int q = lane >> 4, n = lane & 15;
    uint16_t* tile = lds_u16;
    uint4* frag_lds = reinterpret_cast<uint4*>(lds_u16 + (size_t)a.tr * a.tstride);
    int4* rows_lds = reinterpret_cast<int4*>(frag_lds + (size_t)a.nfrag_max * 64);
    const DirectKernel kk = a.k[kn];
    int staged = -1, loaded = -1;
    f32x4 acc[S][NB];
    kernel_potential<S, NB>(acc, tile, frag_lds, rows_lds, staged, loaded, a.state_in + (size_t)kk.src * W * H,
                            kk.src, W, H, x0, y0, a.halo_y, a.halo_x, a.tr, a.tc, a.tstride, a.frags, kk.frag0,
                            kk.nfrag, kk.terms, a.sched + a.sched_base[kn], kk.nrow);
    float* plane = a.gsum + (size_t)kn * W * H;
#pragma unroll
    for (int k = 0; k < S; ++k) {
        const int xo = x0 + 16 * (par + 2 * k) + 4 * q;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int yo = y0 + grp * 16 * NB + 16 * b + n;
            float g[4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
                g[r] = add_weighted(0.0f, kk.weight, growth(potential(acc[k][b][r], kk.scale), kk.mu, kk.inv_sigma));
            *reinterpret_cast<float4*>(plane + (size_t)yo * W + xo) = make_float4(g[0], g[1], g[2], g[3]);
        }
    }
}

// Sum pass of variant 2: thread i takes cells 4i..4i+3 of channel
// blockIdx.y.
__global__ void __launch_bounds__(256) k_direct_sum(DirectArgs a) {
    const size_t cells = (size_t)a.W * a.H, i = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= cells) return;
    const int c = blockIdx.y;
    float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int kn = 0; kn < a.n_kernels; ++kn) {
        if (a.k[kn].dst != c) continue;
        const float4 v = *reinterpret_cast<const float4*>(a.gsum + (size_t)kn * cells + i);
        g[0] = add_weighted(g[0], 1.0f, v.x);
        g[1] = add_weighted(g[1], 1.0f, v.y);
        g[2] = add_weighted(g[2], 1.0f, v.z);
        g[3] = add_weighted(g[3], 1.0f, v.w);
    }
    const size_t o = (size_t)c * cells + i;
    const float4 sv = *reinterpret_cast<const float4*>(a.state_in + o);
    *reinterpret_cast<float4*>(a.state_out + o) =
        make_float4(lenia_update(sv.x, a.dt, g[0]), lenia_update(sv.y, a.dt, g[1]), lenia_update(sv.z, a.dt, g[2]),
                    lenia_update(sv.w, a.dt, g[3]));
}

// Persistent variant for one channel, one kernel and one bf16 term.
// Workgroup b walks a run of vertically adjacent 64-row tiles of one
// 32S-column strip (the torus wraps a run past the last tile row).  Two
// consecutive tiles share 2*halo_y staged rows, so LDS holds a ring of
// tr + 64 rows: tile i of the run reads ring rows (64i + r) mod (tr + 64),
// r < tr, and while its MFMAs run the 64 rows tile i+1 adds are loaded into
// registers and then written to the ring rows tile i-1 used.  Only the run's
// first tile is staged before its MFMAs; the MFMA order per output is that
// of k_direct<S, 1>, so both give the same bits.
template <int S>
__global__ void __launch_bounds__(DTHREADS) k_direct_ring(DirectArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_u16[];
    constexpr int DX = 32 * S, DY = 64, NEW_PER_WAVE = DY / (DTHREADS / 64);
    const int W = a.W, H = a.H, RT = a.tr + DY;
    const int strips = W / DX, TY = H / DY;
    const int segs = gridDim.x / strips;
    const int strip = blockIdx.x / segs, seg = blockIdx.x - strip * segs;
    const int t0 = seg * TY / segs, t1 = (seg + 1) * TY / segs;
    const int x0 = strip * DX;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int par = wave & 1, grp = wave >> 1;
    const int q = lane >> 4, n = lane & 15;
    uint16_t* ring = lds_u16;
    uint4* frag_lds = reinterpret_cast<uint4*>(lds_u16 + (size_t)RT * a.tstride);
    int4* rows_lds = reinterpret_cast<int4*>(frag_lds + (size_t)a.nfrag_max * 64);
    const DirectKernel kk = a.k[0];
    const int4* rows = a.sched + a.sched_base[0];
    // The fragment table goes to LDS by DMA (16 bytes per lane, one wave
    // instruction per 64 entries), so its loads are in flight together with
    // the first tile's instead of one round trip ahead of them; the waits
    // for the tile's loads (in issue order) cover it before the first
    // barrier.
    {
        const uint4* src = a.frags + (size_t)kk.frag0 * 64;
        const int w0 = __builtin_amdgcn_readfirstlane(wave);
        for (int i0 = w0 * 64; i0 < kk.nfrag * 64; i0 += DTHREADS)
            __builtin_amdgcn_global_load_lds(src + i0 + lane, (__attribute__((address_space(3))) void*)(frag_lds + i0),
                                             16, 0, 0);
    }
    for (int i = tid; i < kk.nrow[0] + kk.nrow[1] + kk.nrow[2]; i += DTHREADS) rows_lds[i] = rows[i];
    const int ybase = t0 * DY;
    stage_tile(ring, a.state_in, W, H, x0, ybase, DX, a.halo_y, a.halo_x, a.tr, a.tc, a.tstride, false);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the fragment DMA, whatever the tile's waits covered

    // This lane's column of the streamed rows (the tile is at most 256 wide).
    const int jn = 4 * lane;
    const bool col_ok = jn < a.tc;
    const int cxn = jn - a.halo_x;
    int gxn = (x0 + cxn) % W;
    if (gxn < 0) gxn += W;

    for (int t = t0; t < t1; ++t) {
        const int i = t - t0;
        const bool more = t + 1 < t1;
        // Rows 64i + tr + wave + 8u (u < 8) of the run: the next tile's new rows.
        // Their world rows and ring slots step by 8 from u = 0, so each
        // takes one conditional subtraction (8u < 64 <= H, 8u < RT), not a
        // division; the row indexes are wave-uniform.
        const int v0 = DY * i + a.tr + __builtin_amdgcn_readfirstlane(wave);
        const int gy0 = (ybase - a.halo_y + v0) % H;  // the argument is positive (tr > halo_y)
        const int slot0 = v0 % RT;
        float4 nv[NEW_PER_WAVE];
        if (more && col_ok) {
#pragma unroll
            for (int u = 0; u < NEW_PER_WAVE; ++u) {
                int gy = gy0 + 8 * u;
                if (gy >= H) gy -= H;
                nv[u] = *reinterpret_cast<const float4*>(a.state_in + (size_t)gy * W + gxn);
            }
        }
        // The FP32 state this tile's epilogue updates, loaded now so that it
        // too lands during the MFMAs instead of after them.
        const int yo = t * DY + grp * 16 + n;
        float4 sv[S];
#pragma unroll
        for (int k = 0; k < S; ++k)
            sv[k] = *reinterpret_cast<const float4*>(a.state_in + (size_t)yo * W + x0 + 16 * (par + 2 * k) + 4 * q);
        // This tile's ring rows and the fragments are in LDS.  The barrier
        // does not wait for the loads just issued: they land during the MFMAs.
        lds_sync();
        f32x4 acc[S][1];
#pragma unroll
        for (int k = 0; k < S; ++k) acc[k][0] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        conv_tile<S, 1, true>(acc, ring, frag_lds, rows_lds, kk.nrow, a.halo_y, a.halo_x, a.tstride, RT, (DY * i) % RT);
        if (more && col_ok) {
#pragma unroll
            for (int u = 0; u < NEW_PER_WAVE; ++u) {
                int slot = slot0 + 8 * u;
                if (slot >= RT) slot -= RT;
                float e[4] = {nv[u].x, nv[u].y, nv[u].z, nv[u].w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int cx = cxn + c;
                    if (!(cx >= -a.halo_y && cx < DX + a.halo_y)) e[c] = 0.0f;
                }
                *reinterpret_cast<uint2*>(ring + (size_t)slot * a.tstride + jn) = pack4(e, false);
            }
        }
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int xo = x0 + 16 * (par + 2 * k) + 4 * q;
            const size_t o = (size_t)yo * W + xo;
            float r[4] = {sv[k].x, sv[k].y, sv[k].z, sv[k].w};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                r[c] = lenia_update(r[c], a.dt,
                                    add_weighted(0.0f, kk.weight, growth(potential(acc[k][0][c], kk.scale), kk.mu,
                                                                         kk.inv_sigma)));
            *reinterpret_cast<float4*>(a.state_out + o) = make_float4(r[0], r[1], r[2], r[3]);
        }
    }
}

// Persistent batch kernel: workgroup b walks tiles [seg[b], seg[b+1]) of a
// list grouped by world, cut into segments of near-equal MFMA cost; a tile
// names its world (size, offset, kernel fragments, growth, dt).  States ping-pong between buf[parity] and
// buf[parity^1]; world w's cells live at the same offset in both.
template <int S, int NB>
__global__ void __launch_bounds__(DTHREADS) k_direct_batch(BatchArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint16_t lds_u16[];
    uint16_t* tile = lds_u16;
    uint4* frag_lds = reinterpret_cast<uint4*>(lds_u16 + (size_t)a.tr * a.tstride);
    int4* rows_lds = reinterpret_cast<int4*>(frag_lds + (size_t)a.nfrag_max * 64);
    const float* in = a.parity ? a.buf1 : a.buf0;
    float* out = a.parity ? a.buf0 : a.buf1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int par = wave & 1, grp = wave >> 1;
    const int q = lane >> 4, n = lane & 15;
    int loaded = -1;
    const int t_end = a.seg[blockIdx.x + 1];
    for (int t = a.seg[blockIdx.x]; t < t_end; ++t) {
        const int4 tj = a.tiles[t];  // (world, x0, y0, -)
        const BatchWorld w = a.worlds[tj.x];
        const int x0 = tj.y, y0 = tj.z;
        int staged = -1;  // a new tile always restages
        f32x4 acc[S][NB];
        kernel_potential<S, NB>(acc, tile, frag_lds, rows_lds, staged, loaded, in + w.off, 0, w.W, w.H, x0, y0,
                                a.halo_y, a.halo_x, a.tr, a.tc, a.tstride, a.frags, w.frag0, w.nfrag, w.terms,
                                a.sched + w.sched0, w.nrow);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int xo = x0 + 16 * (par + 2 * k) + 4 * q;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int yo = y0 + grp * 16 * NB + 16 * b + n;
                const size_t o = w.off + (size_t)yo * w.W + xo;
                float4 sv = *reinterpret_cast<const float4*>(in + o);
                float r[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    r[e] = lenia_update(r[e], w.dt,
                                        add_weighted(0.0f, 1.0f, growth(potential(acc[k][b][e], w.scale), w.mu,
                                                                        w.inv_sigma)));
                *reinterpret_cast<float4*>(out + o) = make_float4(r[0], r[1], r[2], r[3]);
            }
        }
    }
}

// Dynamic LDS above 64 KiB must be allowed per kernel.  Done once per kernel
// symbol (up to the 160 KiB a gfx950 workgroup can hold), never inside a
// stream capture.
template <typename K>
void allow_lds(K kernel, size_t bytes) {
    allow_dynamic_lds(reinterpret_cast<const void*>(kernel), bytes);
}

}  // namespace

// LDS: the bf16 tile (or ring), the fragment table, the row list.
size_t direct_lds_bytes(const DirectArgs& a) {
    const int rows = a.variant == 1 ? a.tr + 64 : a.tr;
    return (size_t)rows * a.tstride * sizeof(uint16_t) + (size_t)a.nfrag_max * 64 * sizeof(uint4) +
           DIRECT_MAX_ROWS * sizeof(int4);
}

size_t batch_lds_bytes(const BatchArgs& a) {
    return (size_t)a.tr * a.tstride * sizeof(uint16_t) + (size_t)a.nfrag_max * 64 * sizeof(uint4) +
           DIRECT_MAX_ROWS * sizeof(int4);
}

template <int S, int NB>
void launch_batch_shape(hipStream_t s, const BatchArgs& a, int grid, size_t lds) {
    allow_lds(k_direct_batch<S, NB>, lds);
    hipLaunchKernelGGL((k_direct_batch<S, NB>), dim3(grid), dim3(DTHREADS), lds, s, a);
}

void launch_direct_batch(hipStream_t s, Profiler* p, const BatchArgs& a, int cols, int rows, int grid) {
    const size_t lds = batch_lds_bytes(a);
    Launch L(p, s, "direct_bf16_batch");
    if (cols == 128 && rows == 128) launch_batch_shape<4, 2>(s, a, grid, lds);
    else if (cols == 128) launch_batch_shape<4, 1>(s, a, grid, lds);
    else if (rows == 128) launch_batch_shape<2, 2>(s, a, grid, lds);
    else launch_batch_shape<2, 1>(s, a, grid, lds);
    AH_HIP(hipGetLastError());
}

template <int S, int NB>
void launch_direct_shape(hipStream_t s, const DirectArgs& a, int channels, size_t lds) {
    dim3 grid(a.W / (32 * S), a.H / (64 * NB), a.conv_out ? 1 : channels);
    allow_lds(k_direct<S, NB>, lds);
    hipLaunchKernelGGL((k_direct<S, NB>), grid, dim3(DTHREADS), lds, s, a);
}

template <int S>
void launch_ring_shape(hipStream_t s, const DirectArgs& a, size_t lds) {
    const int strips = a.W / (32 * S), TY = a.H / 64;
    const int segs = std::max(1, std::min(TY, (cu_count() + strips / 2) / strips));
    allow_lds(k_direct_ring<S>, lds);
    hipLaunchKernelGGL((k_direct_ring<S>), dim3(strips * segs), dim3(DTHREADS), lds, s, a);
}

void launch_direct(hipStream_t s, Profiler* p, const DirectArgs& a, int channels) {
    const int cols = a.cols, rows = a.rows;
    Launch L(p, s, a.conv_out ? "direct_bf16_conv" : "direct_bf16_update");
    if (a.variant == 1 && !a.conv_out) {
        const size_t lds = direct_lds_bytes(a);
        if (cols == 128) launch_ring_shape<4>(s, a, lds);
        else launch_ring_shape<2>(s, a, lds);
        AH_HIP(hipGetLastError());
        return;
    }
    DirectArgs plain = a;  // the potential diagnostic runs the tile kernel of the same shape
    plain.variant = 0;
    const size_t lds = direct_lds_bytes(plain);
    if (a.variant == 2 && !a.conv_out) {
        auto split = [&](auto kernel, int S, int NB) {
            allow_lds(kernel, lds);
            hipLaunchKernelGGL(kernel, dim3(a.W / (32 * S), a.H / (64 * NB), a.n_kernels), dim3(DTHREADS), lds, s, a);
        };
        if (cols == 128 && rows == 128) split(k_direct_split<4, 2>, 4, 2);
        else if (cols == 128) split(k_direct_split<4, 1>, 4, 1);
        else if (rows == 128) split(k_direct_split<2, 2>, 2, 2);
        else split(k_direct_split<2, 1>, 2, 1);
        AH_HIP(hipGetLastError());
        const size_t quads = (size_t)a.W * a.H / 4;
        hipLaunchKernelGGL(k_direct_sum, dim3((unsigned)((quads + 255) / 256), channels), dim3(256), 0, s, a);
        AH_HIP(hipGetLastError());
        return;
    }
    if (cols == 128 && rows == 128) launch_direct_shape<4, 2>(s, plain, channels, lds);
    else if (cols == 128) launch_direct_shape<4, 1>(s, plain, channels, lds);
    else if (rows == 128) launch_direct_shape<2, 2>(s, plain, channels, lds);
    else launch_direct_shape<2, 1>(s, plain, channels, lds);
    AH_HIP(hipGetLastError());
}

}  // namespace ah
