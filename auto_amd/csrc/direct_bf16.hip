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

__device__ __forceinline__ uint32_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                      // round to nearest even
    return u >> 16;
}

__device__ __forceinline__ float bf16_value(uint32_t h) { return __uint_as_float(h << 16); }

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
                    const float e[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
                    uint32_t h[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int c = cx + t;
                        uint32_t hv = bf16_bits(e[t]);
                        if (lo) hv = bf16_bits(e[t] - bf16_value(hv));
                        h[t] = (c >= -halo_y && c < DX + halo_y) ? hv : 0u;
                    }
                    *reinterpret_cast<uint2*>(tile + (size_t)r * tstride + j) =
                        make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
                }
            }
        }
    }
}

// One kernel row with NCH chunks: NCH A fragments, then for every shared
// chunk u of the strips' joint window the NB B fragments, each feeding the
// strips k whose own chunk u-k exists.
template <int S, int NB, int NCH>
__device__ __forceinline__ void conv_row(f32x4 (&acc)[S][NB], const uint16_t* brow, const uint4* afrag, int lane,
                                         int tstride) {
    bf16x8 A[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) A[c] = __builtin_bit_cast(bf16x8, afrag[c * 64 + lane]);
#pragma unroll
    for (int u = 0; u < NCH + S - 1; ++u) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const bf16x8 bv =
                __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(brow + (size_t)16 * b * tstride + 32 * u));
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const int c = u - k;
                if (c >= 0 && c < NCH) acc[k][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[c], bv, acc[k][b], 0, 0, 0);
            }
        }
    }
}

// All MFMAs of one kernel term over this wave's S strips x NB blocks,
// accumulated into acc.
template <int S, int NB>
__device__ __forceinline__ void conv_tile(f32x4 (&acc)[S][NB], const uint16_t* tile, const uint4* frag_lds,
                                          const int4* sched, int reach, int halo_y, int halo_x, int tstride) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int par = wave & 1, grp = wave >> 1;
    const int q = lane >> 4, n = lane & 15;
    const int row0 = grp * 16 * NB + n + halo_y;
    const int colP = 16 * par + halo_x + 8 * q;
    for (int dy = -reach; dy <= reach; ++dy) {
        const int4 sc = sched[dy < 0 ? -dy : dy];  // (e, chunks, first fragment)
        const uint16_t* brow = tile + (size_t)(row0 + dy) * tstride + colP + sc.x;
        const uint4* afrag = frag_lds + sc.z * 64;
        if (sc.y == 3) conv_row<S, NB, 3>(acc, brow, afrag, lane, tstride);
        else if (sc.y == 2) conv_row<S, NB, 2>(acc, brow, afrag, lane, tstride);
        else if (sc.y == 1) conv_row<S, NB, 1>(acc, brow, afrag, lane, tstride);
    }
}

// One kernel's potential over the tile: K_hi s_hi, plus K_lo s_hi and
// K_hi s_lo for three terms.  `staged` tracks which (source, hi/lo) tile sits
// in LDS (source*2 + lo) so that consecutive kernels on one source skip the
// restaging; `loaded` tracks the fragment block in LDS the same way.
template <int S, int NB>
__device__ __forceinline__ void kernel_potential(f32x4 (&acc)[S][NB], uint16_t* tile, uint4* frag_lds, int& staged,
                                                 int& loaded,
                                                 const float* src_plane, int src, int W, int H, int x0, int y0,
                                                 int halo_y, int halo_x, int tr, int tc, int tstride,
                                                 const uint4* __restrict__ frags, int frag0, int nfrag, int terms,
                                                 const int4* sched, int reach) {
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
        if (staged != want) {
            stage_tile(tile, src_plane, W, H, x0, y0, 32 * S, halo_y, halo_x, tr, tc, tstride, lo_tile);
            staged = want;
        }
        if (loaded != fbase) {
            for (int i = threadIdx.x; i < nfrag * 64; i += DTHREADS) frag_lds[i] = frags[(size_t)fbase * 64 + i];
            loaded = fbase;
        }
        __syncthreads();
        conv_tile<S, NB>(acc, tile, frag_lds, sched, reach, halo_y, halo_x, tstride);
    }
}

__device__ __forceinline__ float growth(float u, float mu, float inv_sigma) {
    float d = (u - mu) * inv_sigma;
    return 2.0f * __expf(-0.5f * d * d) - 1.0f;
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
        kernel_potential<S, NB>(acc, tile, frag_lds, staged, loaded, a.state_in + (size_t)kk.src * W * H, kk.src, W,
                                H, x0, y0, a.halo_y, a.halo_x, a.tr, a.tc, a.tstride, a.frags, kk.frag0, kk.nfrag,
                                kk.terms, a.sched + a.sched_base[kn], kk.reach);
#pragma unroll
        for (int k = 0; k < S; ++k)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float u = acc[k][b][r] * kk.scale;
                    if (a.conv_out) g[k][b][r] = u;
                    else g[k][b][r] += kk.weight * growth(u, kk.mu, kk.inv_sigma);
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
                for (int t = 0; t < 4; ++t) r[t] = fminf(fmaxf(r[t] + a.dt * g[k][b][t], 0.0f), 1.0f);
                *reinterpret_cast<float4*>(a.state_out + plane + o) = make_float4(r[0], r[1], r[2], r[3]);
            }
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
        kernel_potential<S, NB>(acc, tile, frag_lds, staged, loaded, in + w.off, 0, w.W, w.H, x0, y0, a.halo_y,
                                a.halo_x, a.tr, a.tc, a.tstride, a.frags, w.frag0, w.nfrag, w.terms, a.sched + w.sched0,
                                w.reach);
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
                    r[e] = fminf(fmaxf(r[e] + w.dt * growth(acc[k][b][e] * w.scale, w.mu, w.inv_sigma), 0.0f), 1.0f);
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
    static std::vector<const void*> done;
    const void* key = reinterpret_cast<const void*>(kernel);
    if (bytes <= 64 * 1024 || std::find(done.begin(), done.end(), key) != done.end()) return;
    AH_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    done.push_back(key);
}

}  // namespace

size_t direct_lds_bytes(const DirectArgs& a) {
    return (size_t)a.tr * a.tstride * sizeof(uint16_t) + (size_t)a.nfrag_max * 64 * sizeof(uint4);
}

size_t batch_lds_bytes(const BatchArgs& a) {
    return (size_t)a.tr * a.tstride * sizeof(uint16_t) + (size_t)a.nfrag_max * 64 * sizeof(uint4);
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

void launch_direct(hipStream_t s, Profiler* p, const DirectArgs& a, int channels) {
    const size_t lds = direct_lds_bytes(a);
    const int cols = a.cols, rows = a.rows;
    Launch L(p, s, a.conv_out ? "direct_bf16_conv" : "direct_bf16_update");
    if (cols == 128 && rows == 128) launch_direct_shape<4, 2>(s, a, channels, lds);
    else if (cols == 128) launch_direct_shape<4, 1>(s, a, channels, lds);
    else if (rows == 128) launch_direct_shape<2, 2>(s, a, channels, lds);
    else launch_direct_shape<2, 1>(s, a, channels, lds);
    AH_HIP(hipGetLastError());
}

}  // namespace ah
