// Internal declarations shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/auto_hip.h"

namespace ah {

// Thread-local error message behind ah_last_error().
void set_error(const std::string& msg);

struct Error {
    ah_status code;
    std::string msg;
};

// Runs f and turns exceptions into an ah_status plus ah_last_error() text.
template <typename F>
ah_status guard(F&& f) {
    try {
        f();
        return AH_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return AH_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_error(e.what());
        return AH_ERR_DEVICE;
    } catch (...) {
        set_error("unknown failure");
        return AH_ERR_DEVICE;
    }
}

uint64_t splitmix64(uint64_t& state);
uint16_t f32_to_bf16(float f);
double bf16_to_f64(uint16_t h);
// Appends the row list and the hi fragments (then, for terms == 3, the lo
// fragments) of one kernel; returns the FP64 sum of the bf16 taps used.
// nrow[c] counts the rows with 3 - c chunks (rows listed in that order).
double build_fragments(const std::vector<double>& taps, int R, int terms, std::vector<int>& rows,
                       std::vector<uint16_t>& frags, int& nfrag, int nrow[3]);
// 1 or 3 bf16 terms for a kernel with these taps and growth (see auto_hip.h).
int direct_terms(const std::vector<double>& taps, float dt, float weight, float sigma, uint32_t path);

#define AH_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::ah::Error{AH_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// Per-kernel launch timing on one stream (ah_sim_profile).
struct Profiler {
    bool enabled = false;
    int every = 1;  // time every `every`-th launch of each kernel name
    std::vector<std::pair<const char*, uint64_t>> seen;  // launches per name
    bool sample(const char* name);
    struct Rec {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t get();
    void reset();
    ~Profiler();
};

// Brackets one launch with events when profiling is on.
struct Launch {
    Profiler* p;
    hipStream_t s;
    const char* name;
    hipEvent_t b = nullptr;
    Launch(Profiler* p_, hipStream_t s_, const char* n_);
    ~Launch();
};

// ---------------------------------------------------------------- layouts
//
// Half-spectrum layout of the FFT path (complex64):
//   spec[kb][y][c], kx = 4*kb + c, kb in [0, W/8), y in [0, H)
// Slot kx=0 packs the two real bins (X[0], X[W/2]) as (re, im).
// khat[kb][y][c] holds the real kernel spectrum / (W*H) for kx >= 1;
// kpack[y] = ((K0+KN)/2, (K0-KN)/2) / (W*H) for the packed slot.
constexpr int BC = 4;  // complex columns per spectrum block (32 bytes)
constexpr int RW = 4;  // rows per row-pass workgroup (4 x 32 B = 128 B segments)

struct KernelParams {  // growth of one kernel, as the update kernels read it
    float mu, inv_sigma, weight;  // weight normalised per destination
    uint32_t src, dst;
};

// render.hip: rgba8 view of channel 0 (center, scale: cells per output pixel).
void launch_render(hipStream_t s, Profiler* prof, const float* state, int W, int H, uint32_t rule, float center_x,
                   float center_y, float scale, int out_w, int out_h, uint32_t* rgba);

// ------------------------------------------------------------ launchers
// init.hip
void launch_cos_table64(hipStream_t s, Profiler* p, double* tab, int n);
void launch_twiddles(hipStream_t s, Profiler* p, float2* tw, int n);
// Normalised FP64 taps of one kernel in a (2*reach+1)^2 box; writes taps and raw sum.
void launch_kernel_taps64(hipStream_t s, Profiler* p, const ah_kernel& k, int reach, double* taps, double* sum);
// How the column pass of a world reads its real kernel spectrum:
//   KHAT_BLOCK   khat[kb][p][c], p in [0, H/2] (every column kernel but one);
//   KHAT_DIGITS  khat[kx][digits(p)], p in [0, H/2], digits(p) = (p >> 8) * 256
//                + (p & 15) * 16 + ((p >> 4) & 15) (H = 4096: the block column
//                kernel, whose thread (k1, m) holds k = k1 + 16 m + 256 r; the
//                swapped digits make a wave's loads contiguous, for p = k and
//                for p = H - k alike).  kx = 0 and W/2 always go to kpack.
enum { KHAT_BLOCK = 0, KHAT_DIGITS = 1 };
// khat/kpack for the FFT path from FP64 taps (separable cosine transform); T is scratch.
void launch_khat64(hipStream_t s, Profiler* p, int W, int H, int reach, const double* taps, const double* cosW,
                   const double* cosH, double* T, float* khat, float2* kpack, int layout);
// The K-hat layout the column pass of this world reads, and the floats of
// khat per kernel in it.
int khat_layout(int W, int H, int channels, int n_kernels, const int* ksrc);
size_t khat_floats(int W, int H, int layout);
void launch_seed(hipStream_t s, Profiler* p, const ah_lenia_desc& d, uint32_t seed, const uint32_t* corners,
                 float* state);
int stats_blocks(size_t cells);
void launch_stats(hipStream_t s, Profiler* p, const float* state, size_t cells, int channels, double* pm,
                  float* pmin, float* pmax, unsigned long long* pnf, int nblocks);

// fft_step.hip
void launch_row_r2c(hipStream_t s, Profiler* p, int W, int H, int channels, const float* state, float2* spec,
                    const float2* twW);
void launch_col_conv(hipStream_t s, Profiler* p, int W, int H, int channels, const float2* spec_in,
                     float2* spec_out, const float* khat, const float2* kpack, const float2* twH, const int* ksrc,
                     int n_kernels, bool diag = false, int variant = 0);
// Column-pass kernels of a batch group (diagonal mode): the fused LDS-staged
// batch kernel, the generic kernel, or the register-resident block kernel.
enum { COL_FUSED = 0, COL_GENERIC = 1, COL_BLOCK = 2 };
// Whether a batch group's column pass has the fused kernel; whether the
// block column kernel applies (every single world of that height, and batch
// groups that choose it); whether a batch group's row pass has a fused
// kernel besides the generic one.
bool batch_col_fused_ok(int W, int H);
bool col_blk_ok(int W, int H);
bool batch_row_fused_ok(int W, int H);
// conv_out != nullptr: write kernel conv_k's potential there instead of updating the state.
// spec_next != nullptr (update only): also write the half spectra of the updated
// state there, which is what launch_row_r2c would produce from it next.
// rule: ah_rule applied to the potentials.
// kp_dev, dt_dev (device arrays, one entry per channel): batched same-size
// worlds, channel c being world c with its own kernel c, growth and dt.
void launch_row_c2r(hipStream_t s, Profiler* p, int W, int H, int channels, int n_kernels, const float2* spec,
                    const float* state_in, float* state_out, const KernelParams* kp, float dt, const float2* twW,
                    float* conv_out, int conv_k, float2* spec_next, int rule,
                    const KernelParams* kp_dev = nullptr, const float* dt_dev = nullptr, bool generic = false);
size_t fft_lds_limit();

// direct_bf16.hip
// Direct path row list: per kernel, the rows dy in [-R, R] that hold taps as
// (dy, e, first fragment, 0), those with 3 chunks first, then 2, then 1 (dy
// ascending within a group), so each group runs with a fixed chunk count.
constexpr int DIRECT_MAX_ROWS = 64;  // 2 * 31 + 1 rows, rounded up

struct DirectKernel {  // one kernel as the direct path reads it
    int reach;         // R: rows dy in [-R, R]
    int nrow[3];       // rows with 3, 2, 1 chunks
    int frag0, nfrag;  // hi fragments [frag0, frag0+nfrag); lo fragments follow them
    int terms;         // 1: K_hi s_hi; 3: + K_hi s_lo + K_lo s_hi
    float scale;       // 1 / sum of the bf16-rounded taps
    int src, dst;
    float mu, inv_sigma, weight;
};

struct DirectArgs {
    int W, H;
    int halo_y, halo_x;  // max R over kernels; max R rounded up to a multiple of 8
    int tr, tc, tstride; // LDS tile rows, columns, row stride (bf16 elements)
    int cols, rows;      // output tile shape of the launch (64 or 128 each)
    int variant;         // 0: tile kernel; 1: persistent ring kernel (one channel, one kernel,
                         // one term, rows = 64); 2: a workgroup per (tile, kernel) + sum pass
    float* gsum;         // variant 2: weighted growth per kernel, [n_kernels][H][W]
    int n_kernels;
    DirectKernel k[AH_MAX_KERNELS];
    const int4* sched;   // per kernel, its row list
    int sched_base[AH_MAX_KERNELS];
    int split_order[AH_MAX_KERNELS];  // variant 2: kernel of grid slice z, costliest first
    const uint4* frags;  // [fragment][64 lanes] x 8 bf16
    int nfrag_max;
    const float* state_in;
    float* state_out;
    float dt;
    float* conv_out;     // set: write kernel conv_k's potential instead of updating
    int conv_k;
};
struct BatchWorld {  // one world of a batch as the batch kernel reads it
    int W, H;
    size_t off;      // first cell in both state buffers
    int reach, frag0, nfrag, sched0, terms;
    int nrow[3];     // rows with 3, 2, 1 chunks in the row list at sched0
    float scale, mu, inv_sigma, dt;
};

struct BatchArgs {
    const BatchWorld* worlds;
    const int4* tiles;  // (world, x0, y0, 0); tile = cols x rows of the launch shape
    const int* seg;     // workgroup b runs tiles [seg[b], seg[b+1])
    int ntiles;
    float* buf0;
    float* buf1;
    int parity;         // 0: buf0 -> buf1, 1: buf1 -> buf0
    const int4* sched;
    const uint4* frags;
    int halo_y, halo_x, tr, tc, tstride, nfrag_max;
};
size_t batch_lds_bytes(const BatchArgs& a);
void launch_direct_batch(hipStream_t s, Profiler* p, const BatchArgs& a, int cols, int rows, int grid);
int cu_count();  // compute units of the current device
// Lets `kernel` launch with more than 64 KiB of dynamic LDS on the current
// device (a no-op at or below 64 KiB); thread-safe, once per device.
void allow_dynamic_lds(const void* kernel, size_t bytes);

// Row stride (bf16 elements) of a direct-path LDS tile at least tc wide:
// the smallest multiple of 8 elements (16 bytes) that is 2 or 14 mod 16 in
// 16-byte units.  The B fragments of one MFMA read 16 consecutive tile rows
// at one column; with this stride their 16-byte reads fall in distinct bank
// groups (measured at 4096^2, R = 26: 121 us per step against 129 us for a
// stride of 13 units and 292 us for 32 units, where all rows share banks).
// Where that stride does not fit the LDS, `compact` gives tc + 8 (an odd
// unit count, the next best).
inline int direct_tile_stride(int tc, bool compact) {
    if (compact) return tc + 8;
    int u = (tc + 7) / 8;
    while (u % 16 != 2 && u % 16 != 14) ++u;
    return 8 * u;
}

// Workgroup barrier for LDS traffic only: this wave's LDS operations finish,
// then all waves meet.  Unlike __syncthreads() it does not wait for global
// loads and stores in flight (its release fence would), so loads issued
// before it keep running across it.  Only for workgroups whose threads
// exchange data through LDS, never through global memory.  The memory
// clobber keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
size_t direct_lds_bytes(const DirectArgs& a);
void launch_direct(hipStream_t s, Profiler* p, const DirectArgs& a, int channels);

void launch_copy_probe(hipStream_t s, const float4* src, float4* dst, size_t n4);

// Access to a sim for the surrogate (engine.cpp).
float* sim_state(ah_sim* s);
hipStream_t sim_stream(ah_sim* s);
const ah_lenia_desc& sim_desc(const ah_sim* s);
void sim_advance(ah_sim* s, uint32_t steps);  // plain steps on the sim's own path
void sim_add_generations(ah_sim* s, uint64_t n);

}  // namespace ah
