// Internal declarations shared by the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/auto_hip.h"

namespace ah {

// Thread-local error message behind ah_last_error().
void set_error(const std::string& msg);

struct Error {
    ah_status code;
    std::string msg;
};

#define AH_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::ah::Error{AH_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// Per-kernel launch timing on one stream (ah_sim_profile).
struct Profiler {
    bool enabled = false;
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

// ------------------------------------------------------------ launchers
// init.hip
void launch_cos_table64(hipStream_t s, Profiler* p, double* tab, int n);
void launch_twiddles(hipStream_t s, Profiler* p, float2* tw, int n);
// Normalised FP64 taps of one kernel in a (2*reach+1)^2 box; writes taps and raw sum.
void launch_kernel_taps64(hipStream_t s, Profiler* p, const ah_kernel& k, int reach, double* taps, double* sum);
// khat/kpack for the FFT path from FP64 taps (separable cosine transform); T is scratch.
void launch_khat64(hipStream_t s, Profiler* p, int W, int H, int reach, const double* taps, const double* cosW,
                   const double* cosH, double* T, float* khat, float2* kpack);
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
                     int n_kernels);
// conv_out != nullptr: write kernel conv_k's potential there instead of updating the state.
void launch_row_c2r(hipStream_t s, Profiler* p, int W, int H, int channels, int n_kernels, const float2* spec,
                    const float* state_in, float* state_out, const KernelParams* kp, float dt, const float2* twW,
                    float* conv_out, int conv_k);
size_t fft_lds_limit();

}  // namespace ah
