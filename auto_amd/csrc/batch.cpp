// Batched uneven worlds: hundreds of single-channel Lenia worlds of different
// sizes, rules and paths advanced together.
//
// The reference runs one world per Sim (crates/auto-core/src/sim.rs:202-309);
// it has no batching.  Direct-BF16 worlds share one pair of ping-pong
// buffers; a tile list (grouped by world) is cut into contiguous segments of
// near-equal MFMA cost, one per workgroup of the persistent kernel
// (direct_bf16.hip k_direct_batch).  FFT worlds are grouped by size; a group
// keeps its worlds as the channels of one state, one half spectrum and one
// K-hat table per world, and steps as one column-pass and one row-pass launch
// in diagonal mode (world c's kernel reads and writes channel c only, with
// its own growth and dt).  The direct launch's tile shape and grid are chosen
// by an autotuner that times every admissible candidate on the batch itself
// at creation.  Steps run as hipGraph replays of `graph_steps` steps (an even
// count, so the ping-pong parity is unchanged by a replay), with single eager
// steps for the remainder.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "ah_internal.h"

using namespace ah;

// Same-size FFT worlds, one per channel.
struct FftGroup {
    int W = 0, H = 0, n = 0;
    std::vector<int> worlds;     // batch world index of each channel
    float* state = nullptr;      // [n][H][W], updated in place
    float2* spec = nullptr;      // [n][W/8][H][4] half spectra, in place through the column pass
    float* khat = nullptr;       // [n][W/8][H/2+1][4]
    float2* kpack = nullptr;     // [n][H/2+1]
    float2* twW = nullptr;
    float2* twH = nullptr;
    KernelParams* kp = nullptr;  // [n] growth per world
    float* dt = nullptr;         // [n]
    bool ready = false;          // spec holds the spectra of state
    int col_variant = COL_FUSED; // tuned: which column-pass kernel (COL_FUSED, COL_GENERIC, COL_BLOCK)
    bool row_generic = false;    // tuned: the generic row kernel instead of the fused one
};

struct ah_batch {
    std::vector<ah_lenia_desc> d;
    std::vector<uint32_t> seeds;
    int device = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    std::vector<void*> bufs;
    size_t total_bytes = 0;
    float* buf[2] = {nullptr, nullptr};
    std::vector<BatchWorld> worlds_h;
    BatchWorld* worlds = nullptr;
    std::vector<int> cost;  // MFMA chunks per 16x16 output block, per world
    std::vector<char> fft;  // world runs the FFT path
    std::vector<int> group_of, slot_of;
    std::vector<FftGroup> groups;
    int4* sched = nullptr;
    uint4* frags = nullptr;
    int halo_y = 0, halo_x = 0, nfrag_max = 0;
    uint64_t total_cells = 0;

    // chosen tiling
    int cols = 0, rows = 0, grid = 0, ntiles = 0;
    int4* tiles = nullptr;
    int* seg = nullptr;
    ah_batch_tune tune[AH_BATCH_MAX_CANDIDATES]{};
    uint32_t n_tune = 0;
    ah_batch_group_tune group_tune[AH_BATCH_MAX_GROUP_TUNE]{};
    uint32_t n_group_tune = 0;

    int parity = 0;
    uint64_t generation = 0;
    uint32_t graph_steps = 100;
    hipGraphExec_t graph[2] = {nullptr, nullptr};

    void* alloc(size_t bytes) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            throw Error{AH_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e)};
        bufs.push_back(p);
        total_bytes += bytes;
        return p;
    }
    void release(void* p) {
        auto it = std::find(bufs.begin(), bufs.end(), p);
        if (it != bufs.end()) {
            (void)hipFree(p);
            bufs.erase(it);
        }
    }

    ~ah_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        for (void* p : bufs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    BatchArgs args(int par) const {
        BatchArgs a{};
        a.worlds = worlds;
        a.tiles = tiles;
        a.seg = seg;
        a.ntiles = ntiles;
        a.buf0 = buf[0];
        a.buf1 = buf[1];
        a.parity = par;
        a.sched = sched;
        a.frags = frags;
        a.halo_y = halo_y;
        a.halo_x = halo_x;
        a.tr = rows + 2 * halo_y;
        a.tc = cols + 2 * halo_x + 32;
        a.nfrag_max = nfrag_max;
        a.tstride = direct_tile_stride(a.tc, false);
        if (batch_lds_bytes(a) > 160 * 1024) a.tstride = direct_tile_stride(a.tc, true);
        return a;
    }
};

namespace {

bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

void validate_world(const ah_lenia_desc& d, uint32_t i) {
    auto bad = [&](const std::string& m) { throw Error{AH_ERR_INVALID, "world " + std::to_string(i) + ": " + m}; };
    if (d.path > AH_PATH_AUTO) bad("unknown path");
    if (!pow2(d.width) || !pow2(d.height)) bad("width and height must be powers of two");
    if (d.width < 64 || d.height < 64 || d.width > 8192 || d.height > 8192) bad("width and height 64..8192");
    if (d.channels != 1 || d.n_kernels != 1) bad("batched worlds are single-channel with one kernel");
    const ah_kernel& k = d.kernels[0];
    if (d.rule != AH_RULE_LENIA || k.shape != AH_SHAPE_LENIA) bad("batched worlds run the Lenia rule and shape");
    if (!(k.radius > 0.0f) || !std::isfinite(k.radius)) bad("kernel radius must be positive");
    const int reach = (int)std::ceil((double)k.radius);
    if (2 * reach >= (int)std::min(d.width, d.height)) bad("world smaller than its kernel");
    const bool direct = d.path == AH_PATH_DIRECT_BF16 || d.path == AH_PATH_DIRECT_BF16_X1 || d.path == AH_PATH_DIRECT_BF16_X3;
    if (direct && reach > 31) throw Error{AH_ERR_UNSUPPORTED, "world " + std::to_string(i) + ": direct path radius above 31"};
    if (reach > 127) throw Error{AH_ERR_UNSUPPORTED, "world " + std::to_string(i) + ": radius above 127"};
    if (k.rings < 1 || k.rings > AH_MAX_RINGS) bad("rings must be 1..3");
    if (!(k.sigma > 0.0f) || !std::isfinite(k.mu)) bad("sigma must be positive, mu finite");
    if (!std::isfinite(d.dt) || d.dt <= 0.0f) bad("dt must be a positive finite number");
    if (d.init > AH_INIT_ZERO || d.init == AH_INIT_BLOBS) bad("batched worlds take ORBIUM, PATCH or ZERO inits");
    if (d.init == AH_INIT_PATCH && (d.init_size == 0 || d.init_size > std::min(d.width, d.height)))
        bad("patch size must be 1..side");
}

// Tiles (cols x rows) of every world, and contiguous segments of near-equal
// cost for `grid` workgroups.
void build_tiling(ah_batch* b, int cols, int rows, int grid, std::vector<int4>& tiles, std::vector<int>& seg) {
    tiles.clear();
    std::vector<double> tcost;
    for (size_t w = 0; w < b->d.size(); ++w) {
        if (b->fft[w]) continue;
        const int W = (int)b->d[w].width, H = (int)b->d[w].height;
        for (int y0 = 0; y0 < H; y0 += rows)
            for (int x0 = 0; x0 < W; x0 += cols) {
                tiles.push_back(make_int4((int)w, x0, y0, 0));
                tcost.push_back((double)b->cost[w]);
            }
    }
    double total = 0.0;
    for (double c : tcost) total += c;
    seg.assign(grid + 1, (int)tiles.size());
    seg[0] = 0;
    double acc = 0.0;
    int g = 1;
    for (size_t t = 0; t < tiles.size() && g < grid; ++t) {
        acc += tcost[t];
        while (g < grid && acc >= total * g / grid) seg[g++] = (int)t + 1;
    }
}

void install_tiling(ah_batch* b, int cols, int rows, int grid) {
    std::vector<int4> tiles;
    std::vector<int> seg;
    build_tiling(b, cols, rows, grid, tiles, seg);
    if (b->tiles) b->release(b->tiles);
    if (b->seg) b->release(b->seg);
    b->tiles = (int4*)b->alloc(sizeof(int4) * tiles.size());
    b->seg = (int*)b->alloc(sizeof(int) * seg.size());
    AH_HIP(hipMemcpy(b->tiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
    AH_HIP(hipMemcpy(b->seg, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice));
    b->cols = cols;
    b->rows = rows;
    b->grid = grid;
    b->ntiles = (int)tiles.size();
    for (auto& g : b->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Device bytes one FFT world holds in its group: state, spectra, kernel
// spectra and the packed column.
size_t fft_world_bytes(int W, int H) {
    const size_t cells = (size_t)W * H, spec = (size_t)(W / 2) * H, np = (size_t)H / 2 + 1;
    return sizeof(float) * cells + sizeof(float2) * spec + sizeof(float) * (size_t)(W / 8) * np * BC +
           sizeof(float2) * np;
}

// The largest working set of one FFT group: 256 MiB, the Infinity Cache
// (the memory-side last-level cache) of an MI355X.  Measured on 100 worlds
// of 1024^2 (9.1 MiB each) stepped group-major: caps of 96, 192, 256, 320
// and 512 MiB gave 0.70, 0.64, 0.62, 0.67 and 0.70 ms per step of the
// 384-world batch.  AUTO_HIP_BATCH_GROUP_MB overrides it (0: one group per
// size).
size_t group_cap_bytes() {
    const char* e = std::getenv("AUTO_HIP_BATCH_GROUP_MB");
    if (e && *e) return (size_t)std::strtoull(e, nullptr, 10) << 20;
    return (size_t)256 << 20;
}

// AUTO_HIP_BATCH_ORDER=step issues the launches step by step instead.
bool group_major() {
    const char* e = std::getenv("AUTO_HIP_BATCH_ORDER");
    return !(e && std::strcmp(e, "step") == 0);
}

// One step of one FFT group: the column pass, then the row pass (which also
// writes the next step's spectra).
void launch_group_step(ah_batch* b, FftGroup& g, Profiler* p, int col_variant, bool row_generic) {
    launch_col_conv(b->stream, p, g.W, g.H, g.n, g.spec, g.spec, g.khat, g.kpack, g.twH, nullptr, 0, true,
                    col_variant);
    launch_row_c2r(b->stream, p, g.W, g.H, g.n, g.n, g.spec, g.state, g.state, nullptr, 0.0f, g.twW, nullptr, 0,
                   g.spec, AH_RULE_LENIA, g.kp, g.dt, row_generic);
}

void launch_direct_step(ah_batch* b, Profiler* p) {
    if (b->ntiles) launch_direct_batch(b->stream, p, b->args(b->parity), b->cols, b->rows, b->grid);
    b->parity ^= 1;
}

void launch_step(ah_batch* b, Profiler* p) {
    launch_direct_step(b, p);
    for (FftGroup& g : b->groups) launch_group_step(b, g, p, g.col_variant, g.row_generic);
}

// `steps` steps of every world.  The worlds are independent, so the order
// of launches is free: group-major order (the direct launches, then each
// FFT group's `steps` steps back to back) keeps a group's state, spectra and
// kernel spectra in the 256 MiB Infinity Cache from one step to the next,
// where step-major order streams the whole batch through HBM every step.
// Each world's arithmetic is the same in either order (bit for bit).
void launch_steps(ah_batch* b, uint32_t steps, Profiler* p) {
    if (!group_major()) {
        for (uint32_t i = 0; i < steps; ++i) launch_step(b, p);
        return;
    }
    for (uint32_t i = 0; i < steps; ++i) launch_direct_step(b, p);
    for (FftGroup& g : b->groups)
        for (uint32_t i = 0; i < steps; ++i) launch_group_step(b, g, p, g.col_variant, g.row_generic);
}

// Times every admissible (column kernel, row kernel) pair of each size's
// first FFT group on the batch (zero states at creation: the work does not
// depend on the data), keeps the fastest for every group of that size and
// records every candidate.
// AUTO_HIP_BATCH_FFT_KERNELS=fused|block|generic skips the timing and takes
// the fused (or, for the column pass, block) kernels wherever they apply, or
// the generic ones everywhere (tests use it to run each variant on the same
// batch).
void tune_groups(ah_batch* b) {
    b->n_group_tune = 0;
    if (b->groups.empty()) return;
    const char* force = std::getenv("AUTO_HIP_BATCH_FFT_KERNELS");
    if (force && *force && std::strcmp(force, "tune") != 0) {
        const bool generic = std::strcmp(force, "generic") == 0, block = std::strcmp(force, "block") == 0;
        if (!generic && !block && std::strcmp(force, "fused") != 0)
            throw Error{AH_ERR_INVALID, "AUTO_HIP_BATCH_FFT_KERNELS must be fused, block, generic or tune"};
        for (FftGroup& g : b->groups) {
            g.col_variant = generic ? COL_GENERIC
                            : block && col_blk_ok(g.W, g.H) ? COL_BLOCK
                            : batch_col_fused_ok(g.W, g.H)          ? COL_FUSED
                                                                    : COL_GENERIC;
            g.row_generic = generic || !batch_row_fused_ok(g.W, g.H);
        }
        return;
    }
    hipEvent_t e0, e1;
    AH_HIP(hipEventCreate(&e0));
    AH_HIP(hipEventCreate(&e1));
    try {
        for (size_t gi = 0; gi < b->groups.size(); ++gi) {
            FftGroup& g = b->groups[gi];
            bool done = false;  // groups of one size take the first one's choice
            for (size_t e = 0; e < gi && !done; ++e)
                if (b->groups[e].W == g.W && b->groups[e].H == g.H) {
                    g.col_variant = b->groups[e].col_variant;
                    g.row_generic = b->groups[e].row_generic;
                    done = true;
                }
            if (done) continue;
            double best = 1e300;
            for (int cg : {COL_FUSED, COL_GENERIC, COL_BLOCK}) {
                if (cg == COL_FUSED && !batch_col_fused_ok(g.W, g.H)) continue;
                if (cg == COL_BLOCK && !col_blk_ok(g.W, g.H)) continue;
                for (int rg = 0; rg < 2; ++rg) {
                    if (!rg && !batch_row_fused_ok(g.W, g.H)) continue;
                    // Warm-up, then enough steps for about 2 ms of timing.
                    launch_group_step(b, g, nullptr, cg, rg);
                    AH_HIP(hipEventRecord(e0, b->stream));
                    launch_group_step(b, g, nullptr, cg, rg);
                    AH_HIP(hipEventRecord(e1, b->stream));
                    AH_HIP(hipEventSynchronize(e1));
                    float est = 0.0f;
                    AH_HIP(hipEventElapsedTime(&est, e0, e1));
                    const int reps = std::max(5, std::min(200, (int)(2.0f / std::max(est, 1e-3f))));
                    AH_HIP(hipEventRecord(e0, b->stream));
                    for (int r = 0; r < reps; ++r) launch_group_step(b, g, nullptr, cg, rg);
                    AH_HIP(hipEventRecord(e1, b->stream));
                    AH_HIP(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    AH_HIP(hipEventElapsedTime(&ms, e0, e1));
                    const double per = ms / reps;
                    if (b->n_group_tune < AH_BATCH_MAX_GROUP_TUNE)
                        b->group_tune[b->n_group_tune++] =
                            ah_batch_group_tune{(uint32_t)g.W, (uint32_t)g.H, (uint32_t)cg, (uint32_t)rg, per};
                    if (per < best) {
                        best = per;
                        g.col_variant = cg;
                        g.row_generic = rg;
                    }
                }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// Pass 1 for FFT groups whose state changed outside a step (seeded or
// written); every step's row pass leaves the next spectra behind.
void ensure_spectra(ah_batch* b) {
    for (FftGroup& g : b->groups)
        if (!g.ready) {
            launch_row_r2c(b->stream, nullptr, g.W, g.H, g.n, g.state, g.spec, g.twW);
            g.ready = true;
        }
}

// Times each admissible (tile shape, grid) on the batch and keeps the fastest.
void autotune(ah_batch* b) {
    int min_h = 1 << 30, min_w = 1 << 30;
    for (size_t w = 0; w < b->d.size(); ++w) {
        if (b->fft[w]) continue;
        min_h = std::min(min_h, (int)b->d[w].height);
        min_w = std::min(min_w, (int)b->d[w].width);
    }
    b->n_tune = 0;
    if (min_h == 1 << 30) {  // every world runs the FFT path
        b->cols = b->rows = b->grid = b->ntiles = 0;
        return;
    }
    hipDeviceProp_t prop;
    AH_HIP(hipGetDeviceProperties(&prop, b->device));
    const int cus = prop.multiProcessorCount;
    hipEvent_t e0, e1;
    AH_HIP(hipEventCreate(&e0));
    AH_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int best_c = 0, best_r = 0, best_grid = 0;
    b->n_tune = 0;
    try {
        for (int cols : {128, 64})
            for (int rows : {128, 64}) {
                if (cols > min_w || rows > min_h) continue;
                b->cols = cols;
                b->rows = rows;
                const size_t lds = batch_lds_bytes(b->args(0));
                if (lds > 160 * 1024) continue;
                const int per_cu = std::max(1, (int)((160 * 1024) / lds));
                for (int k = 1; k <= std::min(per_cu, 2); ++k) {
                    const int grid = cus * k;
                    install_tiling(b, cols, rows, grid);
                    launch_step(b, nullptr);  // warm-up
                    const int reps = 4;
                    AH_HIP(hipEventRecord(e0, b->stream));
                    for (int r = 0; r < reps; ++r) launch_step(b, nullptr);
                    AH_HIP(hipEventRecord(e1, b->stream));
                    AH_HIP(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    AH_HIP(hipEventElapsedTime(&ms, e0, e1));
                    const double per = ms / reps;
                    if (b->n_tune < AH_BATCH_MAX_CANDIDATES)
                        b->tune[b->n_tune++] = ah_batch_tune{(uint32_t)cols, (uint32_t)rows, (uint32_t)grid, per};
                    if (per < best) {
                        best = per;
                        best_c = cols;
                        best_r = rows;
                        best_grid = grid;
                    }
                }
            }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!best_c) throw Error{AH_ERR_UNSUPPORTED, "no tile shape fits the batch in LDS"};
    install_tiling(b, best_c, best_r, best_grid);
    b->parity = 0;
}

void setup(ah_batch* b) {
    const size_t n = b->d.size();
    AH_HIP(hipGetDevice(&b->device));
    AH_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->fft.assign(n, 0);
    b->group_of.assign(n, -1);
    b->slot_of.assign(n, -1);
    // Direct worlds: fragments and schedules; every world: its path.
    std::vector<int> sched;
    std::vector<uint16_t> frags;
    b->worlds_h.assign(n, BatchWorld{});
    b->cost.assign(n, 0);
    size_t off = 0;
    int rmax = 0, nfrag_total = 0, reach_max = 0;
    for (auto& d : b->d) reach_max = std::max(reach_max, (int)std::ceil((double)d.kernels[0].radius));
    const size_t box = (size_t)(2 * reach_max + 1) * (2 * reach_max + 1);
    std::vector<void*> scratch;
    auto tmp = [&](size_t bytes) {
        void* p = nullptr;
        AH_HIP(hipMalloc(&p, bytes));
        scratch.push_back(p);
        return p;
    };
    try {
        double* taps = (double*)tmp(sizeof(double) * box);
        double* sum = (double*)tmp(sizeof(double));
        // Paths first.  AUTO: FFT when the world's direct cost (32-column
        // chunks over the kernel rows, times the bf16 terms) exceeds
        // AH_BATCH_FFT_ABOVE or its radius exceeds 31.  A direct launch has a
        // fixed cost (~35 us at 4096-tile scale, measured) that a few small
        // worlds do not repay, while an FFT group absorbs them almost for
        // free: when no world asks for the direct path itself and the AUTO
        // worlds that chose it hold under 1/20 of the cells, they join the
        // FFT groups too.
        std::vector<std::vector<double>> tapv(n);
        std::vector<char> fft_path(n, 0);
        size_t auto_direct_cells = 0, all_cells = 0;
        bool explicit_direct = false;
        for (size_t w = 0; w < n; ++w) {
            const ah_lenia_desc& d = b->d[w];
            const int R = (int)std::ceil((double)d.kernels[0].radius), side = 2 * R + 1;
            launch_kernel_taps64(b->stream, nullptr, d.kernels[0], R, taps, sum);
            std::vector<double>& t = tapv[w];
            t.resize((size_t)side * side);
            AH_HIP(hipMemcpyAsync(t.data(), taps, sizeof(double) * t.size(), hipMemcpyDeviceToHost, b->stream));
            AH_HIP(hipStreamSynchronize(b->stream));
            bool use_fft = d.path == AH_PATH_FFT_FP32;
            if (d.path == AH_PATH_AUTO) {
                if (R > 31) {
                    use_fft = true;
                } else {
                    std::vector<int> s_tmp;
                    std::vector<uint16_t> f_tmp;
                    int nf = 0;
                    const int terms = direct_terms(t, d.dt, 1.0f, d.kernels[0].sigma, AH_PATH_DIRECT_BF16);
                    int nrow[3];
                    (void)build_fragments(t, R, terms, s_tmp, f_tmp, nf, nrow);
                    const int chunks = 3 * nrow[0] + 2 * nrow[1] + nrow[2];
                    use_fft = chunks * terms > AH_BATCH_FFT_ABOVE;
                    if (!use_fft) auto_direct_cells += (size_t)d.width * d.height;
                }
            } else if (!use_fft) {
                explicit_direct = true;
            }
            fft_path[w] = use_fft;
            all_cells += (size_t)d.width * d.height;
        }
        if (!explicit_direct && auto_direct_cells * 20 < all_cells)
            for (size_t w = 0; w < n; ++w)
                if (b->d[w].path == AH_PATH_AUTO) fft_path[w] = 1;
        // FFT worlds of one size form groups whose working set (state,
        // spectra, kernel spectra) stays under the group cap, split evenly.
        std::vector<std::pair<std::pair<int, int>, int>> per_size;  // (W, H) -> worlds
        for (size_t w = 0; w < n; ++w) {
            if (!fft_path[w]) continue;
            const std::pair<int, int> wh((int)b->d[w].width, (int)b->d[w].height);
            auto it = std::find_if(per_size.begin(), per_size.end(), [&](auto& e) { return e.first == wh; });
            if (it == per_size.end()) per_size.emplace_back(wh, 1);
            else ++it->second;
        }
        const size_t cap = group_cap_bytes();
        auto group_max = [&](int W, int H) {
            int count = 0;
            for (auto& e : per_size)
                if (e.first == std::make_pair(W, H)) count = e.second;
            const size_t ws = fft_world_bytes(W, H);
            if (!cap) return count;
            const int chunks = (int)std::max<size_t>(1, ((size_t)count * ws + cap - 1) / cap);
            return (count + chunks - 1) / chunks;
        };
        for (size_t w = 0; w < n; ++w) {
            const ah_lenia_desc& d = b->d[w];
            const int R = (int)std::ceil((double)d.kernels[0].radius);
            const std::vector<double>& t = tapv[w];
            const bool use_fft = fft_path[w];
            if (use_fft) {
                b->fft[w] = 1;
                int gi = -1;
                const int gmax = group_max((int)d.width, (int)d.height);
                for (size_t g = 0; g < b->groups.size(); ++g)
                    if (b->groups[g].W == (int)d.width && b->groups[g].H == (int)d.height && b->groups[g].n < gmax)
                        gi = (int)g;
                if (gi < 0) {
                    b->groups.emplace_back();
                    gi = (int)b->groups.size() - 1;
                    b->groups[gi].W = (int)d.width;
                    b->groups[gi].H = (int)d.height;
                }
                FftGroup& g = b->groups[gi];
                b->group_of[w] = gi;
                b->slot_of[w] = g.n++;
                g.worlds.push_back((int)w);
                continue;
            }
            const uint32_t path = d.path == AH_PATH_AUTO ? (uint32_t)AH_PATH_DIRECT_BF16 : d.path;
            rmax = std::max(rmax, R);
            BatchWorld& bw = b->worlds_h[w];
            bw.W = (int)d.width;
            bw.H = (int)d.height;
            bw.off = off;
            bw.reach = R;
            bw.frag0 = nfrag_total;
            bw.sched0 = (int)(sched.size() / 4);
            bw.terms = direct_terms(t, d.dt, 1.0f, d.kernels[0].sigma, path);
            bw.scale = (float)(1.0 / build_fragments(t, R, bw.terms, sched, frags, bw.nfrag, bw.nrow));
            bw.mu = d.kernels[0].mu;
            bw.inv_sigma = 1.0f / d.kernels[0].sigma;
            bw.dt = d.dt;
            b->cost[w] = (3 * bw.nrow[0] + 2 * bw.nrow[1] + bw.nrow[2]) * bw.terms;
            nfrag_total += bw.nfrag * (bw.terms == 3 ? 2 : 1);
            b->nfrag_max = std::max(b->nfrag_max, bw.nfrag);
            off += (size_t)d.width * d.height;
        }

        // FFT groups: buffers, twiddles, kernel spectra and growth per world.
        for (FftGroup& g : b->groups) {
            const size_t cells = (size_t)g.W * g.H, spec = (size_t)(g.W / 2) * g.H, np = (size_t)g.H / 2 + 1;
            const size_t khat_n = (size_t)(g.W / 8) * np * BC;
            g.state = (float*)b->alloc(sizeof(float) * cells * g.n);
            g.spec = (float2*)b->alloc(sizeof(float2) * spec * g.n);
            g.khat = (float*)b->alloc(sizeof(float) * khat_n * g.n);
            g.kpack = (float2*)b->alloc(sizeof(float2) * np * g.n);
            g.twW = (float2*)b->alloc(sizeof(float2) * g.W);
            g.twH = (float2*)b->alloc(sizeof(float2) * g.H);
            g.kp = (KernelParams*)b->alloc(sizeof(KernelParams) * g.n);
            g.dt = (float*)b->alloc(sizeof(float) * g.n);
            AH_HIP(hipMemsetAsync(g.state, 0, sizeof(float) * cells * g.n, b->stream));
            AH_HIP(hipMemsetAsync(g.spec, 0, sizeof(float2) * spec * g.n, b->stream));
            AH_HIP(hipMemsetAsync(g.khat, 0, sizeof(float) * khat_n * g.n, b->stream));
            launch_twiddles(b->stream, nullptr, g.twW, g.W);
            launch_twiddles(b->stream, nullptr, g.twH, g.H);
            double* cosW = (double*)tmp(sizeof(double) * g.W);
            double* cosH = (double*)tmp(sizeof(double) * g.H);
            launch_cos_table64(b->stream, nullptr, cosW, g.W);
            launch_cos_table64(b->stream, nullptr, cosH, g.H);
            double* T = (double*)tmp(sizeof(double) * (2 * reach_max + 1) * (g.W / 2 + 1));
            std::vector<KernelParams> kp(g.n);
            std::vector<float> dt(g.n);
            for (int c = 0; c < g.n; ++c) {
                const ah_lenia_desc& d = b->d[g.worlds[c]];
                const ah_kernel& k = d.kernels[0];
                const int R = (int)std::ceil((double)k.radius);
                launch_kernel_taps64(b->stream, nullptr, k, R, taps, sum);
                launch_khat64(b->stream, nullptr, g.W, g.H, R, taps, cosW, cosH, T, g.khat + khat_n * c,
                              g.kpack + np * c, KHAT_BLOCK);
                kp[c] = KernelParams{k.mu, 1.0f / k.sigma, 1.0f, 0u, 0u};
                dt[c] = d.dt;
            }
            AH_HIP(hipMemcpyAsync(g.kp, kp.data(), sizeof(KernelParams) * g.n, hipMemcpyHostToDevice, b->stream));
            AH_HIP(hipMemcpyAsync(g.dt, dt.data(), sizeof(float) * g.n, hipMemcpyHostToDevice, b->stream));
            AH_HIP(hipStreamSynchronize(b->stream));  // taps, cos tables and T are reused
        }
        AH_HIP(hipStreamSynchronize(b->stream));
    } catch (...) {
        (void)hipStreamSynchronize(b->stream);
        for (void* p : scratch) (void)hipFree(p);
        throw;
    }
    for (void* p : scratch) (void)hipFree(p);
    b->total_cells = 0;
    for (auto& d : b->d) b->total_cells += (uint64_t)d.width * d.height;
    b->halo_y = rmax;
    b->halo_x = 8 * ((rmax + 7) / 8);
    const size_t dcells = std::max<size_t>(off, 1);
    b->buf[0] = (float*)b->alloc(sizeof(float) * dcells);
    b->buf[1] = (float*)b->alloc(sizeof(float) * dcells);
    AH_HIP(hipMemsetAsync(b->buf[0], 0, sizeof(float) * dcells, b->stream));
    AH_HIP(hipMemsetAsync(b->buf[1], 0, sizeof(float) * dcells, b->stream));
    b->worlds = (BatchWorld*)b->alloc(sizeof(BatchWorld) * n);
    b->sched = (int4*)b->alloc(sizeof(int) * std::max<size_t>(sched.size(), 4));
    b->frags = (uint4*)b->alloc(sizeof(uint16_t) * std::max<size_t>(frags.size(), 8));
    AH_HIP(hipMemcpy(b->worlds, b->worlds_h.data(), sizeof(BatchWorld) * n, hipMemcpyHostToDevice));
    if (!sched.empty()) AH_HIP(hipMemcpy(b->sched, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice));
    if (!frags.empty())
        AH_HIP(hipMemcpy(b->frags, frags.data(), sizeof(uint16_t) * frags.size(), hipMemcpyHostToDevice));
    AH_HIP(hipStreamSynchronize(b->stream));

    // The tuner steps zero states (the work is data-independent); one eager
    // step warms the FFT launches too.  Then the worlds are seeded.
    autotune(b);
    tune_groups(b);
    launch_step(b, nullptr);
    for (size_t w = 0; w < n; ++w) {
        float* dst = b->fft[w] ? b->groups[b->group_of[w]].state +
                                     (size_t)b->slot_of[w] * b->d[w].width * b->d[w].height
                               : b->buf[0] + b->worlds_h[w].off;
        launch_seed(b->stream, nullptr, b->d[w], b->seeds[w], nullptr, dst);
    }
    AH_HIP(hipStreamSynchronize(b->stream));
    for (FftGroup& g : b->groups) g.ready = false;
    b->parity = 0;
    b->generation = 0;
}

// The state of world w: its slot of the ping-pong buffer or of its FFT group.
float* world_state(ah_batch* b, uint32_t w) {
    if (b->fft[w]) return b->groups[b->group_of[w]].state + (size_t)b->slot_of[w] * b->d[w].width * b->d[w].height;
    return b->buf[b->parity] + b->worlds_h[w].off;
}

hipGraphExec_t graph_for(ah_batch* b) {
    hipGraphExec_t& g = b->graph[b->parity];
    if (g) return g;
    hipGraph_t graph = nullptr;
    AH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int start = b->parity;
    try {
        launch_steps(b, b->graph_steps, nullptr);
    } catch (...) {
        (void)hipStreamEndCapture(b->stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        b->parity = start;
        throw;
    }
    b->parity = start;  // an even number of steps returns to the same parity
    AH_HIP(hipStreamEndCapture(b->stream, &graph));
    hipError_t e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g = nullptr;
        throw Error{AH_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)};
    }
    return g;
}

void run_steps(ah_batch* b, uint32_t steps) {
    ensure_spectra(b);
    for (; steps >= b->graph_steps; steps -= b->graph_steps) {
        AH_HIP(hipGraphLaunch(graph_for(b), b->stream));
        b->generation += b->graph_steps;
    }
    if (steps) {
        launch_steps(b, steps, &b->prof);
        b->generation += steps;
    }
}

}  // namespace

extern "C" {

ah_status ah_batch_new(const ah_lenia_desc* descs, uint32_t n_worlds, const uint32_t* seeds, ah_batch** out) {
    return guard([&] {
        if (!descs || !seeds || !out || n_worlds == 0) throw Error{AH_ERR_INVALID, "null argument or no worlds"};
        *out = nullptr;
        for (uint32_t i = 0; i < n_worlds; ++i) validate_world(descs[i], i);
        std::unique_ptr<ah_batch> b(new ah_batch());
        b->d.assign(descs, descs + n_worlds);
        b->seeds.assign(seeds, seeds + n_worlds);
        setup(b.get());
        *out = b.release();
    });
}

ah_status ah_batch_step(ah_batch* b, uint32_t steps) {
    return guard([&] {
        if (!b) throw Error{AH_ERR_INVALID, "null batch"};
        AH_HIP(hipSetDevice(b->device));
        run_steps(b, steps);
    });
}

ah_status ah_batch_wait(ah_batch* b) {
    return guard([&] {
        if (!b) throw Error{AH_ERR_INVALID, "null batch"};
        AH_HIP(hipStreamSynchronize(b->stream));
    });
}

ah_status ah_batch_read_world(ah_batch* b, uint32_t world, float* dst, size_t bytes) {
    return guard([&] {
        if (!b || !dst) throw Error{AH_ERR_INVALID, "null argument"};
        if (world >= b->d.size()) throw Error{AH_ERR_INVALID, "world index out of range"};
        const ah_lenia_desc& d = b->d[world];
        if (bytes != sizeof(float) * (size_t)d.width * d.height) throw Error{AH_ERR_SIZE, "world is height*width*4 bytes"};
        AH_HIP(hipMemcpyAsync(dst, world_state(b, world), bytes, hipMemcpyDeviceToHost, b->stream));
        AH_HIP(hipStreamSynchronize(b->stream));
    });
}

ah_status ah_batch_write_world(ah_batch* b, uint32_t world, const float* src, size_t bytes) {
    return guard([&] {
        if (!b || !src) throw Error{AH_ERR_INVALID, "null argument"};
        if (world >= b->d.size()) throw Error{AH_ERR_INVALID, "world index out of range"};
        const ah_lenia_desc& d = b->d[world];
        if (bytes != sizeof(float) * (size_t)d.width * d.height) throw Error{AH_ERR_SIZE, "world is height*width*4 bytes"};
        if (b->fft[world]) b->groups[b->group_of[world]].ready = false;
        AH_HIP(hipMemcpyAsync(world_state(b, world), src, bytes, hipMemcpyHostToDevice, b->stream));
        AH_HIP(hipStreamSynchronize(b->stream));
    });
}

ah_status ah_batch_info_get(const ah_batch* b, ah_batch_info* out) {
    return guard([&] {
        if (!b || !out) throw Error{AH_ERR_INVALID, "null argument"};
        ah_batch_info i{};
        i.n_worlds = (uint32_t)b->d.size();
        i.tile_cols = (uint32_t)b->cols;
        i.tile_rows = (uint32_t)b->rows;
        i.grid = (uint32_t)b->grid;
        i.ntiles = (uint32_t)b->ntiles;
        i.graph_steps = b->graph_steps;
        i.total_cells = b->total_cells;
        i.total_bytes = b->total_bytes;
        i.generation = b->generation;
        i.n_tune = b->n_tune;
        for (uint32_t k = 0; k < b->n_tune; ++k) i.tune[k] = b->tune[k];
        for (char f : b->fft) i.fft_worlds += f ? 1 : 0;
        i.fft_groups = (uint32_t)b->groups.size();
        i.n_group_tune = b->n_group_tune;
        for (uint32_t k = 0; k < b->n_group_tune; ++k) i.group_tune[k] = b->group_tune[k];
        *out = i;
    });
}

ah_status ah_batch_measure(ah_batch* b, uint32_t steps, double* kernel_ms, double* graph_wall_ms,
                           double* eager_wall_ms) {
    return guard([&] {
        if (!b || !kernel_ms || !graph_wall_ms || !eager_wall_ms) throw Error{AH_ERR_INVALID, "null argument"};
        if (steps == 0 || steps % b->graph_steps) throw Error{AH_ERR_INVALID, "steps must be a multiple of graph_steps"};
        AH_HIP(hipSetDevice(b->device));
        // Eager launches with per-launch events: summed device time and wall time.
        ensure_spectra(b);
        AH_HIP(hipStreamSynchronize(b->stream));
        b->prof.reset();
        b->prof.enabled = true;
        auto t0 = std::chrono::steady_clock::now();
        launch_steps(b, steps, &b->prof);
        AH_HIP(hipStreamSynchronize(b->stream));
        auto t1 = std::chrono::steady_clock::now();
        double ksum = 0.0;
        for (auto& r : b->prof.recs) {
            float ms = 0.0f;
            AH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            ksum += ms;
        }
        b->prof.reset();
        b->prof.enabled = false;
        b->generation += steps;
        // The same number of steps through graph replays.
        (void)graph_for(b);
        AH_HIP(hipStreamSynchronize(b->stream));
        auto t2 = std::chrono::steady_clock::now();
        run_steps(b, steps);
        AH_HIP(hipStreamSynchronize(b->stream));
        auto t3 = std::chrono::steady_clock::now();
        *kernel_ms = ksum;
        *eager_wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        *graph_wall_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
    });
}

void ah_batch_destroy(ah_batch* b) { delete b; }

}  // extern "C"
