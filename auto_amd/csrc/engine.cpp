// Host engine behind include/auto_hip.h.
//
// One ah_sim owns its device buffers and one HIP stream (default priority).
// It plays the part of auto-core's Sim for the Lenia program
// (crates/auto-core/src/sim.rs:202-309,559-595): creation builds the kernels
// and seeds the state in one go (Sim::new runs the init passes, :286-294),
// stepping is asynchronous (Sim::step, :297-309), reads block.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "ah_internal.h"

namespace ah {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

// ------------------------------------------------------------ profiler
hipEvent_t Profiler::get() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    // Timing-only events: no system-scope fence when recorded.  With the
    // fence each event pair adds a cache writeback to the launch it
    // brackets (~2.5 us on a 64 MiB-writing kernel), which the unprofiled
    // step never pays.  profile_read synchronises the stream itself.
    hipEvent_t e;
    AH_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
}

void Profiler::reset() {
    for (auto& r : recs) {
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    recs.clear();
    seen.clear();
}

bool Profiler::sample(const char* name) {
    for (auto& e : seen)
        if (std::strcmp(e.first, name) == 0) return e.second++ % (uint64_t)every == 0;
    seen.push_back({name, 1});
    return true;
}

Profiler::~Profiler() {
    reset();
    for (auto e : pool) (void)hipEventDestroy(e);
}

Launch::Launch(Profiler* p_, hipStream_t s_, const char* n_) : p(p_), s(s_), name(n_) {
    if (p && p->enabled && p->sample(name)) {
        b = p->get();
        AH_HIP(hipEventRecord(b, s));
    }
}

Launch::~Launch() {
    if (p && p->enabled && b) {
        hipEvent_t e = p->get();
        if (hipEventRecord(e, s) == hipSuccess) p->recs.push_back({name, b, e});
    }
}


// ------------------------------------------------------------ shared helpers
// SplitMix64, crates/auto-life-gpu/src/world.rs:153-159.
uint64_t splitmix64(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
double bf16_to_f64(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

// Row schedule and bf16 fragments of one kernel for the direct path.  For
// kernel row dy with chord half-width dxm the window starts at
// e = -8*ceil(dxm/8) (16-byte aligned B reads) and needs
// floor((dxm + 15 - e)/32) + 1 chunks of 32 input columns.  Fragment lane l,
// element t holds bf16(K[dy][e + 32c + 8q + t - i]) with i = l%16, q = l/16.
// Fragments are numbered by |dy| (rows dy and -dy share them).  Appends the
// row list (dy, e, first fragment, 0) of the rows with 3, then 2, then 1
// chunks (each group: row 0 if present, then the pairs -|dy|, +|dy|) to `rows`, the hi fragments to `frags` and, for three terms, the lo
// fragments bf16(K - bf16(K)) right after them (same numbering, offset nfrag).
double build_fragments(const std::vector<double>& t, int R, int terms, std::vector<int>& rowlist,
                       std::vector<uint16_t>& frags, int& nfrag, int nrow[3]) {
    const int side = 2 * R + 1;
    std::vector<uint16_t> hi(t.size()), lo(t.size());
    double sum = 0.0;
    for (size_t i = 0; i < t.size(); ++i) {
        hi[i] = f32_to_bf16((float)t[i]);
        lo[i] = f32_to_bf16((float)(t[i] - bf16_to_f64(hi[i])));
        sum += bf16_to_f64(hi[i]);
        if (terms == 3) sum += bf16_to_f64(lo[i]);
    }
    std::vector<int> rows;  // (e, chunks, first fragment) per |dy|
    int nf = 0;
    for (int dy = 0; dy <= R; ++dy) {
        int dxm = -1;
        for (int dx = 0; dx <= R; ++dx)
            if (t[(size_t)(dy + R) * side + (dx + R)] != 0.0) dxm = dx;
        int e = 0, chunks = 0;
        if (dxm >= 0) {
            e = -8 * ((dxm + 7) / 8);
            chunks = (dxm + 15 - e) / 32 + 1;
        }
        if (chunks > 3) throw Error{AH_ERR_UNSUPPORTED, "direct path: more than 3 chunks in a kernel row"};
        rows.insert(rows.end(), {e, chunks, nf});
        nf += chunks;
    }
    // Within a group, row 0 (if the group has it) comes first and rows -|dy|
    // and +|dy| follow as adjacent pairs, which share their fragments.
    for (int c = 3; c >= 1; --c) {
        nrow[3 - c] = 0;
        for (int ady = 0; ady <= R; ++ady) {
            const int* r = &rows[3 * ady];
            if (r[1] != c) continue;
            for (int dy : {-ady, ady}) {
                rowlist.insert(rowlist.end(), {dy, r[0], r[2], 0});
                ++nrow[3 - c];
                if (ady == 0) break;  // row 0 once
            }
        }
    }
    for (int part = 0; part < (terms == 3 ? 2 : 1); ++part) {
        const std::vector<uint16_t>& tb = part ? lo : hi;
        for (int dy = 0; dy <= R; ++dy) {
            const int e = rows[3 * dy], chunks = rows[3 * dy + 1];
            for (int c = 0; c < chunks; ++c)
                for (int lane = 0; lane < 64; ++lane) {
                    const int i = lane & 15, q = lane >> 4;
                    for (int el = 0; el < 8; ++el) {
                        const int dx = e + 32 * c + 8 * q + el - i;
                        frags.push_back(dx < -R || dx > R ? 0 : tb[(size_t)(dy + R) * side + (dx + R)]);
                    }
                }
        }
    }
    nfrag = nf;
    return sum;
}

int direct_terms(const std::vector<double>& t, float dt, float weight, float sigma, uint32_t path) {
    if (path == AH_PATH_DIRECT_BF16_X1) return 1;
    if (path == AH_PATH_DIRECT_BF16_X3) return 3;
    double sq = 0.0;
    for (double v : t) sq += v * v;
    const double neff = 1.0 / sq;
    const double slope = 1.213 * (double)dt * (double)weight / (double)sigma;
    return slope * 0.0078 / std::sqrt(neff) <= 1e-3 ? 1 : 3;
}

}  // namespace ah

using namespace ah;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct ah_sim {
    ah_lenia_desc d{};
    uint32_t seed = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    std::vector<DevBuf> bufs;
    size_t total_bytes = 0;
    uint64_t generation = 0;

    float* state = nullptr;      // [C][H][W]
    float* state_alt = nullptr;  // the last rule step's input (steps write out of place)
    bool stepped = false;        // a rule step has run: state_alt is its input
    KernelParams kp[AH_MAX_KERNELS]{};
    int ksrc[AH_MAX_KERNELS]{};
    int reach[AH_MAX_KERNELS]{};

    // FFT path
    float2* spec_src = nullptr;  // [C][W/8][H][4]
    bool spec_ready = false;     // spec_src holds the half spectra of the current state
    float2* spec_k = nullptr;    // [K][W/8][H][4] (aliases spec_src when K == C == 1)
    float* khat = nullptr;       // [K][W/8][H/2+1][4]
    float2* kpack = nullptr;     // [K][H/2+1]
    float2* twW = nullptr;
    float2* twH = nullptr;

    // direct path
    DirectArgs direct{};
    std::vector<ah_direct_tile> direct_tiles;  // admissible variants

    // stats
    double* pmass = nullptr;
    float* pmin = nullptr;
    float* pmax = nullptr;
    unsigned long long* pnf = nullptr;
    int stat_blocks = 0;

    void* alloc(size_t bytes) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess)
            throw Error{AH_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e)};
        bufs.push_back({p, bytes});
        total_bytes += bytes;
        return p;
    }

    ~ah_sim() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& b : bufs) (void)hipFree(b.p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    size_t cells() const { return (size_t)d.width * d.height; }
};

namespace {

bool pow2(uint32_t v) { return v && !(v & (v - 1)); }

// Reject what the reference asserts on (spectral.rs:103-116) and what this engine cannot run.
// Largest |dx|, |dy| a kernel touches: its profile is 0 beyond.
int kernel_reach(const ah_kernel& k) {
    if (k.shape == AH_SHAPE_SMOOTH_DISC || k.shape == AH_SHAPE_SMOOTH_RING)
        return (int)std::ceil((double)k.radius + (double)k.aa);
    return (int)std::ceil((double)k.radius);
}

void validate(const ah_lenia_desc& d) {
    auto bad = [](const std::string& m) { throw Error{AH_ERR_INVALID, m}; };
    if (!pow2(d.width) || !pow2(d.height)) bad("width and height must be powers of two");
    if (d.width < 16 || d.height < 16 || d.width > 8192 || d.height > 8192) bad("width and height must be in 16..8192");
    if (d.channels < 1 || d.channels > AH_MAX_CHANNELS) bad("channels must be 1..3");
    if (d.n_kernels < 1 || d.n_kernels > AH_MAX_KERNELS) bad("n_kernels must be 1..9");
    if (!std::isfinite(d.dt) || d.dt <= 0.0f) bad("dt must be a positive finite number");
    const uint32_t side = std::min(d.width, d.height);
    if (d.rule > AH_RULE_MNCA) bad("unknown rule");
    for (uint32_t k = 0; k < d.n_kernels; ++k) {
        const ah_kernel& kd = d.kernels[k];
        if (kd.shape > AH_SHAPE_BAND) bad("unknown kernel shape");
        if (!(kd.radius > 0.0f) || !std::isfinite(kd.radius)) bad("kernel radius must be positive");
        if (kd.shape == AH_SHAPE_SMOOTH_DISC || kd.shape == AH_SHAPE_SMOOTH_RING)
            if (!(kd.aa > 0.0f) || !std::isfinite(kd.aa)) bad("antialiasing width must be positive");
        if (kd.shape == AH_SHAPE_SMOOTH_RING || kd.shape == AH_SHAPE_BAND)
            if (!(kd.inner >= 0.0f) || !(kd.inner < kd.radius)) bad("inner radius must be in [0, radius)");
        const int reach = kernel_reach(kd);
        if (2 * reach >= (int)side) bad("world smaller than its kernel");
        if (reach > 127) throw Error{AH_ERR_UNSUPPORTED, "kernel reach above 127 cells"};
        if (kd.shape == AH_SHAPE_LENIA && (kd.rings < 1 || kd.rings > AH_MAX_RINGS)) bad("rings must be 1..3");
        if (d.rule == AH_RULE_LENIA) {
            if (!(kd.sigma > 0.0f) || !std::isfinite(kd.mu)) bad("sigma must be positive, mu finite");
            if (!(kd.weight > 0.0f)) bad("kernel weight must be positive");
        }
        if (kd.src >= d.channels || kd.dst >= d.channels) bad("kernel src/dst out of range");
    }
    if (d.rule != AH_RULE_LENIA) {
        // The reference's SmoothLife and MNCA programs: one channel, two kernels (spectral.rs:354-480).
        if (d.channels != 1 || d.n_kernels != 2) bad("SmoothLife and MNCA rules take one channel and two kernels");
        if (d.path != AH_PATH_FFT_FP32)
            throw Error{AH_ERR_UNSUPPORTED, "SmoothLife and MNCA rules run on the FFT path"};
    }
    for (uint32_t k = 0; k < d.n_kernels; ++k)
        if (d.path != AH_PATH_FFT_FP32 && d.kernels[k].shape != AH_SHAPE_LENIA)
            throw Error{AH_ERR_UNSUPPORTED, "direct BF16 path takes Lenia-shaped kernels"};
    if (d.path == AH_PATH_AUTO) bad("AH_PATH_AUTO chooses a path per world in a batch; a Sim takes a fixed path");
    if (d.path > AH_PATH_AUTO) bad("unknown path");
    if (d.path != AH_PATH_FFT_FP32) {
        if (d.width < 64 || d.height < 64)
            throw Error{AH_ERR_UNSUPPORTED, "direct BF16 path needs width >= 64 and height >= 64"};
        for (uint32_t k = 0; k < d.n_kernels; ++k)
            if (std::ceil((double)d.kernels[k].radius) > 31)
                throw Error{AH_ERR_UNSUPPORTED, "direct BF16 path supports kernel radius <= 31"};
    }
    if (d.init > AH_INIT_NOISE) bad("unknown init");
    if (d.init == AH_INIT_ORBIUM && d.channels != 1) bad("Orbium init is single-channel");
    if (d.init == AH_INIT_PATCH && (d.init_size == 0 || d.init_size > side)) bad("patch size must be 1..side");
    if (d.init == AH_INIT_BLOBS && (d.init_size == 0 || d.init_size > side || d.init_count > 65536))
        bad("blob size must be 1..side and count <= 65536");
}


// LDS tile geometry of one direct-path variant: tr rows x tc bf16 columns.
void direct_shape(DirectArgs& a, int cols, int rows, int variant) {
    a.cols = cols;
    a.rows = rows;
    a.variant = variant;
    a.tr = rows + 2 * a.halo_y;
    a.tc = cols + 2 * a.halo_x + 32;
    a.tstride = direct_tile_stride(a.tc, false);
    if (direct_lds_bytes(a) > 160 * 1024) a.tstride = direct_tile_stride(a.tc, true);
}

// Fragments and row schedule of every kernel for the direct path, from the
// FP64 taps on the device.  For kernel row dy with chord half-width dxm the
// window starts at e = -8*ceil(dxm/8) (16-byte aligned B reads) and needs
// floor((dxm + 15 - e)/32) + 1 chunks of 32 input columns.  Fragment lane l,
// element t holds bf16(K[dy][e + 32c + 8q + t - i]) with i = l%16, q = l/16.
void setup_direct(ah_sim* s, const std::vector<double*>& taps) {
    const ah_lenia_desc& d = s->d;
    const int K = (int)d.n_kernels;
    AH_HIP(hipStreamSynchronize(s->stream));  // taps are written on the sim's stream
    DirectArgs& a = s->direct;
    a = DirectArgs{};
    a.W = (int)d.width;
    a.H = (int)d.height;
    int rmax = 0;
    for (int k = 0; k < K; ++k) rmax = std::max(rmax, s->reach[k]);
    a.halo_y = rmax;
    a.halo_x = 8 * ((rmax + 7) / 8);
    // Output tile: the largest of 128x128, 128x64, 64x64 that the world and LDS allow.
    auto shape = [&](int cols, int rows, int ring) { direct_shape(a, cols, rows, ring); };
    a.n_kernels = K;
    std::vector<int> sched;
    std::vector<uint16_t> frags;  // [fragment][64][8]
    int nfrag_total = 0;
    for (int k = 0; k < K; ++k) {
        const int R = s->reach[k], side = 2 * R + 1;
        std::vector<double> t(side * side);
        AH_HIP(hipMemcpy(t.data(), taps[k], sizeof(double) * t.size(), hipMemcpyDeviceToHost));
        DirectKernel& dk = a.k[k];
        dk.reach = R;
        dk.frag0 = nfrag_total;
        a.sched_base[k] = (int)(sched.size() / 4);
        dk.terms = direct_terms(t, d.dt, s->kp[k].weight, d.kernels[k].sigma, d.path);
        dk.scale = (float)(1.0 / build_fragments(t, R, dk.terms, sched, frags, dk.nfrag, dk.nrow));
        dk.src = (int)d.kernels[k].src;
        dk.dst = (int)d.kernels[k].dst;
        dk.mu = s->kp[k].mu;
        dk.inv_sigma = s->kp[k].inv_sigma;
        dk.weight = s->kp[k].weight;
        nfrag_total += dk.nfrag * (dk.terms == 3 ? 2 : 1);
        a.nfrag_max = std::max(a.nfrag_max, dk.nfrag);
    }
    // Candidates: the tile kernel at 128x128, 128x64, 64x64; for several
    // kernels the same tiles one workgroup per (tile, kernel) with a sum
    // pass (variant 2, which needs a growth plane per kernel); for one
    // channel, one kernel and one bf16 term the persistent ring kernel at
    // 64 rows (128 or 64 columns).
    using Shape = ah_direct_tile;
    std::vector<Shape> shapes;
    const bool ring_ok = K == 1 && d.channels == 1 && a.k[0].terms == 1;
    for (auto cr : {std::make_pair(128, 128), std::make_pair(128, 64), std::make_pair(64, 64)}) {
        if (cr.first > (int)d.width || cr.second > (int)d.height) continue;
        shape(cr.first, cr.second, 0);
        if (direct_lds_bytes(a) > 160 * 1024) continue;
        shapes.push_back({(uint32_t)cr.first, (uint32_t)cr.second, 0});
        if (K > 1) shapes.push_back({(uint32_t)cr.first, (uint32_t)cr.second, 2});
    }
    if (K > 1) {
        a.gsum = (float*)s->alloc(sizeof(float) * s->cells() * K);
        // Slices of the costliest kernels (MFMAs per tile) go first, so the
        // last workgroups to start are the cheap ones.
        std::vector<int> order(K);
        for (int k = 0; k < K; ++k) order[k] = k;
        auto cost = [&](int k) {
            const DirectKernel& dk = a.k[k];
            return (3 * dk.nrow[0] + 2 * dk.nrow[1] + dk.nrow[2]) * (dk.terms == 3 ? 3 : 1);
        };
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
        for (int k = 0; k < K; ++k) a.split_order[k] = order[k];
    }
    for (int cols : {128, 64}) {
        if (!ring_ok || cols > (int)d.width || d.width % cols || d.height % 64) continue;
        shape(cols, 64, 1);
        if (a.tc <= 256 && direct_lds_bytes(a) <= 160 * 1024) shapes.push_back({(uint32_t)cols, 64, 1});
    }
    s->direct_tiles = shapes;
    if (shapes.empty())
        throw Error{AH_ERR_UNSUPPORTED, "direct BF16 path: kernel tile and fragments exceed 160 KiB of LDS"};
    int4* dsched = (int4*)s->alloc(sizeof(int) * sched.size());
    uint4* dfrags = (uint4*)s->alloc(sizeof(uint16_t) * frags.size());
    AH_HIP(hipMemcpy(dsched, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice));
    AH_HIP(hipMemcpy(dfrags, frags.data(), sizeof(uint16_t) * frags.size(), hipMemcpyHostToDevice));
    a.sched = dsched;
    a.frags = dfrags;
    a.dt = d.dt;
    s->state_alt = (float*)s->alloc(sizeof(float) * s->cells() * d.channels);
    // Tile shape: the largest that fits LDS when it is the only candidate;
    // otherwise each admissible shape is timed on this world (zeroed states,
    // the work is data-independent) and the fastest kept.  Small worlds can
    // prefer small tiles: more workgroups than CUs.
    shape(shapes[0].cols, shapes[0].rows, shapes[0].variant);
    if (shapes.size() > 1) {
        const size_t bytes = sizeof(float) * s->cells() * d.channels;
        AH_HIP(hipMemsetAsync(s->state, 0, bytes, s->stream));
        AH_HIP(hipMemsetAsync(s->state_alt, 0, bytes, s->stream));
        hipEvent_t e0, e1;
        AH_HIP(hipEventCreate(&e0));
        AH_HIP(hipEventCreate(&e1));
        double best = 1e300;
        Shape pick = shapes[0];
        try {
            for (const Shape& cr : shapes) {
                shape(cr.cols, cr.rows, cr.variant);
                DirectArgs t = a;
                t.state_in = s->state;
                t.state_out = s->state_alt;
                launch_direct(s->stream, nullptr, t, (int)d.channels);  // warm-up
                AH_HIP(hipEventRecord(e0, s->stream));
                for (int r = 0; r < 3; ++r) launch_direct(s->stream, nullptr, t, (int)d.channels);
                AH_HIP(hipEventRecord(e1, s->stream));
                AH_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                AH_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    pick = cr;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        shape(pick.cols, pick.rows, pick.variant);
    }
}

void setup(ah_sim* s) {
    const ah_lenia_desc& d = s->d;
    const int W = (int)d.width, H = (int)d.height, C = (int)d.channels, K = (int)d.n_kernels;
    const size_t cells = s->cells();
    AH_HIP(hipGetDevice(&s->device));
    AH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    hipStream_t st = s->stream;

    s->state = (float*)s->alloc(sizeof(float) * cells * C);

    // Normalised growth parameters.
    float wsum[AH_MAX_CHANNELS] = {0, 0, 0};
    for (int k = 0; k < K; ++k) wsum[d.kernels[k].dst] += d.kernels[k].weight;
    for (int k = 0; k < K; ++k) {
        const ah_kernel& kd = d.kernels[k];
        if (d.rule == AH_RULE_LENIA) s->kp[k] = KernelParams{kd.mu, 1.0f / kd.sigma, kd.weight / wsum[kd.dst], kd.src, kd.dst};
        else s->kp[k] = KernelParams{0.0f, 0.0f, 0.0f, kd.src, kd.dst};
        s->ksrc[k] = (int)kd.src;
        s->reach[k] = kernel_reach(kd);
    }

    // FP64 taps of every kernel (scratch for set-up).
    int max_reach = 0;
    for (int k = 0; k < K; ++k) max_reach = std::max(max_reach, s->reach[k]);
    const int max_side = 2 * max_reach + 1;
    std::vector<double*> taps(K);
    double* tap_sums = nullptr;
    std::vector<void*> scratch;
    auto tmp = [&](size_t bytes) {
        void* p = nullptr;
        AH_HIP(hipMalloc(&p, bytes));
        scratch.push_back(p);
        return p;
    };
    try {
        tap_sums = (double*)tmp(sizeof(double) * K);
        for (int k = 0; k < K; ++k) {
            const int side = 2 * s->reach[k] + 1;
            taps[k] = (double*)tmp(sizeof(double) * side * side);
            launch_kernel_taps64(st, &s->prof, d.kernels[k], s->reach[k], taps[k], tap_sums + k);
        }
        (void)max_side;

        if (d.path == AH_PATH_FFT_FP32) {
            const size_t spec_elems = (size_t)(W / 2) * H;  // complex per channel / kernel
            s->spec_src = (float2*)s->alloc(sizeof(float2) * spec_elems * C);
            s->state_alt = (float*)s->alloc(sizeof(float) * s->cells() * C);
            bool alias = (K == 1 && C == 1);
            s->spec_k = alias ? s->spec_src : (float2*)s->alloc(sizeof(float2) * spec_elems * K);
            const size_t np = (size_t)H / 2 + 1;
            const int klayout = khat_layout(W, H, C, K, s->ksrc);
            const size_t kfloats = khat_floats(W, H, klayout);
            s->khat = (float*)s->alloc(sizeof(float) * kfloats * K);
            s->kpack = (float2*)s->alloc(sizeof(float2) * np * K);
            AH_HIP(hipMemsetAsync(s->khat, 0, sizeof(float) * kfloats * K, st));
            s->twW = (float2*)s->alloc(sizeof(float2) * W);
            s->twH = (float2*)s->alloc(sizeof(float2) * H);
            launch_twiddles(st, &s->prof, s->twW, W);
            launch_twiddles(st, &s->prof, s->twH, H);
            double* cosW = (double*)tmp(sizeof(double) * W);
            double* cosH = (double*)tmp(sizeof(double) * H);
            launch_cos_table64(st, &s->prof, cosW, W);
            launch_cos_table64(st, &s->prof, cosH, H);
            for (int k = 0; k < K; ++k) {
                const int side = 2 * s->reach[k] + 1;
                double* T = (double*)tmp(sizeof(double) * side * (W / 2 + 1));
                launch_khat64(st, &s->prof, W, H, s->reach[k], taps[k], cosW, cosH, T, s->khat + (size_t)k * kfloats,
                              s->kpack + (size_t)k * np, klayout);
            }
        } else {
            setup_direct(s, taps);
        }

        // Seed.
        uint32_t* corners = nullptr;
        if (d.init == AH_INIT_BLOBS && d.init_count > 0) {
            std::vector<uint32_t> host(2 * d.init_count);
            uint64_t rs = s->seed;
            for (uint32_t i = 0; i < d.init_count; ++i) {
                host[2 * i] = (uint32_t)(splitmix64(rs) % d.width);
                host[2 * i + 1] = (uint32_t)(splitmix64(rs) % d.height);
            }
            corners = (uint32_t*)tmp(sizeof(uint32_t) * host.size());
            AH_HIP(hipMemcpyAsync(corners, host.data(), sizeof(uint32_t) * host.size(), hipMemcpyHostToDevice, st));
        }
        launch_seed(st, &s->prof, d, s->seed, corners, s->state);

        s->stat_blocks = stats_blocks(cells);
        size_t np = (size_t)s->stat_blocks * C;
        s->pmass = (double*)s->alloc(sizeof(double) * np);
        s->pmin = (float*)s->alloc(sizeof(float) * np);
        s->pmax = (float*)s->alloc(sizeof(float) * np);
        s->pnf = (unsigned long long*)s->alloc(sizeof(unsigned long long) * np);

        AH_HIP(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        for (void* p : scratch) (void)hipFree(p);
        throw;
    }
    for (void* p : scratch) (void)hipFree(p);
}

// One rule step (conv_out == nullptr): reads `state`, writes `state_alt`,
// then the two swap, so the step's input stays readable as state_alt.  With
// conv_out set, only kernel conv_k's potential of `input` is written there
// and the state is left alone.
void step_fft(ah_sim* s, float* conv_out, int conv_k, const float* input) {
    const ah_lenia_desc& d = s->d;
    const int W = (int)d.width, H = (int)d.height, C = (int)d.channels, K = (int)d.n_kernels;
    // Pass 1 runs only when the state changed outside a step; otherwise the
    // previous step's pass 3 already left the spectra of the current state.
    if (conv_out || !s->spec_ready) launch_row_r2c(s->stream, &s->prof, W, H, C, input, s->spec_src, s->twW);
    // The column pass consumes spec_src (in place when it aliases spec_k).
    s->spec_ready = false;
    launch_col_conv(s->stream, &s->prof, W, H, C, s->spec_src, s->spec_k, s->khat, s->kpack, s->twH, s->ksrc, K);
    launch_row_c2r(s->stream, &s->prof, W, H, C, K, s->spec_k, input, s->state_alt, s->kp, d.dt, s->twW, conv_out,
                   conv_k, s->spec_src, (int)d.rule);
    if (!conv_out) {
        std::swap(s->state, s->state_alt);
        s->spec_ready = true;
    }
}

void step_direct(ah_sim* s, float* conv_out, int conv_k, const float* input) {
    DirectArgs a = s->direct;
    a.state_in = input;
    a.state_out = s->state_alt;
    a.conv_out = conv_out;
    a.conv_k = conv_k;
    launch_direct(s->stream, &s->prof, a, (int)s->d.channels);
    if (!conv_out) std::swap(s->state, s->state_alt);
}

void step_once(ah_sim* s, float* conv_out = nullptr, int conv_k = 0, const float* input = nullptr) {
    if (!input) input = s->state;
    if (s->d.path == AH_PATH_FFT_FP32) step_fft(s, conv_out, conv_k, input);
    else step_direct(s, conv_out, conv_k, input);
    if (!conv_out) s->stepped = true;
}


}  // namespace

namespace ah {
// Callers may write through the returned pointer, so the spectra of the
// previous state are dropped.
float* sim_state(ah_sim* s) {
    s->spec_ready = false;
    return s->state;
}
hipStream_t sim_stream(ah_sim* s) { return s->stream; }
const ah_lenia_desc& sim_desc(const ah_sim* s) { return s->d; }
void sim_advance(ah_sim* s, uint32_t steps) {
    for (uint32_t i = 0; i < steps; ++i) step_once(s);
    s->generation += steps;
}
void sim_add_generations(ah_sim* s, uint64_t n) { s->generation += n; }
}  // namespace ah

extern "C" {

const char* ah_last_error(void) { return ah::g_error.c_str(); }

ah_status ah_lenia_build(const char* id, uint32_t width, uint32_t height, ah_lenia_desc* out) {
    return guard([&] {
        if (!id || !out) throw Error{AH_ERR_INVALID, "null argument"};
        // automata::fit_size rounds to powers of two (mod.rs:51-61).
        auto p2 = [](uint32_t v) {
            uint32_t r = 1;
            while (r < v) r <<= 1;
            return r;
        };
        ah_lenia_desc d{};
        d.width = p2(width);
        d.height = p2(height);
        d.channels = 1;
        d.path = AH_PATH_FFT_FP32;
        auto ring = [](float lo, float hi) {
            ah_kernel k{};
            k.shape = AH_SHAPE_BAND;
            k.inner = lo;
            k.radius = hi;
            k.weight = 1.0f;
            return k;
        };
        if (std::strcmp(id, "lenia") == 0) {
            // lenia(w, h, 13.0) (spectral.rs:286-352,489).
            d.n_kernels = 1;
            d.kernels[0] = ah_kernel{13.0f, 1, {1.0f, 0.0f, 0.0f}, 0.15f, 0.015f, 1.0f, 0, 0, AH_SHAPE_LENIA, 0.0f, 0.0f};
            d.dt = 0.1f;
            d.init = AH_INIT_ORBIUM;
            d.rule = AH_RULE_LENIA;
        } else if (std::strcmp(id, "smoothlife") == 0) {
            // smoothlife(w, h, 21.0) (spectral.rs:354-413,499): ri = ra/3, aa = 0.5.
            const float ra = 21.0f, ri = ra / 3.0f, aa = 0.5f;
            d.n_kernels = 2;
            d.kernels[0] = ah_kernel{};
            d.kernels[0].shape = AH_SHAPE_SMOOTH_DISC;
            d.kernels[0].radius = ri;
            d.kernels[0].aa = aa;
            d.kernels[0].weight = 1.0f;
            d.kernels[1] = ah_kernel{};
            d.kernels[1].shape = AH_SHAPE_SMOOTH_RING;
            d.kernels[1].radius = ra;
            d.kernels[1].inner = ri;
            d.kernels[1].aa = aa;
            d.kernels[1].weight = 1.0f;
            d.dt = 0.1f;
            d.init = AH_INIT_DISCS;
            d.rule = AH_RULE_SMOOTHLIFE;
        } else if (std::strcmp(id, "mnca") == 0 || std::strncmp(id, "mnca.", 5) == 0) {
            // mnca(w, h) = mnca_shaped(w, h, [12, 18, 3, 8]) (spectral.rs:415-428);
            // "mnca.i0,o0,i1,o1" gives the rings (mod.rs:86-92).
            float v[4] = {12.0f, 18.0f, 3.0f, 8.0f};
            if (id[4] == '.') {
                const char* p = id + 5;
                for (int i = 0; i < 4; ++i) {
                    char* end = nullptr;
                    v[i] = std::strtof(p, &end);
                    if (end == p || (i < 3 ? *end != ',' : *end != 0))
                        throw Error{AH_ERR_INVALID, std::string("mnca ids are mnca.i0,o0,i1,o1: ") + id};
                    p = end + 1;
                }
            }
            d.n_kernels = 2;
            d.kernels[0] = ring(v[0], v[1]);
            d.kernels[1] = ring(v[2], v[3]);
            d.dt = 1.0f;  // unused by the rule
            d.init = AH_INIT_NOISE;
            d.rule = AH_RULE_MNCA;
        } else {
            throw Error{AH_ERR_INVALID, std::string("unknown automaton id: ") + id};
        }
        *out = d;
    });
}

ah_status ah_sim_new(const ah_lenia_desc* desc, uint32_t seed, ah_sim** out) {
    return guard([&] {
        if (!desc || !out) throw Error{AH_ERR_INVALID, "null argument"};
        *out = nullptr;
        validate(*desc);
        std::unique_ptr<ah_sim> s(new ah_sim());
        s->d = *desc;
        s->seed = seed;
        setup(s.get());
        *out = s.release();
    });
}

ah_status ah_sim_step(ah_sim* s, uint32_t steps, uint64_t* generations) {
    return guard([&] {
        if (!s) throw Error{AH_ERR_INVALID, "null sim"};
        AH_HIP(hipSetDevice(s->device));
        for (uint32_t i = 0; i < steps; ++i) step_once(s);
        s->generation += steps;
        if (generations) *generations = steps;
    });
}

ah_status ah_sim_wait(ah_sim* s) {
    return guard([&] {
        if (!s) throw Error{AH_ERR_INVALID, "null sim"};
        AH_HIP(hipStreamSynchronize(s->stream));
    });
}

ah_status ah_sim_read_field(ah_sim* s, const char* name, void* dst, size_t bytes) {
    return guard([&] {
        if (!s || !name || !dst) throw Error{AH_ERR_INVALID, "null argument"};
        AH_HIP(hipSetDevice(s->device));
        const size_t plane = sizeof(float) * s->cells();
        if (std::strcmp(name, "state") == 0) {
            if (bytes != plane * s->d.channels) throw Error{AH_ERR_SIZE, "state is channels*height*width*4 bytes"};
            AH_HIP(hipMemcpyAsync(dst, s->state, bytes, hipMemcpyDeviceToHost, s->stream));
            AH_HIP(hipStreamSynchronize(s->stream));
            return;
        }
        // "conv<k>": the reference's conv_i (spectral.rs:227-234), which each
        // step writes: kernel k's potential of that step's input, zero before
        // the first step; a state write leaves it alone.  It is recomputed
        // here from the last step's input, which stepping out of place keeps.
        // "potential<k>": kernel k's potential of the current state.
        const bool conv = std::strncmp(name, "conv", 4) == 0;
        const bool pot = std::strncmp(name, "potential", 9) == 0;
        const char* num = conv ? name + 4 : pot ? name + 9 : nullptr;
        if (num && num[0] >= '0' && num[0] <= '9' && num[1] == 0) {
            int k = num[0] - '0';
            if (k >= (int)s->d.n_kernels) throw Error{AH_ERR_UNKNOWN_FIELD, std::string("no kernel for ") + name};
            if (bytes != plane) throw Error{AH_ERR_SIZE, "conv and potential fields are height*width*4 bytes"};
            if (conv && !s->stepped) {
                std::memset(dst, 0, bytes);
                return;
            }
            float* u = nullptr;
            AH_HIP(hipMallocAsync((void**)&u, plane, s->stream));
            step_once(s, u, k, conv ? s->state_alt : s->state);
            hipError_t e = hipMemcpyAsync(dst, u, bytes, hipMemcpyDeviceToHost, s->stream);
            (void)hipFreeAsync(u, s->stream);
            AH_HIP(e);
            AH_HIP(hipStreamSynchronize(s->stream));
            return;
        }
        throw Error{AH_ERR_UNKNOWN_FIELD, std::string("unknown field: ") + name};
    });
}

ah_status ah_sim_write_field(ah_sim* s, const char* name, const void* src, size_t bytes) {
    return guard([&] {
        if (!s || !name || !src) throw Error{AH_ERR_INVALID, "null argument"};
        AH_HIP(hipSetDevice(s->device));
        if (std::strcmp(name, "state") != 0) throw Error{AH_ERR_UNKNOWN_FIELD, std::string("not writable: ") + name};
        if (bytes != sizeof(float) * s->cells() * s->d.channels)
            throw Error{AH_ERR_SIZE, "state is channels*height*width*4 bytes"};
        s->spec_ready = false;
        AH_HIP(hipMemcpyAsync(s->state, src, bytes, hipMemcpyHostToDevice, s->stream));
        AH_HIP(hipStreamSynchronize(s->stream));
    });
}

ah_status ah_sim_render(ah_sim* s, uint32_t out_w, uint32_t out_h, float center_x, float center_y, float scale,
                        uint8_t* rgba, size_t bytes) {
    return guard([&] {
        if (!s || !rgba) throw Error{AH_ERR_INVALID, "null argument"};
        if (out_w == 0 || out_h == 0 || out_w > 16384 || out_h > 16384)
            throw Error{AH_ERR_INVALID, "output size must be 1..16384"};
        if (!(scale > 0.0f) || !std::isfinite(scale) || !std::isfinite(center_x) || !std::isfinite(center_y))
            throw Error{AH_ERR_INVALID, "camera must be finite with a positive scale"};
        const size_t n = (size_t)out_w * out_h * 4;
        if (bytes != n) throw Error{AH_ERR_SIZE, "rgba buffer is out_w*out_h*4 bytes"};
        AH_HIP(hipSetDevice(s->device));
        uint32_t* d = nullptr;
        AH_HIP(hipMallocAsync((void**)&d, n, s->stream));
        launch_render(s->stream, &s->prof, s->state, (int)s->d.width, (int)s->d.height, s->d.rule, center_x,
                      center_y, scale, (int)out_w, (int)out_h, d);
        hipError_t e = hipMemcpyAsync(rgba, d, n, hipMemcpyDeviceToHost, s->stream);
        (void)hipFreeAsync(d, s->stream);
        AH_HIP(e);
        AH_HIP(hipStreamSynchronize(s->stream));
    });
}

ah_status ah_sim_stats(ah_sim* s, ah_stats* out) {
    return guard([&] {
        if (!s || !out) throw Error{AH_ERR_INVALID, "null argument"};
        AH_HIP(hipSetDevice(s->device));
        const int C = (int)s->d.channels, nb = s->stat_blocks;
        launch_stats(s->stream, &s->prof, s->state, s->cells(), C, s->pmass, s->pmin, s->pmax, s->pnf, nb);
        ah_stats r{};
        for (int c = 0; c < C; ++c) {
            size_t o = (size_t)c * nb;
            unsigned long long nf = 0;
            AH_HIP(hipMemcpyAsync(&r.mass[c], s->pmass + o, sizeof(double), hipMemcpyDeviceToHost, s->stream));
            AH_HIP(hipMemcpyAsync(&r.min[c], s->pmin + o, sizeof(float), hipMemcpyDeviceToHost, s->stream));
            AH_HIP(hipMemcpyAsync(&r.max[c], s->pmax + o, sizeof(float), hipMemcpyDeviceToHost, s->stream));
            AH_HIP(hipMemcpyAsync(&nf, s->pnf + o, sizeof(nf), hipMemcpyDeviceToHost, s->stream));
            AH_HIP(hipStreamSynchronize(s->stream));
            r.nonfinite += nf;
        }
        r.generation = s->generation;
        *out = r;
    });
}

ah_status ah_sim_check(ah_sim* s) {
    ah_stats st{};
    const ah_status e = ah_sim_stats(s, &st);
    if (e != AH_OK) return e;
    return guard([&] {
        if (st.nonfinite)
            throw Error{AH_ERR_DIVERGED, "state holds " + std::to_string(st.nonfinite) + " non-finite cells at generation " +
                                             std::to_string(st.generation)};
    });
}

ah_status ah_sim_info(const ah_sim* s, ah_info* out) {
    return guard([&] {
        if (!s || !out) throw Error{AH_ERR_INVALID, "null argument"};
        ah_info i{};
        i.width = s->d.width;
        i.height = s->d.height;
        i.channels = s->d.channels;
        i.n_kernels = s->d.n_kernels;
        i.path = s->d.path;
        i.cells_per_step = (uint64_t)s->cells();
        i.generations_per_step = 1;
        i.total_bytes = s->total_bytes;
        i.generation = s->generation;
        *out = i;
    });
}

void ah_sim_destroy(ah_sim* s) { delete s; }

ah_status ah_sim_direct_tile(const ah_sim* s, ah_direct_tile* out) {
    return guard([&] {
        if (!s || !out) throw Error{AH_ERR_INVALID, "null argument"};
        if (s->d.path == AH_PATH_FFT_FP32) throw Error{AH_ERR_INVALID, "FFT sims have no direct-path tile"};
        *out = ah_direct_tile{(uint32_t)s->direct.cols, (uint32_t)s->direct.rows, (uint32_t)s->direct.variant};
    });
}

ah_status ah_sim_set_direct_tile(ah_sim* s, const ah_direct_tile* t) {
    return guard([&] {
        if (!s || !t) throw Error{AH_ERR_INVALID, "null argument"};
        if (s->d.path == AH_PATH_FFT_FP32) throw Error{AH_ERR_INVALID, "FFT sims have no direct-path tile"};
        for (const ah_direct_tile& c : s->direct_tiles)
            if (c.cols == t->cols && c.rows == t->rows && c.variant == t->variant) {
                AH_HIP(hipStreamSynchronize(s->stream));
                direct_shape(s->direct, (int)c.cols, (int)c.rows, (int)c.variant);
                return;
            }
        throw Error{AH_ERR_UNSUPPORTED, "direct tile " + std::to_string(t->cols) + "x" + std::to_string(t->rows) +
                                            " variant " + std::to_string(t->variant) +
                                            " is not admissible for this world"};
    });
}

ah_status ah_sim_profile(ah_sim* s, int enable) {
    return guard([&] {
        if (!s) throw Error{AH_ERR_INVALID, "null sim"};
        AH_HIP(hipStreamSynchronize(s->stream));
        if (enable < 0) throw Error{AH_ERR_INVALID, "enable must be >= 0"};
        s->prof.reset();
        s->prof.enabled = enable != 0;
        s->prof.every = enable > 1 ? enable : 1;
    });
}

ah_status ah_sim_profile_read(ah_sim* s, ah_kernel_time* out, uint32_t capacity, uint32_t* count) {
    return guard([&] {
        if (!s || !count || (capacity && !out)) throw Error{AH_ERR_INVALID, "null argument"};
        AH_HIP(hipStreamSynchronize(s->stream));
        std::vector<ah_kernel_time> agg;
        for (auto& r : s->prof.recs) {
            float ms = 0.0f;
            AH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            auto it = std::find_if(agg.begin(), agg.end(),
                                   [&](const ah_kernel_time& t) { return std::strcmp(t.name, r.name) == 0; });
            if (it == agg.end()) {
                ah_kernel_time t{};
                std::snprintf(t.name, sizeof(t.name), "%s", r.name);
                agg.push_back(t);
                it = agg.end() - 1;
            }
            it->launches += 1;
            it->total_ms += ms;
        }
        uint32_t n = (uint32_t)std::min<size_t>(agg.size(), capacity);
        for (uint32_t i = 0; i < n; ++i) out[i] = agg[i];
        *count = (uint32_t)agg.size();
    });
}

ah_status ah_device_name(char* out, size_t capacity) {
    return guard([&] {
        if (!out || capacity == 0) throw Error{AH_ERR_INVALID, "null or empty buffer"};
        int dev = 0;
        AH_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        AH_HIP(hipGetDeviceProperties(&p, dev));
        std::string name = std::string(p.name) + " (" + p.gcnArchName + ")";
        std::snprintf(out, capacity, "%s", name.c_str());
    });
}

ah_status ah_copy_probe(size_t bytes, uint32_t iters, double* ms_per_copy) {
    void* a = nullptr;
    void* b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ah_status r = guard([&] {
        if (!ms_per_copy || bytes < 16 || bytes % 16 || iters == 0) throw Error{AH_ERR_INVALID, "bytes must be a multiple of 16, iters > 0"};
        AH_HIP(hipMalloc(&a, bytes));
        AH_HIP(hipMalloc(&b, bytes));
        AH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        AH_HIP(hipEventCreate(&e0));
        AH_HIP(hipEventCreate(&e1));
        AH_HIP(hipMemsetAsync(a, 0, bytes, st));
        const size_t n4 = bytes / 16;
        launch_copy_probe(st, (const float4*)a, (float4*)b, n4);
        AH_HIP(hipEventRecord(e0, st));
        for (uint32_t i = 0; i < iters; ++i) launch_copy_probe(st, (const float4*)a, (float4*)b, n4);
        AH_HIP(hipEventRecord(e1, st));
        AH_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        AH_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = (double)ms / iters;
    });
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (st) (void)hipStreamDestroy(st);
    return r;
}

}  // extern "C"
