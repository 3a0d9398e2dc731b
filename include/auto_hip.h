/*
 * auto_hip.h -- C ABI of the MI355X (gfx950) Lenia world-step engine.
 *
 * The engine replaces the Lenia path of auto-core: the registry entry that
 * builds the Lenia program (crates/auto-core/src/automata/mod.rs:63-95,
 * crates/auto-core/src/automata/spectral.rs:286-352,482-490) and the `Sim`
 * object that runs it (crates/auto-core/src/sim.rs:202-309,559-595).  The
 * reference describes a rule as WGSL text; a HIP engine cannot consume that,
 * so the rule crosses this boundary as data (ah_lenia_desc).
 *
 * Conventions
 *   - Every entry point returns ah_status; AH_OK is 0.  On failure the
 *     message is available from ah_last_error() (thread-local), mirroring the
 *     reference's Result<_, String> (sim.rs:202) and replacing its panics on
 *     invalid descriptors (spectral.rs:103-116) with AH_ERR_INVALID.
 *   - The engine owns all device memory; callers own host buffers.
 *   - State is float32, row-major [channel][y][x], little-endian, exactly as
 *     read_field/write_field move it in the reference (sim.rs:559-595).
 *   - Stepping is asynchronous on the engine's own HIP stream; read_field,
 *     stats and wait synchronise.  One handle must not be used from two
 *     threads at once; distinct handles may be used from distinct threads.
 *   - No exceptions cross this boundary.
 */
#ifndef AUTO_HIP_H
#define AUTO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AH_OK = 0,
    AH_ERR_INVALID = 1,       /* bad descriptor or argument            */
    AH_ERR_DEVICE = 2,        /* HIP runtime or launch failure         */
    AH_ERR_ALLOC = 3,         /* device allocation failed              */
    AH_ERR_UNKNOWN_FIELD = 4, /* read_field/write_field name unknown   */
    AH_ERR_SIZE = 5,          /* host buffer size does not match       */
    AH_ERR_UNSUPPORTED = 6,   /* valid request this engine cannot serve */
    AH_ERR_DIVERGED = 7       /* surrogate judge exceeded its bound    */
} ah_status;

enum { AH_MAX_CHANNELS = 3, AH_MAX_KERNELS = 9, AH_MAX_RINGS = 3 };

/* How the periodic convolution is computed.  The direct paths need width
 * and height >= 64 and kernel radius <= 31. */
typedef enum {
    AH_PATH_FFT_FP32 = 0,       /* fused FP32 real FFT, two passes per step              */
    AH_PATH_DIRECT_BF16 = 1,    /* direct convolution on BF16 MFMA, FP32 accumulate; the
                                   engine picks 1 or 3 bf16 terms per kernel (below)      */
    AH_PATH_DIRECT_BF16_X1 = 2, /* force one term: bf16(K) * bf16(s)                      */
    AH_PATH_DIRECT_BF16_X3 = 3, /* force three terms: K_hi s_hi + K_hi s_lo + K_lo s_hi,
                                   with x = x_hi + x_lo split into two bf16 values        */
    AH_PATH_AUTO = 4            /* batched worlds only: the batch picks FFT or direct per
                                   world (see ah_batch_new)                              */
} ah_path;
/*
 * AH_PATH_DIRECT_BF16 chooses per kernel: one term when the growth slope
 * times the bf16 error bound stays within 1e-3, i.e.
 *   1.213 * dt * h / sigma * 0.0078 / sqrt(N_eff) <= 1e-3,  N_eff = 1 / sum(K^2),
 * else three terms (about FP32 accuracy at three times the MFMA work).
 */

/* How the state is seeded at creation. */
typedef enum {
    AH_INIT_ORBIUM = 0, /* reference Orbium lattice (spectral.rs:313-335)             */
    AH_INIT_PATCH = 1,  /* centred init_size^2 patch, unit_f(rand3(x,y,seed+c))       */
    AH_INIT_BLOBS = 2,  /* init_count blobs of init_size^2, corners from SplitMix64    */
    AH_INIT_ZERO = 3,
    AH_INIT_DISCS = 4,  /* SmoothLife seed: a radius-24 disc of 1s in each 96x96 tile
                           whose rand3(tx,ty,seed) is odd (spectral.rs:381-391)       */
    AH_INIT_NOISE = 5   /* MNCA seed: 1 where unit_f(rand3(x/3,y/3,seed)) < 0.32
                           (spectral.rs:456-460)                                      */
} ah_init;

/* Radial profile of a kernel, as a function of the distance r in cells. */
typedef enum {
    AH_SHAPE_LENIA = 0,       /* `rings` exp-bump rings over [0, radius) with weights beta
                                 (spectral.rs:288-303; Chan's multi-ring form)          */
    AH_SHAPE_SMOOTH_DISC = 1, /* clamp((radius + aa - r) / (2 aa), 0, 1): SmoothLife's
                                 inner disc (spectral.rs:357-358)                       */
    AH_SHAPE_SMOOTH_RING = 2, /* clamp((radius + aa - r) / (2 aa), 0, 1) *
                                 clamp((r - inner + aa) / (2 aa), 0, 1): its outer ring
                                 (spectral.rs:359-361)                                  */
    AH_SHAPE_BAND = 3         /* 1 for inner <= r < radius, r = sqrt in float32: an MNCA
                                 ring (spectral.rs:432-440)                             */
} ah_shape;

/* What a step does with the kernels' potentials. */
typedef enum {
    AH_RULE_LENIA = 0,      /* s += dt * sum_k h_k (2 exp(-((u_k - mu_k)/sigma_k)^2 / 2) - 1),
                               clipped to [0, 1] (spectral.rs:337-345)                  */
    AH_RULE_SMOOTHLIFE = 1, /* Rafler's transition on u_0 (inner) and u_1 (outer) with
                               B1=.278 B2=.365 D1=.267 D2=.445 AN=.028 AM=.147
                               (spectral.rs:393-405); FFT path, one channel, 2 kernels */
    AH_RULE_MNCA = 2        /* threshold bands on u_0 and u_1, each overruling the one
                               before (spectral.rs:462-473); FFT path, one channel,
                               2 kernels                                                */
} ah_rule;

/*
 * One convolution kernel: a radial profile (`shape`), normalised to sum 1 on
 * the torus as the reference divides by its box sum (spectral.rs:75-83,
 * 150-171).  For AH_SHAPE_LENIA: `rings` exp-bump rings with ring weights
 * beta (rings = 1, beta = {1} is exactly the reference's core,
 * spectral.rs:288-303); mu, sigma and weight give its Gaussian growth and
 * mixing weight into `dst` under AH_RULE_LENIA.  The reach of a kernel (the
 * largest |dx|, |dy| it touches) is ceil(radius) for LENIA and BAND and
 * ceil(radius + aa) for the smooth shapes; 2*reach < min(width, height).
 */
typedef struct {
    float radius;             /* R (LENIA), outer radius (SMOOTH_*, BAND)     */
    uint32_t rings;           /* 1..AH_MAX_RINGS (LENIA)                      */
    float beta[AH_MAX_RINGS]; /* ring weights (LENIA)                         */
    float mu, sigma;          /* growth centre and width (RULE_LENIA)         */
    float weight;             /* mixing weight h; normalised per dst at create */
    uint32_t src, dst;        /* source and destination channel               */
    uint32_t shape;           /* ah_shape                                     */
    float inner;              /* inner radius (SMOOTH_RING, BAND)             */
    float aa;                 /* antialiasing half-width (SMOOTH_*), > 0      */
} ah_kernel;

typedef struct {
    uint32_t width, height;   /* powers of two, 16..8192                      */
    uint32_t channels;        /* 1..AH_MAX_CHANNELS                           */
    uint32_t n_kernels;       /* 1..AH_MAX_KERNELS                            */
    ah_kernel kernels[AH_MAX_KERNELS];
    float dt;                 /* time step                                    */
    uint32_t path;            /* ah_path                                      */
    uint32_t init;            /* ah_init                                      */
    uint32_t init_size;       /* patch side (PATCH) or blob side (BLOBS)      */
    uint32_t init_count;      /* blob count (BLOBS)                           */
    uint32_t rule;            /* ah_rule                                      */
} ah_lenia_desc;

typedef struct {
    uint32_t width, height, channels, n_kernels, path;
    uint64_t cells_per_step;       /* program.rs:117                      */
    uint64_t generations_per_step; /* program.rs:119                      */
    uint64_t total_bytes;          /* device bytes held (program.rs:123)  */
    uint64_t generation;           /* generations advanced so far         */
} ah_info;

/* Mass and range of the current state: FP64 sums, per channel. */
typedef struct {
    double mass[AH_MAX_CHANNELS];
    float min[AH_MAX_CHANNELS];
    float max[AH_MAX_CHANNELS];
    uint64_t nonfinite;            /* NaN/Inf cells over all channels     */
    uint64_t generation;
} ah_stats;

typedef struct ah_sim ah_sim;

/*
 * Fill `out` with a registry rule, as automata::build(id, w, h) does
 * (mod.rs:63-95): sizes are rounded up to powers of two (mod.rs:51-61).
 * Known ids: "lenia" (the reference rule: R=13, one ring, mu=0.15,
 * sigma=0.015, dt=0.1, Orbium seed; spectral.rs:286-352,482-490);
 * "smoothlife" (ra=21, ri=ra/3, aa=0.5, dt=0.1, disc seed;
 * spectral.rs:354-413,492-500); "mnca" (rings [12,18) and [3,8), noise
 * seed; spectral.rs:415-480,502-510) and "mnca.<i0>,<o0>,<i1>,<o1>" with
 * rings [i0,o0) and [i1,o1) (mod.rs:86-92).
 */
ah_status ah_lenia_build(const char* id, uint32_t width, uint32_t height, ah_lenia_desc* out);

/* Sim::new (sim.rs:202-282): allocate, build kernels, seed the state. */
ah_status ah_sim_new(const ah_lenia_desc* desc, uint32_t seed, ah_sim** out);

/* Sim::step (sim.rs:297-309): enqueue `steps` steps; asynchronous. */
ah_status ah_sim_step(ah_sim* sim, uint32_t steps, uint64_t* generations);

/* Gpu::wait (gpu.rs:133-137): block until queued work is done. */
ah_status ah_sim_wait(ah_sim* sim);

/*
 * Sim::read_field / write_field (sim.rs:559-595).  Fields:
 *   "state"  float32 [channels][height][width], read/write;
 *   "conv<k>" float32 [height][width], read-only: the reference's conv_i
 *            field (spectral.rs:227-234): kernel k convolved with its source
 *            channel of the last step's input state, i.e. the potential that
 *            step used; zero before the first step; a state write leaves it
 *            unchanged.  Computed on read by the sim's path (steps write the
 *            state out of place, so their input is still held).
 *   "potential<k>" float32 [height][width], read-only (not in the reference):
 *            kernel k convolved with its source channel of the current state.
 * `bytes` must equal the field size exactly.
 */
ah_status ah_sim_read_field(ah_sim* sim, const char* name, void* dst, size_t bytes);
ah_status ah_sim_write_field(ah_sim* sim, const char* name, const void* src, size_t bytes);

/*
 * Sim::render (sim.rs:452-514) into host memory: channel 0 of the current
 * state through the rule's colour ramp (spectral.rs:346-350,407-413,473-478),
 * rgba8 [out_h][out_w][4], camera centre (center_x, center_y) in cells and
 * `scale` cells per output pixel (Camera::fit, sim.rs:52-61: scale =
 * max(width/out_w, height/out_h), centre = world centre).  Synchronises.
 */
ah_status ah_sim_render(ah_sim* sim, uint32_t out_w, uint32_t out_h, float center_x, float center_y, float scale,
                        uint8_t* rgba, size_t bytes);

/* Mass/min/max/non-finite reduction of the current state (synchronises). */
ah_status ah_sim_stats(ah_sim* sim, ah_stats* out);

/*
 * Failure check: the same reduction, AH_ERR_DIVERGED when any cell of the
 * current state is NaN or infinite (the message gives the count and the
 * generation), AH_OK otherwise.  No counterpart in the reference, whose
 * Lenia has no failure detection.  A step maps a non-finite input to 0
 * (the clip's fmax/fmin return the non-NaN operand, and a NaN reaches every
 * potential through the transform), so the check belongs after a state
 * write and before stepping it.
 */
ah_status ah_sim_check(ah_sim* sim);

/*
 * The reduction of ah_sim_stats, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc) and hold a state in device memory of their own.  `state` is a
 * device pointer to float32 [channels][cells], 16-byte aligned; cells is a
 * positive multiple of 4; channels 1..3 (AH_ERR_INVALID otherwise).  The
 * partial results go to a fixed library-owned device array of 72 KiB per
 * GPU (at most 1024 partials per channel, 3 channels), kept until the
 * process ends, instead of a sim's buffers; the results equal ah_sim_stats
 * on the same state bit for bit.  `stream` is a hipStream_t (NULL: the null
 * stream); the call synchronises it.  Calls on one GPU share the array:
 * calls that overlap on one GPU race.  out->generation is 0.
 */
ah_status ah_state_stats_noalloc(const float* state, uint64_t cells, uint32_t channels, void* stream,
                                 ah_stats* out);

ah_status ah_sim_info(const ah_sim* sim, ah_info* out);

void ah_sim_destroy(ah_sim* sim);

/* Message of the last failed call on this thread ("" if none). */
const char* ah_last_error(void);

/*
 * Kernel timing.  While enabled, kernels the sim launches are bracketed by
 * HIP events on the sim's stream; ah_sim_profile_read returns per-kernel
 * counts of timed launches and their summed device milliseconds (it
 * synchronises).  enable = 1 times every launch; enable = N > 1 times every
 * N-th launch of each kernel, which keeps the events' own cost (a few
 * microseconds per recorded pair) out of the step rate.  0 turns timing off.
 * Enabling resets the counters.  This replaces the reference's unused
 * TIMESTAMP_QUERY (gpu.rs:83) and complements the wall-clock method of
 * bench.rs:126-158.
 */
typedef struct {
    char name[40];
    uint64_t launches;
    double total_ms;
} ah_kernel_time;

ah_status ah_sim_profile(ah_sim* sim, int enable);
ah_status ah_sim_profile_read(ah_sim* sim, ah_kernel_time* out, uint32_t capacity, uint32_t* count);

/*
 * Direct-path kernel variant.  No counterpart in the reference (its only
 * path is the FFT).  A DIRECT_BF16* sim times every admissible variant on its
 * world at create and keeps the fastest, for a cols x rows output tile
 * (128x128, 128x64, 64x64): variant 0, the tile kernel (a workgroup computes
 * every kernel ending in its channel); variant 1, for one channel, one kernel
 * and one bf16 term, the persistent ring kernel (rows = 64) that carries the
 * shared halo rows from one tile to the next; variant 2, for several
 * kernels, one workgroup per (tile, kernel) writing that kernel's weighted
 * growth to a scratch plane, and a second pass that sums them per channel
 * in kernel order.  Every variant gives the same bits.
 * ah_sim_direct_tile reports the variant in use;
 * ah_sim_set_direct_tile forces one (AH_ERR_UNSUPPORTED if it is not
 * admissible for this world; AH_ERR_INVALID on an FFT sim).
 */
typedef struct {
    uint32_t cols, rows, variant;
} ah_direct_tile;

ah_status ah_sim_direct_tile(const ah_sim* sim, ah_direct_tile* out);
ah_status ah_sim_set_direct_tile(ah_sim* sim, const ah_direct_tile* tile);

/*
 * Batched uneven worlds.  No counterpart in the reference, which runs one
 * world per Sim (sim.rs:202-309).  Every world is single-channel with one
 * Lenia kernel, width and height >= 64, and an ORBIUM, PATCH or ZERO init
 * seeded with its own seed; worlds differ in size, rule, dt and path.
 * Direct BF16 worlds (radius <= 31) advance together in one persistent launch
 * per step, over tiles of every such world.  FFT worlds are
 * grouped by size; each group runs as one column-pass and one row-pass launch
 * per step, one world per channel.  AH_PATH_AUTO worlds take the FFT path when
 * their direct cost (32-column MFMA chunks summed over the kernel rows, times
 * the bf16 terms) exceeds AH_BATCH_FFT_ABOVE, or their radius exceeds 31.
 * The threshold is measured: on the 384-world benchmark batch, 30 and 0 (all
 * FFT) run within 2% of each other, 60 at 0.9x and 240 at 0.25x of it.  When
 * no world names a direct path itself and the AUTO worlds below the threshold
 * hold under 1/20 of the batch's cells, they take the FFT path too: their
 * own direct launch would cost more than they save.
 * ah_batch_step replays hipGraph-captured chunks of `graph_steps` steps and
 * runs the remainder one by one.  At creation an autotuner times every
 * admissible (tile shape, grid) pair of the direct launch, and for each FFT
 * group every admissible pair of column-pass and row-pass kernels (the
 * fused batch kernels and the generic ones), on the batch itself, and keeps
 * the fastest; the candidates and their times are reported by
 * ah_batch_info_get.
 */
typedef struct ah_batch ah_batch;

enum { AH_BATCH_MAX_CANDIDATES = 8, AH_BATCH_FFT_ABOVE = 30, AH_BATCH_MAX_GROUP_TUNE = 64 };

/* One timed FFT-group candidate: the group's world size, which column and
 * row kernels ran (0: the fused batch kernel, 1: the generic kernel) and the
 * measured time of one step of the group. */
typedef struct {
    uint32_t width, height;
    uint32_t column, row;
    double ms_per_step;
} ah_batch_group_tune;

typedef struct {
    uint32_t tile_cols;  /* output columns per tile (64 or 128)               */
    uint32_t tile_rows;  /* output rows per tile (64 or 128)                  */
    uint32_t grid;       /* persistent workgroups                             */
    double ms_per_step;  /* measured by the autotuner                         */
} ah_batch_tune;

typedef struct {
    uint32_t n_worlds;
    uint32_t tile_cols, tile_rows, grid, ntiles, graph_steps;
    uint64_t total_cells;  /* cell-updates per step */
    uint64_t total_bytes;
    uint64_t generation;
    uint32_t n_tune;
    ah_batch_tune tune[AH_BATCH_MAX_CANDIDATES];
    uint32_t fft_worlds;   /* worlds on the FFT path (the rest run direct BF16) */
    uint32_t fft_groups;   /* same-size FFT groups, two launches each per step  */
    /* FFT group tuning: every timed (column kernel, row kernel) pair of every
     * group, in group order (the first AH_BATCH_MAX_GROUP_TUNE of them); the
     * step uses each group's fastest. */
    uint32_t n_group_tune;
    ah_batch_group_tune group_tune[AH_BATCH_MAX_GROUP_TUNE];
} ah_batch_info;

ah_status ah_batch_new(const ah_lenia_desc* descs, uint32_t n_worlds, const uint32_t* seeds, ah_batch** out);
ah_status ah_batch_step(ah_batch* batch, uint32_t steps);
ah_status ah_batch_wait(ah_batch* batch);
/* One world's state, float32 [height][width]; `bytes` must match exactly. */
ah_status ah_batch_read_world(ah_batch* batch, uint32_t world, float* dst, size_t bytes);
ah_status ah_batch_write_world(ah_batch* batch, uint32_t world, const float* src, size_t bytes);
ah_status ah_batch_info_get(const ah_batch* batch, ah_batch_info* out);
/*
 * Launch-overhead measurement: runs `steps` (a multiple of graph_steps) eager
 * launches with per-launch HIP events (summed device time -> kernel_ms, host
 * wall time -> eager_wall_ms), then the same number of steps as graph
 * replays (host wall time -> graph_wall_ms).  Advances the batch 2*steps.
 */
ah_status ah_batch_measure(ah_batch* batch, uint32_t steps, double* kernel_ms, double* graph_wall_ms,
                           double* eager_wall_ms);
void ah_batch_destroy(ah_batch* batch);

/*
 * Learned k-step surrogate.  No counterpart in the reference.  A small
 * convolutional network maps a single-channel state to the state k steps
 * ahead: a 2-level encoder-decoder of circular 5x5 convolutions, widths
 * 32 -> 64 -> 32, GELU, an additive skip, residual output clipped to [0, 1]
 * (the full definition is in auto_amd/csrc/surrogate.hip).  Weights and
 * activations are bf16 with FP32 accumulation; FP32 master weights and Adam
 * state live on the device.  Sizes fed to the network need height % 8 == 0
 * and width % 128 == 0.
 */
typedef struct ah_surrogate ah_surrogate;

enum { AH_SURROGATE_PARAMS = 104129 };

typedef struct {
    uint64_t lenia_steps;      /* steps the call advanced the sim             */
    uint64_t surrogate_calls;  /* network calls whose output was kept         */
    uint64_t plain_steps;      /* plain steps (judges, fallback, remainder)   */
    uint64_t judges, fallbacks;
    double max_judge_rel_l2, last_judge_rel_l2;
    double ms;                 /* host wall time of the call                  */
} ah_surrogate_report;

/* Parameters drawn from SplitMix64(seed): uniform +-sqrt(3/fan_in), the
 * output layer scaled by 0.1, biases zero. */
ah_status ah_surrogate_new(uint32_t seed, ah_surrogate** out);
void ah_surrogate_destroy(ah_surrogate* s);
/* The flat FP32 parameter vector (W1, b1, W2, b2, W3, b3, W4, b4). */
ah_status ah_surrogate_get_params(const ah_surrogate* s, float* dst, size_t n);
ah_status ah_surrogate_set_params(ah_surrogate* s, const float* src, size_t n);
/* Network output for host states x [n][h][w] -> y [n][h][w]. */
ah_status ah_surrogate_forward(ah_surrogate* s, const float* x, uint32_t n, uint32_t h, uint32_t w, float* y);
/* MSE loss against target and, if grad is set, its gradient (n_params floats). */
ah_status ah_surrogate_loss_grad(ah_surrogate* s, const float* x, const float* target, uint32_t n, uint32_t h,
                                 uint32_t w, double* loss, float* grad);
/* Training pairs (x_t, x_{t+horizon}): `pairs` crops of side `crop` at
 * SplitMix64(seed) positions, from the sim's own trajectory (it advances). */
ah_status ah_surrogate_collect(ah_surrogate* s, ah_sim* sim, uint32_t pairs, uint32_t crop, uint32_t horizon,
                               uint32_t seed);
ah_status ah_surrogate_dataset(ah_surrogate* s, uint32_t first, uint32_t count, float* x, float* y);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on minibatches drawn from
 * pairs [0, n_train); losses[i] is step i's minibatch MSE (may be NULL). */
ah_status ah_surrogate_train(ah_surrogate* s, uint32_t steps, uint32_t batch, uint32_t n_train, float lr,
                             uint32_t seed, float* losses);
/* MSE of the network on pairs [first, first + count). */
ah_status ah_surrogate_eval(ah_surrogate* s, uint32_t first, uint32_t count, double* mse);
/*
 * Advance `sim` by `steps` steps with the network (horizon steps per call).
 * The first call of every window of judge_every calls (calls 1, 1 +
 * judge_every, ...) is judged before the window runs: `horizon` plain steps
 * from the same input, relative L2 of the network's output against them; the
 * sim keeps the plain result, and past `tol` the rest of the call falls back
 * to plain steps (a network that fails its first judge leaves the plain
 * trajectory unchanged).
 */
ah_status ah_surrogate_advance(ah_surrogate* s, ah_sim* sim, uint32_t steps, uint32_t judge_every, float tol,
                               ah_surrogate_report* report);

/*
 * Bandwidth probe: a plain vectorised device-to-device copy of `bytes`
 * (read `bytes`, write `bytes`), run `iters` times on a private stream after
 * one warm-up.  Returns the mean device milliseconds per copy.  It gives the
 * practical HBM ceiling beside the 8 TB/s spec figure and the known-size
 * kernel ("copy_probe") that byte counters are calibrated against.  No
 * counterpart in the reference.
 */
ah_status ah_copy_probe(size_t bytes, uint32_t iters, double* ms_per_copy);

/*
 * Name of the current HIP device ("AMD Instinct MI355X" and its gfx target),
 * the `adapter` a benchmark row records (the reference takes it from wgpu,
 * crates/auto-cli/src/bench.rs:163).  Truncated to capacity-1 bytes.
 */
ah_status ah_device_name(char* out, size_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* AUTO_HIP_H */
