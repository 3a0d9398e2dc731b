// Learned k-step surrogate: host side of ah_surrogate_* (see surrogate.hip
// for the network and its kernels).  No counterpart in the reference.
//
// Training data are (x_t, x_{t+k}) crops taken from a sim's own trajectory
// (the plain step generates them); training is Adam on the MSE loss with the
// project's forward and backward kernels; inference advances a sim k steps
// per network call.  The first call of every window of `judge_every` calls is
// judged before the window is trusted: k plain steps from the same input,
// compared by relative L2.  Past the tolerance the run falls back to plain
// stepping for the rest of the call.  At a judged call the sim continues from
// the plain result, so a network that fails its first judge never touches the
// state and the run is the plain trajectory bit for bit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "ah_internal.h"
#include "surrogate.h"

using namespace ah;
using namespace ah::surr;

namespace {

// Activation and gradient buffers for one (N, H, W).
struct Work {
    int N = 0, H = 0, W = 0;
    bool train = false;
    std::vector<void*> bufs;
    uint16_t *xb = nullptr, *a1 = nullptr, *p = nullptr, *a2 = nullptr, *s = nullptr;
    uint16_t *g1 = nullptr, *g2 = nullptr, *g3 = nullptr;  // bf16 gelu' of each layer's pre-activation (training)
    float *r = nullptr, *y = nullptr;
    float *dr = nullptr, *dp = nullptr;
    uint16_t *dz3 = nullptr, *dz2 = nullptr;
    float* part = nullptr;
    float* bpart = nullptr;  // bias-gradient partials, 2048 workgroups x 64 channels
    double* lpart = nullptr;
    void release() {
        for (void* b : bufs) (void)hipFree(b);
        bufs.clear();
        N = H = W = 0;
    }
    template <typename T>
    T* get(size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, n * sizeof(T));
        if (e != hipSuccess) throw Error{AH_ERR_ALLOC, std::string("surrogate workspace: ") + hipGetErrorString(e)};
        bufs.push_back(q);
        return (T*)q;
    }
};

}  // namespace

struct ah_surrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    std::vector<void*> bufs;
    float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    PackedWeights pw{};
    int adam_t = 0;
    Work work;
    // dataset
    float *data_x = nullptr, *data_y = nullptr;
    uint32_t n_pairs = 0, crop = 0, horizon = 4;
    float *bx = nullptr, *by = nullptr;
    int* bidx = nullptr;
    uint32_t bcap = 0;

    void* alloc(size_t bytes) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) throw Error{AH_ERR_ALLOC, std::string("surrogate: ") + hipGetErrorString(e)};
        bufs.push_back(q);
        return q;
    }
    void free_one(void*& q) {
        if (!q) return;
        auto it = std::find(bufs.begin(), bufs.end(), q);
        if (it != bufs.end()) bufs.erase(it);
        (void)hipFree(q);
        q = nullptr;
    }
    ~ah_surrogate() {
        if (stream) (void)hipStreamSynchronize(stream);
        work.release();
        for (void* b : bufs) (void)hipFree(b);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void ensure_work(ah_surrogate* S, int N, int H, int W, bool train) {
    Work& w = S->work;
    if (w.N == N && w.H == H && w.W == W && (w.train || !train)) return;
    if (H % 8 || W % 128 || H < 8) throw Error{AH_ERR_INVALID, "surrogate: height % 8 == 0 and width % 128 == 0"};
    AH_HIP(hipStreamSynchronize(S->stream));
    w.release();
    const size_t P = (size_t)N * H * W, Ph = P / 4;
    w.N = N;
    w.H = H;
    w.W = W;
    w.train = train;
    w.xb = w.get<uint16_t>(P);
    w.a1 = w.get<uint16_t>(P * C1);
    w.p = w.get<uint16_t>(Ph * C1);
    w.a2 = w.get<uint16_t>(Ph * C2);
    w.s = w.get<uint16_t>(P * C1);
    w.r = w.get<float>(P);
    w.y = w.get<float>(P);
    w.lpart = w.get<double>(1024 * 2);
    if (train) {
        w.g1 = w.get<uint16_t>(P * C1);
        w.g2 = w.get<uint16_t>(Ph * C2);
        w.g3 = w.get<uint16_t>(P * C1);
        w.dr = w.get<float>(P);
        w.dz3 = w.get<uint16_t>(P * C1);
        w.dz2 = w.get<uint16_t>(Ph * C2);
        w.dp = w.get<float>(Ph * C1);
        w.part = w.get<float>(WGRAD_PART_FLOATS);
        w.bpart = w.get<float>(WGRAD_BPART_FLOATS);
    }
}

// x (device, [N][H][W]) -> work.y; gelu' of every pre-activation kept when training.
void forward(ah_surrogate* S, const float* x, bool train) {
    Work& w = S->work;
    hipStream_t st = S->stream;
    Profiler* pf = &S->prof;
    const int N = w.N, H = w.H, W = w.W;
    ConvArgs c{};
    c.N = N;
    c.H = H;
    c.W = W;
    c.Cin = 32;
    c.ks = 5;
    c.Wt = S->pw.W1x;
    c.Cout = C1;
    c.bias = S->theta + OFF_B1;
    c.mode = CONV_GELU;
    c.dgelu = train ? w.g1 : nullptr;
    c.a = w.a1;
    conv1(st, pf, c, x, w.xb, w.p);
    c = ConvArgs{};
    c.X = w.p;
    c.N = N;
    c.H = H / 2;
    c.W = W / 2;
    c.Cin = C1;
    c.ks = 5;
    c.Wt = S->pw.W2;
    c.Cout = C2;
    c.bias = S->theta + OFF_B2;
    c.mode = CONV_GELU;
    c.dgelu = train ? w.g2 : nullptr;
    c.a = w.a2;
    conv(st, pf, c, "surr_conv2");
    c = ConvArgs{};
    c.X = w.a2;
    c.N = N;
    c.H = H;
    c.W = W;
    c.Cin = C2;
    c.up = 1;
    c.ks = 5;
    c.Wt = S->pw.W3;
    c.Cout = C1;
    c.bias = S->theta + OFF_B3;
    c.mode = CONV_GELU_ADD;
    c.dgelu = train ? w.g3 : nullptr;
    c.a = w.s;
    c.add = w.a1;
    conv(st, pf, c, "surr_conv3");
    out_layer(st, pf, w.s, S->pw.W4b, S->theta + OFF_B4, x, w.r, w.y, N, H, W);
}

// Loss of work.y against `target` and, when `grad` is set, the full
// gradient into S->grad.  Returns the MSE (synchronises).
double backward(ah_surrogate* S, const float* x, const float* target, bool grad) {
    Work& w = S->work;
    hipStream_t st = S->stream;
    Profiler* pf = &S->prof;
    const int N = w.N, H = w.H, W = w.W;
    const size_t P = (size_t)N * H * W;
    const int lb = loss_blocks(P);
    surr::loss(st, pf, w.y, target, x, w.r, grad ? w.dr : nullptr, w.lpart, P);
    if (grad) {
        float* g = S->grad;
        WGradArgs a{};
        // layer 4
        a.dz_f = w.dr;
        a.X = w.s;
        a.xstride = C1;
        a.N = N;
        a.H = H;
        a.W = W;
        a.Cin = C1;
        a.Cout = 1;
        wgrad(st, pf, a, w.part, g + OFF_W4, w.bpart, g + OFF_B4);
        // layer 3 (ds, the layer-4 input gradient, is formed from dr where it is used)
        dz3_from_dr(st, pf, w.dr, S->pw.W4, w.g3, w.dz3, N, H, W);
        a = WGradArgs{};
        a.dz_h = w.dz3;
        a.X = w.a2;
        a.xstride = C2;
        a.N = N;
        a.H = H;
        a.W = W;
        a.Cin = C2;
        a.Cout = C1;
        a.up = 1;
        wgrad(st, pf, a, w.part, g + OFF_W3, w.bpart, g + OFF_B3);
        ConvArgs c{};
        c.X = w.dz3;
        c.N = N;
        c.H = H;
        c.W = W;
        c.Cin = C1;
        c.ks = 5;
        c.Wt = S->pw.W3T;
        c.Cout = C2;
        // dZ2 = (the input gradient summed over 2 x 2 blocks) * gelu'(z2),
        // formed in the convolution's epilogue.
        c.mode = CONV_POOL_DGELU;
        c.dgelu_in = w.g2;
        c.dpool = w.dz2;
        conv(st, pf, c, "surr_conv3_dgrad");
        // layer 2
        a = WGradArgs{};
        a.dz_h = w.dz2;
        a.X = w.p;
        a.xstride = C1;
        a.N = N;
        a.H = H / 2;
        a.W = W / 2;
        a.Cin = C1;
        a.Cout = C2;
        wgrad(st, pf, a, w.part, g + OFF_W2, w.bpart, g + OFF_B2);
        c = ConvArgs{};
        c.X = w.dz2;
        c.N = N;
        c.H = H / 2;
        c.W = W / 2;
        c.Cin = C2;
        c.ks = 5;
        c.Wt = S->pw.W2T;
        c.Cout = C1;
        c.mode = CONV_LINEAR;
        c.z = w.dp;
        conv(st, pf, c, "surr_conv2_dgrad");
        // layer 1 (dz1 is formed inside the weight-gradient kernel)
        dz1_wgrad(st, pf, w.dr, S->pw.W4, w.dp, w.g1, w.xb, N, H, W, w.part, g + OFF_W1, w.bpart, g + OFF_B1);
    }
    std::vector<double> part(lb);
    AH_HIP(hipMemcpyAsync(part.data(), w.lpart, sizeof(double) * lb, hipMemcpyDeviceToHost, st));
    AH_HIP(hipStreamSynchronize(st));
    double sum = 0.0;
    for (double v : part) sum += v;
    return sum / (double)P;
}

void repack(ah_surrogate* S) { pack_weights(S->stream, &S->prof, S->theta, S->pw); }

void init_params(uint32_t seed, std::vector<float>& th) {
    th.assign(N_PARAMS, 0.0f);
    uint64_t st = seed;
    auto u01 = [&]() { return (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0); };
    struct Blk {
        int off, n, fan_in;
        double scale;
    } blks[] = {{OFF_W1, C1 * 25, 25, 1.0},
                {OFF_W2, C2 * 25 * C1, 25 * C1, 1.0},
                {OFF_W3, C1 * 25 * C2, 25 * C2, 1.0},
                {OFF_W4, 25 * C1, 25 * C1, 0.1}};
    // Draw order follows the parameter order; biases draw nothing.
    for (const Blk& b : blks) {
        const double lim = std::sqrt(3.0 / b.fan_in) * b.scale;
        for (int i = 0; i < b.n; ++i) th[b.off + i] = (float)((2.0 * u01() - 1.0) * lim);
    }
}

void check_dims(uint32_t n, uint32_t h, uint32_t w) {
    if (n == 0 || h == 0 || w == 0) throw Error{AH_ERR_INVALID, "empty batch"};
    if (h % 8 || w % 128) throw Error{AH_ERR_INVALID, "surrogate: height % 8 == 0 and width % 128 == 0"};
}

double rel_l2_dev(ah_surrogate* S, const float* a, const float* b, size_t n) {
    const int nb = rel_l2_blocks(n);
    rel_l2(S->stream, a, b, n, S->work.lpart);
    std::vector<double> part(2 * nb);
    AH_HIP(hipMemcpyAsync(part.data(), S->work.lpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, S->stream));
    AH_HIP(hipStreamSynchronize(S->stream));
    double d2 = 0.0, b2 = 0.0;
    for (int i = 0; i < nb; ++i) {
        d2 += part[2 * i];
        b2 += part[2 * i + 1];
    }
    return b2 > 0.0 ? std::sqrt(d2 / b2) : std::sqrt(d2);
}

}  // namespace

extern "C" {

ah_status ah_surrogate_new(uint32_t seed, ah_surrogate** out) {
    return guard([&] {
        if (!out) throw Error{AH_ERR_INVALID, "null argument"};
        *out = nullptr;
        std::unique_ptr<ah_surrogate> S(new ah_surrogate());
        AH_HIP(hipGetDevice(&S->device));
        AH_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        S->theta = (float*)S->alloc(sizeof(float) * N_PARAMS);
        S->m = (float*)S->alloc(sizeof(float) * N_PARAMS);
        S->v = (float*)S->alloc(sizeof(float) * N_PARAMS);
        S->grad = (float*)S->alloc(sizeof(float) * N_PARAMS);
        S->pw.W1x = (uint16_t*)S->alloc(2 * 32 * 32);
        S->pw.W2 = (uint16_t*)S->alloc(2 * 64 * 25 * 32);
        S->pw.W2T = (uint16_t*)S->alloc(2 * 64 * 25 * 32);
        S->pw.W3 = (uint16_t*)S->alloc(2 * 64 * 25 * 32);
        S->pw.W3T = (uint16_t*)S->alloc(2 * 64 * 25 * 32);
        S->pw.W4b = (uint16_t*)S->alloc(2 * 25 * 32);
        S->pw.W4 = (float*)S->alloc(4 * 25 * 32);
        std::vector<float> th;
        init_params(seed, th);
        AH_HIP(hipMemcpy(S->theta, th.data(), sizeof(float) * N_PARAMS, hipMemcpyHostToDevice));
        AH_HIP(hipMemset(S->m, 0, sizeof(float) * N_PARAMS));
        AH_HIP(hipMemset(S->v, 0, sizeof(float) * N_PARAMS));
        repack(S.get());
        AH_HIP(hipStreamSynchronize(S->stream));
        *out = S.release();
    });
}

void ah_surrogate_destroy(ah_surrogate* S) { delete S; }

ah_status ah_surrogate_get_params(const ah_surrogate* S, float* dst, size_t n) {
    return guard([&] {
        if (!S || !dst) throw Error{AH_ERR_INVALID, "null argument"};
        if (n != (size_t)N_PARAMS) throw Error{AH_ERR_SIZE, "the surrogate has 104129 parameters"};
        AH_HIP(hipMemcpyAsync(dst, S->theta, sizeof(float) * n, hipMemcpyDeviceToHost, S->stream));
        AH_HIP(hipStreamSynchronize(S->stream));
    });
}

ah_status ah_surrogate_set_params(ah_surrogate* S, const float* src, size_t n) {
    return guard([&] {
        if (!S || !src) throw Error{AH_ERR_INVALID, "null argument"};
        if (n != (size_t)N_PARAMS) throw Error{AH_ERR_SIZE, "the surrogate has 104129 parameters"};
        AH_HIP(hipMemcpyAsync(S->theta, src, sizeof(float) * n, hipMemcpyHostToDevice, S->stream));
        repack(S);
        AH_HIP(hipStreamSynchronize(S->stream));
    });
}

ah_status ah_surrogate_forward(ah_surrogate* S, const float* x, uint32_t n, uint32_t h, uint32_t w, float* y) {
    return guard([&] {
        if (!S || !x || !y) throw Error{AH_ERR_INVALID, "null argument"};
        check_dims(n, h, w);
        ensure_work(S, (int)n, (int)h, (int)w, false);
        const size_t P = (size_t)n * h * w;
        float* dx = (float*)S->alloc(sizeof(float) * P);
        try {
            AH_HIP(hipMemcpyAsync(dx, x, sizeof(float) * P, hipMemcpyHostToDevice, S->stream));
            forward(S, dx, false);
            AH_HIP(hipMemcpyAsync(y, S->work.y, sizeof(float) * P, hipMemcpyDeviceToHost, S->stream));
            AH_HIP(hipStreamSynchronize(S->stream));
        } catch (...) {
            S->free_one((void*&)dx);
            throw;
        }
        S->free_one((void*&)dx);
    });
}

ah_status ah_surrogate_loss_grad(ah_surrogate* S, const float* x, const float* target, uint32_t n, uint32_t h,
                                 uint32_t w, double* loss_out, float* grad_out) {
    return guard([&] {
        if (!S || !x || !target || !loss_out) throw Error{AH_ERR_INVALID, "null argument"};
        check_dims(n, h, w);
        ensure_work(S, (int)n, (int)h, (int)w, true);
        const size_t P = (size_t)n * h * w;
        float* dx = (float*)S->alloc(sizeof(float) * P);
        float* dt = (float*)S->alloc(sizeof(float) * P);
        try {
            AH_HIP(hipMemcpyAsync(dx, x, sizeof(float) * P, hipMemcpyHostToDevice, S->stream));
            AH_HIP(hipMemcpyAsync(dt, target, sizeof(float) * P, hipMemcpyHostToDevice, S->stream));
            forward(S, dx, true);
            *loss_out = backward(S, dx, dt, true);
            if (grad_out) {
                AH_HIP(hipMemcpyAsync(grad_out, S->grad, sizeof(float) * N_PARAMS, hipMemcpyDeviceToHost, S->stream));
                AH_HIP(hipStreamSynchronize(S->stream));
            }
        } catch (...) {
            S->free_one((void*&)dx);
            S->free_one((void*&)dt);
            throw;
        }
        S->free_one((void*&)dx);
        S->free_one((void*&)dt);
    });
}

ah_status ah_surrogate_collect(ah_surrogate* S, ah_sim* sim, uint32_t pairs, uint32_t crop_side, uint32_t horizon,
                               uint32_t seed) {
    return guard([&] {
        if (!S || !sim) throw Error{AH_ERR_INVALID, "null argument"};
        const ah_lenia_desc& d = sim_desc(sim);
        if (d.channels != 1) throw Error{AH_ERR_INVALID, "the surrogate learns single-channel worlds"};
        if (pairs == 0 || horizon == 0 || crop_side == 0 || crop_side > d.width || crop_side > d.height)
            throw Error{AH_ERR_INVALID, "pairs, horizon > 0 and crop <= world side"};
        check_dims(1, crop_side, crop_side);
        AH_HIP(hipStreamSynchronize(S->stream));
        S->free_one((void*&)S->data_x);
        S->free_one((void*&)S->data_y);
        const size_t per = (size_t)crop_side * crop_side;
        S->data_x = (float*)S->alloc(sizeof(float) * per * pairs);
        S->data_y = (float*)S->alloc(sizeof(float) * per * pairs);
        S->n_pairs = pairs;
        S->crop = crop_side;
        S->horizon = horizon;
        // Crops are taken in groups of up to 128 per start time t, with the
        // targets k steps later; consecutive groups start k+1 steps apart.
        const uint32_t group = std::min<uint32_t>(pairs, 128);
        int2* corners = (int2*)S->alloc(sizeof(int2) * group);
        uint64_t rs = seed;
        hipStream_t st = sim_stream(sim);
        try {
            std::vector<int2> host(group);
            for (uint32_t done = 0; done < pairs;) {
                const uint32_t cnt = std::min(group, pairs - done);
                for (uint32_t i = 0; i < cnt; ++i)
                    host[i] = make_int2((int)(splitmix64(rs) % d.width), (int)(splitmix64(rs) % d.height));
                AH_HIP(hipMemcpyAsync(corners, host.data(), sizeof(int2) * cnt, hipMemcpyHostToDevice, st));
                crop(st, sim_state(sim), (int)d.width, (int)d.height, corners, (int)cnt, (int)crop_side,
                     S->data_x + per * done);
                sim_advance(sim, horizon);
                crop(st, sim_state(sim), (int)d.width, (int)d.height, corners, (int)cnt, (int)crop_side,
                     S->data_y + per * done);
                sim_advance(sim, 1);
                AH_HIP(hipStreamSynchronize(st));  // `host` is reused for the next group
                done += cnt;
            }
        } catch (...) {
            S->free_one((void*&)corners);
            throw;
        }
        S->free_one((void*&)corners);
    });
}

ah_status ah_surrogate_dataset(ah_surrogate* S, uint32_t first, uint32_t count, float* x, float* y) {
    return guard([&] {
        if (!S || !x || !y) throw Error{AH_ERR_INVALID, "null argument"};
        if ((uint64_t)first + count > S->n_pairs) throw Error{AH_ERR_INVALID, "pair range outside the dataset"};
        const size_t per = (size_t)S->crop * S->crop;
        AH_HIP(hipMemcpyAsync(x, S->data_x + per * first, sizeof(float) * per * count, hipMemcpyDeviceToHost, S->stream));
        AH_HIP(hipMemcpyAsync(y, S->data_y + per * first, sizeof(float) * per * count, hipMemcpyDeviceToHost, S->stream));
        AH_HIP(hipStreamSynchronize(S->stream));
    });
}

ah_status ah_surrogate_train(ah_surrogate* S, uint32_t steps, uint32_t batch, uint32_t n_train, float lr,
                             uint32_t seed, float* losses) {
    return guard([&] {
        if (!S) throw Error{AH_ERR_INVALID, "null surrogate"};
        if (!S->n_pairs) throw Error{AH_ERR_INVALID, "no dataset: call ah_surrogate_collect first"};
        if (batch == 0 || n_train == 0 || n_train > S->n_pairs) throw Error{AH_ERR_INVALID, "bad batch or n_train"};
        const int side = (int)S->crop;
        ensure_work(S, (int)batch, side, side, true);
        const size_t per = (size_t)side * side;
        if (S->bcap < batch) {
            S->free_one((void*&)S->bx);
            S->free_one((void*&)S->by);
            S->free_one((void*&)S->bidx);
            S->bx = (float*)S->alloc(sizeof(float) * per * batch);
            S->by = (float*)S->alloc(sizeof(float) * per * batch);
            S->bidx = (int*)S->alloc(sizeof(int) * batch);
            S->bcap = batch;
        }
        uint64_t rs = seed;
        std::vector<int> idx(batch);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (uint32_t it = 0; it < steps; ++it) {
            for (uint32_t b = 0; b < batch; ++b) idx[b] = (int)(splitmix64(rs) % n_train);
            AH_HIP(hipMemcpyAsync(S->bidx, idx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, S->stream));
            gather(S->stream, S->data_x, S->bidx, (int)batch, per, S->bx);
            gather(S->stream, S->data_y, S->bidx, (int)batch, per, S->by);
            forward(S, S->bx, true);
            const double l = backward(S, S->bx, S->by, true);  // synchronises: idx may be rewritten
            ++S->adam_t;
            const float c1 = 1.0f - std::pow(b1, (float)S->adam_t), c2 = 1.0f - std::pow(b2, (float)S->adam_t);
            adam(S->stream, &S->prof, S->theta, S->m, S->v, S->grad, N_PARAMS, lr, b1, b2, eps, c1, c2);
            repack(S);
            if (losses) losses[it] = (float)l;
        }
        AH_HIP(hipStreamSynchronize(S->stream));
    });
}

ah_status ah_surrogate_eval(ah_surrogate* S, uint32_t first, uint32_t count, double* mse) {
    return guard([&] {
        if (!S || !mse) throw Error{AH_ERR_INVALID, "null argument"};
        if (count == 0 || (uint64_t)first + count > S->n_pairs) throw Error{AH_ERR_INVALID, "pair range outside the dataset"};
        const int side = (int)S->crop;
        const size_t per = (size_t)side * side;
        const uint32_t chunk = std::min<uint32_t>(count, 64);
        ensure_work(S, (int)chunk, side, side, S->work.train);
        double sum = 0.0;
        for (uint32_t done = 0; done < count;) {
            const uint32_t cnt = std::min(chunk, count - done);
            if ((uint32_t)S->work.N != cnt) ensure_work(S, (int)cnt, side, side, false);
            const float* x = S->data_x + per * (first + done);
            forward(S, x, false);
            sum += backward(S, x, S->data_y + per * (first + done), false) * (double)cnt;
            done += cnt;
        }
        *mse = sum / count;
    });
}

ah_status ah_surrogate_advance(ah_surrogate* S, ah_sim* sim, uint32_t steps, uint32_t judge_every, float tol,
                               ah_surrogate_report* rep) {
    return guard([&] {
        if (!S || !sim || !rep) throw Error{AH_ERR_INVALID, "null argument"};
        const ah_lenia_desc& d = sim_desc(sim);
        if (d.channels != 1) throw Error{AH_ERR_INVALID, "the surrogate advances single-channel worlds"};
        check_dims(1, d.height, d.width);
        if (judge_every == 0) throw Error{AH_ERR_INVALID, "judge_every must be > 0"};
        const uint32_t k = S->horizon;
        ensure_work(S, 1, (int)d.height, (int)d.width, false);
        const size_t P = (size_t)d.width * d.height;
        ah_surrogate_report r{};
        hipStream_t sst = sim_stream(sim);
        auto t0 = std::chrono::steady_clock::now();
        try {
            AH_HIP(hipStreamSynchronize(sst));
            bool fallen = false;
            uint32_t calls = 0;
            uint32_t left = steps;
            while (left >= k && !fallen) {
                float* state = sim_state(sim);
                forward(S, state, false);
                ++calls;
                if ((calls - 1) % judge_every == 0) {
                    // Judge (calls 1, 1 + judge_every, ...): k plain steps from
                    // the same input, on the sim's path.
                    AH_HIP(hipStreamSynchronize(S->stream));
                    sim_advance(sim, k);
                    AH_HIP(hipStreamSynchronize(sst));
                    const double e = rel_l2_dev(S, S->work.y, sim_state(sim), P);
                    r.judges++;
                    r.last_judge_rel_l2 = e;
                    r.max_judge_rel_l2 = std::max(r.max_judge_rel_l2, e);
                    r.plain_steps += k;
                    if (e > tol) {
                        r.fallbacks++;
                        fallen = true;
                    }
                } else {
                    AH_HIP(hipMemcpyAsync(state, S->work.y, sizeof(float) * P, hipMemcpyDeviceToDevice, S->stream));
                    sim_add_generations(sim, k);
                    r.surrogate_calls++;
                }
                left -= k;
            }
            AH_HIP(hipStreamSynchronize(S->stream));
            if (left) {
                sim_advance(sim, left);
                r.plain_steps += left;
            }
            AH_HIP(hipStreamSynchronize(sst));
        } catch (...) {
            (void)hipStreamSynchronize(S->stream);
            throw;
        }
        r.lenia_steps = steps;
        r.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *rep = r;
    });
}

}  // extern "C"
