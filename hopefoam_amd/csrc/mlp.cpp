// Host side of the learned troubled-cell indicator behind dgh_mlp_* (include/dgh.h): parameters,
// training loop (Adam over shuffled mini-batches, all arithmetic in the kernels of mlp.hip), BF16
// fragments for the MFMA inference of the limiter pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "dgh.h"
#include "dgh_internal.hpp"
#include "mlp.hpp"
#include "mlp_infer.hpp"

using namespace dg;

namespace dg {

struct MlpDev {
    int device = 0;
    int N = 0, Np = 0, nin = 0, np = 0;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *loss = nullptr;
    unsigned short *W1f = nullptr, *W2f = nullptr;
    float b3 = 0.0f;
    int adam_t = 0;
    std::vector<void*> allocs;
    ~MlpDev() {
        (void)hipSetDevice(device);
        for (void* a : allocs) (void)hipFree(a);
    }
};

}  // namespace dg

namespace {

template <typename T>
bool dalloc(dgh_ctx c, MlpDev& M, T** p, size_t n) {
    if (!c->hip(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc mlp")) return false;
    M.allocs.push_back(*p);
    return true;
}

// scratch buffers of one call (freed on return)
struct Scratch {
    std::vector<void*> v;
    template <typename T>
    bool get(dgh_ctx c, T** p, size_t n) {
        if (!c->hip(hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc")) return false;
        v.push_back(*p);
        return true;
    }
    ~Scratch() {
        for (void* p : v) (void)hipFree(p);
    }
};

MlpDev* get_mlp(dgh_ctx c, int h) {
    if (h < 0 || h >= int(c->mlps.size()) || !c->mlps[h]) {
        c->fail(DGH_ERR_ARG, "invalid mlp handle %d", h);
        return nullptr;
    }
    return c->mlps[h].get();
}

// BF16 A-operand fragments of W1 (64 x nin, K padded to mlp_pad(nin)) and W2 (64 x 64): lane l of tile (m, s)
// holds row 16 m + l % 16, k = 16 s + 4 (l / 16) + j
bool refresh_infer(dgh_ctx c, MlpDev& M) {
    std::vector<float> p(M.np);
    if (!c->hip(hipMemcpyAsync(p.data(), M.p, M.np * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") ||
        !c->hip(hipStreamSynchronize(c->stream), "sync"))
        return false;
    const int H = kMlpH, nin = M.nin, KS = mlp_pad(nin) / 16;
    std::vector<unsigned short> w1(4 * KS * 64 * 4, 0), w2(4 * 4 * 64 * 4, 0);
    for (int m = 0; m < 4; ++m)
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < 4; ++j) {
                const int row = 16 * m + (l & 15);
                for (int s = 0; s < KS; ++s) {
                    const int k = 16 * s + 4 * (l >> 4) + j;
                    w1[((m * KS + s) * 64 + l) * 4 + j] = k < nin ? f2bf(p[row * nin + k]) : 0;
                }
                for (int s = 0; s < 4; ++s) {
                    const int k = 16 * s + 4 * (l >> 4) + j;
                    w2[((m * 4 + s) * 64 + l) * 4 + j] = f2bf(p[H * nin + H + row * H + k]);
                }
            }
    M.b3 = p[M.np - 1];
    return c->hip(hipMemcpyAsync(M.W1f, w1.data(), w1.size() * 2, hipMemcpyHostToDevice, c->stream), "h2d") &&
           c->hip(hipMemcpyAsync(M.W2f, w2.data(), w2.size() * 2, hipMemcpyHostToDevice, c->stream), "h2d") &&
           c->hip(hipStreamSynchronize(c->stream), "sync");
}

}  // namespace

// inference view used by the limiter (limiter.cpp)
bool dg_mlp_infer(dgh_ctx c, int h, int N, double threshold, MlpInfer& w) {
    MlpDev* M = get_mlp(c, h);
    if (!M) return false;
    if (M->N != N) return c->fail(DGH_ERR_ARG, "mlp built for N=%d used on an N=%d mesh", M->N, N);
    if (!(threshold > 0.0 && threshold < 1.0)) return c->fail(DGH_ERR_ARG, "mlp threshold must be in (0, 1)");
    const int H = kMlpH;
    w.W1f = M->W1f;
    w.W2f = M->W2f;
    w.b1 = M->p + H * M->nin;
    w.b2 = M->p + H * M->nin + H + H * H;
    w.w3 = w.b2 + H;
    w.b3 = M->b3;
    w.zthr = float(std::log(threshold / (1.0 - threshold)));
    w.nin = M->nin;
    return true;
}

extern "C" {

int dgh_mlp_create(dgh_ctx c, int N, unsigned seed, int* out) {
    if (!c || !out) return DGH_ERR_ARG;
    const int Np = (N + 1) * (N + 2) / 2;
    if (N < 1 || mlp_inputs(Np) > kMlpInMax)
        return c->fail(DGH_ERR_UNSUPPORTED, "dgh_mlp_create: orders 1..6 (2 Np + 8 <= %d inputs)", kMlpInMax),
               DGH_ERR_UNSUPPORTED;
    (void)hipSetDevice(c->device);
    auto M = std::make_shared<MlpDev>();
    M->device = c->device;
    M->N = N;
    M->Np = Np;
    M->nin = mlp_inputs(Np);
    M->np = mlp_params(M->nin);
    // He-uniform weights, zero biases
    std::mt19937 rng(seed);
    std::vector<float> p(M->np, 0.0f);
    const int H = kMlpH;
    auto fill = [&](int off, int rows, int fan_in) {
        std::uniform_real_distribution<float> u(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
        for (int i = 0; i < rows * fan_in; ++i) p[off + i] = u(rng);
    };
    fill(0, H, M->nin);
    fill(H * M->nin + H, H, H);
    fill(H * M->nin + H + H * H + H, 1, H);
    std::vector<float> zero(M->np, 0.0f);
    bool ok = dalloc(c, *M, &M->p, M->np) && dalloc(c, *M, &M->m, M->np) && dalloc(c, *M, &M->v, M->np) &&
              dalloc(c, *M, &M->g, M->np) && dalloc(c, *M, &M->loss, 1) && dalloc(c, *M, &M->W1f, 4 * (mlp_pad(M->nin) / 16) * 64 * 4) &&
              dalloc(c, *M, &M->W2f, 4 * 4 * 64 * 4);
    ok = ok && c->hip(hipMemcpy(M->p, p.data(), M->np * sizeof(float), hipMemcpyHostToDevice), "h2d") &&
         c->hip(hipMemcpy(M->m, zero.data(), M->np * sizeof(float), hipMemcpyHostToDevice), "h2d") &&
         c->hip(hipMemcpy(M->v, zero.data(), M->np * sizeof(float), hipMemcpyHostToDevice), "h2d") &&
         refresh_infer(c, *M);
    if (!ok) return c->last;
    c->mlps.push_back(M);
    *out = int(c->mlps.size()) - 1;
    return DGH_OK;
}

int dgh_mlp_destroy(dgh_ctx c, int h) {
    if (!c) return DGH_ERR_ARG;
    if (!get_mlp(c, h)) return c->last;
    (void)hipStreamSynchronize(c->stream);
    c->mlps[h].reset();
    return DGH_OK;
}

int dgh_mlp_info(dgh_ctx c, int h, int* info) {
    if (!c || !info) return DGH_ERR_ARG;
    MlpDev* M = get_mlp(c, h);
    if (!M) return c->last;
    info[0] = M->N;
    info[1] = M->nin;
    info[2] = M->np;
    return DGH_OK;
}

int dgh_mlp_get_params(dgh_ctx c, int h, float* out) {
    if (!c || !out) return DGH_ERR_ARG;
    MlpDev* M = get_mlp(c, h);
    if (!M) return c->last;
    return c->hip(hipMemcpyAsync(out, M->p, M->np * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") &&
                   c->hip(hipStreamSynchronize(c->stream), "sync")
               ? DGH_OK
               : c->last;
}

int dgh_mlp_set_params(dgh_ctx c, int h, const float* in) {
    if (!c || !in) return DGH_ERR_ARG;
    MlpDev* M = get_mlp(c, h);
    if (!M) return c->last;
    (void)hipSetDevice(c->device);
    M->adam_t = 0;
    std::vector<float> zero(M->np, 0.0f);
    const bool ok = c->hip(hipMemcpyAsync(M->p, in, M->np * sizeof(float), hipMemcpyHostToDevice, c->stream), "h2d") &&
                    c->hip(hipMemcpyAsync(M->m, zero.data(), M->np * sizeof(float), hipMemcpyHostToDevice, c->stream),
                           "h2d") &&
                    c->hip(hipMemcpyAsync(M->v, zero.data(), M->np * sizeof(float), hipMemcpyHostToDevice, c->stream),
                           "h2d") &&
                    c->hip(hipStreamSynchronize(c->stream), "sync") && refresh_infer(c, *M);
    return ok ? DGH_OK : c->last;
}

int dgh_mlp_forward(dgh_ctx c, int h, const float* X, int n, int prec, float* z) {
    if (!c || !X || !z || n < 1) return DGH_ERR_ARG;
    MlpDev* M = get_mlp(c, h);
    if (!M) return c->last;
    if (prec != DGH_MLP_FP32 && prec != DGH_MLP_BF16)
        return c->fail(DGH_ERR_ARG, "dgh_mlp_forward: precision %d", prec), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    Scratch S;
    float *dX, *dz;
    if (!S.get(c, &dX, size_t(n) * M->nin) || !S.get(c, &dz, n)) return c->last;
    bool ok = c->hip(hipMemcpyAsync(dX, X, size_t(n) * M->nin * sizeof(float), hipMemcpyHostToDevice, c->stream), "h2d");
    if (ok && prec == DGH_MLP_FP32) {
        ok = c->hip(mlp_forward32(M->nin, n, dX, M->p, dz, c->stream), "mlp forward");
    } else if (ok) {
        MlpInfer w;
        ok = dg_mlp_infer(c, h, M->N, 0.5, w) && c->hip(mlp_forward_bf16(w, n, dX, dz, c->stream), "mlp forward bf16");
    }
    ok = ok && c->hip(hipMemcpyAsync(z, dz, n * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") &&
         c->hip(hipStreamSynchronize(c->stream), "sync");
    return ok ? DGH_OK : c->last;
}

int dgh_mlp_grad(dgh_ctx c, int h, const float* X, const float* y, int n, float pos_weight, float* grad, float* loss) {
    if (!c || !X || !y || n < 1) return DGH_ERR_ARG;
    MlpDev* M = get_mlp(c, h);
    if (!M) return c->last;
    (void)hipSetDevice(c->device);
    const int nb = (n + 63) / 64;
    Scratch S;
    float *dX, *dy, *part, *dl;
    if (!S.get(c, &dX, size_t(n) * M->nin) || !S.get(c, &dy, n) || !S.get(c, &part, size_t(nb) * (M->np + 1)) ||
        !S.get(c, &dl, 1))
        return c->last;
    MlpTrainArgs a;
    a.nin = M->nin;
    a.n = n;
    a.X = dX;
    a.y = dy;
    a.p = M->p;
    a.pos_weight = pos_weight;
    a.part = part;
    float l = 0.0f;
    const bool ok =
        c->hip(hipMemcpyAsync(dX, X, size_t(n) * M->nin * sizeof(float), hipMemcpyHostToDevice, c->stream), "h2d") &&
        c->hip(hipMemcpyAsync(dy, y, n * sizeof(float), hipMemcpyHostToDevice, c->stream), "h2d") &&
        c->hip(mlp_grad_partials(a, c->stream), "mlp grad") &&
        c->hip(mlp_reduce_adam(M->np, nb, part, M->g, M->p, M->m, M->v, 1, 0.0f, 0, dl, c->stream), "mlp reduce") &&
        (!grad || c->hip(hipMemcpyAsync(grad, M->g, M->np * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h")) &&
        c->hip(hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") &&
        c->hip(hipStreamSynchronize(c->stream), "sync");
    if (!ok) return c->last;
    if (loss) *loss = l;
    return DGH_OK;
}

int dgh_mlp_train(dgh_ctx c, int h, const float* X, const float* y, int n, const dgh_mlp_train_desc* d,
                  float* epoch_loss) {
    if (!c || !X || !y || !d || n < 1 || d->epochs < 0 || d->batch < 1 || !(d->lr > 0.0f)) return DGH_ERR_ARG;
    MlpDev* M = get_mlp(c, h);
    if (!M) return c->last;
    (void)hipSetDevice(c->device);
    const int B = std::min(d->batch, n), nbmax = (B + 63) / 64, steps = (n + B - 1) / B;
    Scratch S;
    float *dX, *dy, *Xb, *yb, *part, *losses;
    int* idx;
    if (!S.get(c, &dX, size_t(n) * M->nin) || !S.get(c, &dy, n) || !S.get(c, &Xb, size_t(B) * M->nin) ||
        !S.get(c, &yb, B) || !S.get(c, &part, size_t(nbmax) * (M->np + 1)) || !S.get(c, &idx, n) ||
        !S.get(c, &losses, std::max(1, steps)))
        return c->last;
    if (!c->hip(hipMemcpyAsync(dX, X, size_t(n) * M->nin * sizeof(float), hipMemcpyHostToDevice, c->stream), "h2d") ||
        !c->hip(hipMemcpyAsync(dy, y, n * sizeof(float), hipMemcpyHostToDevice, c->stream), "h2d"))
        return c->last;
    std::vector<int> perm(n);
    std::vector<float> hl(steps);
    std::mt19937 rng(d->seed);
    for (int ep = 0; ep < d->epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), rng);
        if (!c->hip(hipMemcpyAsync(idx, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, c->stream), "h2d"))
            return c->last;
        for (int st = 0; st < steps; ++st) {
            const int nb_s = std::min(B, n - st * B);
            MlpTrainArgs a;
            a.nin = M->nin;
            a.n = nb_s;
            a.X = Xb;
            a.y = yb;
            a.p = M->p;
            a.pos_weight = d->pos_weight;
            a.part = part;
            ++M->adam_t;
            if (!c->hip(mlp_gather(M->nin, nb_s, idx + size_t(st) * B, dX, dy, Xb, yb, c->stream), "gather") ||
                !c->hip(mlp_grad_partials(a, c->stream), "mlp grad") ||
                !c->hip(mlp_reduce_adam(M->np, (nb_s + 63) / 64, part, M->g, M->p, M->m, M->v, M->adam_t, d->lr, 1,
                                        losses + st, c->stream),
                        "adam"))
                return c->last;
        }
        // the host copy of the permutation must not change while the upload may still be in flight
        if (!c->hip(hipMemcpyAsync(hl.data(), losses, steps * sizeof(float), hipMemcpyDeviceToHost, c->stream), "d2h") ||
            !c->hip(hipStreamSynchronize(c->stream), "sync"))
            return c->last;
        if (epoch_loss) {
            double s = 0.0;
            for (int st = 0; st < steps; ++st) s += hl[st];
            epoch_loss[ep] = float(s / steps);
        }
    }
    return refresh_infer(c, *M) ? DGH_OK : c->last;
}

}  // extern "C"
