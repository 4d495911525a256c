// Host side of the learned shortlist network: parameters and Adam state, the training
// pool, the training loop (every step on the GPU: gather, forward, loss, backward,
// weight gradients, Adam), gradient and embedding entry points used by the tests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "v3dm_context.h"
#include "v3dm_learned.h"

namespace v3dm {

struct LearnedState {
    MlpWeights w;
    float* M = nullptr;
    float* V = nullptr;
    float* grad = nullptr;
    int step = 0;
    bool haveParams = false;
    float codeScale = 0.0f;  // quantisation scale of the codes (0: not calibrated)

    // training pool and triplets
    float* pool = nullptr;
    int nPool = 0;
    int* tri = nullptr;
    int nTri = 0;
    int* order = nullptr;

    // minibatch workspace for Rp rows
    int capRp = 0;
    __bf16 *X = nullptr, *XT = nullptr, *HT = nullptr, *dZ = nullptr, *dZT = nullptr, *dHT = nullptr;
    float *Z = nullptr, *dC = nullptr, *dZf = nullptr, *part = nullptr, *lossPart = nullptr;
    int* hard = nullptr;
};

namespace {

#define LCHECK(h, call)                                                       \
    do {                                                                      \
        hipError_t const e_ = (call);                                         \
        if (e_ != hipSuccess) return ctx_fail((h), V3DM_ERR_HIP, #call);      \
    } while (0)

int roundUp(int n, int m) { return (n + m - 1) / m * m; }

template <class T>
void freeDev(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int ensure_state(v3dm_context* h, LearnedState*& L)
{
    if (!L) L = new LearnedState();
    if (L->w.P) return V3DM_OK;
    LCHECK(h, hipMalloc(&L->w.P, kParams * sizeof(float)));
    LCHECK(h, hipMalloc(&L->w.W1b, kHid * 128 * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->w.W2b, kCode * kHid * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->w.W2Tb, kHid * kCode * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->w.W2p, kHid * kCode * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->M, kParams * sizeof(float)));
    LCHECK(h, hipMalloc(&L->V, kParams * sizeof(float)));
    LCHECK(h, hipMalloc(&L->grad, kParams * sizeof(float)));
    return V3DM_OK;
}

int ensure_workspace(v3dm_context* h, LearnedState* L, int Rp)
{
    if (L->capRp >= Rp) return V3DM_OK;
    for (__bf16** p : {&L->X, &L->XT, &L->HT, &L->dZ, &L->dZT, &L->dHT}) freeDev(*p);
    for (float** p : {&L->Z, &L->dC, &L->dZf, &L->part, &L->lossPart}) freeDev(*p);
    freeDev(L->hard);
    size_t const r = (size_t)Rp;
    LCHECK(h, hipMalloc(&L->X, r * 128 * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->XT, r * 128 * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->HT, r * kHid * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->dZ, r * kCode * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->dZT, r * kCode * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->dHT, r * kHid * sizeof(__bf16)));
    LCHECK(h, hipMalloc(&L->Z, r * kCode * sizeof(float)));
    LCHECK(h, hipMalloc(&L->dC, r * kCode * sizeof(float)));
    LCHECK(h, hipMalloc(&L->dZf, r * kCode * sizeof(float)));
    LCHECK(h, hipMalloc(&L->part, (size_t)mlp_wgrad_chunks(Rp) * kParams * sizeof(float)));
    LCHECK(h, hipMalloc(&L->lossPart, (r / 64 + 1) * sizeof(float)));
    LCHECK(h, hipMalloc(&L->hard, r * sizeof(int)));
    L->capRp = Rp;
    return V3DM_OK;
}

// One minibatch: B triplets given by order[orderOff .. orderOff+B) -> gradients (in
// L->grad) and, with apply, one Adam step.  Leaves the per-work-group loss sums in
// L->lossPart.
int enqueue_step(v3dm_context* h, LearnedState* L, int orderOff, int B, float lr, int apply, float tau)
{
    int const R = 3 * B;
    int const Rp = roundUp(R, kRowPadMlp);
    int rc = ensure_workspace(h, L, Rp);
    if (rc) return rc;
    hipStream_t const st = h->stream;
    launch_mlp_gather(L->pool, L->tri, L->order, orderOff, B, Rp, L->X, L->XT, st);
    launch_mlp_forward(L->X, Rp, L->w, L->HT, L->Z, st);
    launch_fill(L->dC, (size_t)Rp * kCode, 0.0f, st);
    launch_mlp_loss(L->Z, B, tau, L->dC, L->hard, L->lossPart, st);
    launch_mlp_norm_backward(L->Z, L->dC, R, Rp, L->dZf, L->dZ, L->dZT, st);
    launch_mlp_backward_h(L->dZ, Rp, L->w, L->HT, L->dHT, st);
    launch_mlp_wgrad(L->XT, L->HT, L->dZT, L->dHT, L->dZf, Rp, L->part, st);
    if (apply) L->step += 1;
    launch_mlp_adam(L->part, mlp_wgrad_chunks(Rp), L->grad, L->w, L->M, L->V, lr, std::max(L->step, 1), apply, st);
    LCHECK(h, hipGetLastError());
    return V3DM_OK;
}

int check_learned(v3dm_context* h, LearnedState*& L, bool needParams)
{
    if (!h) return V3DM_ERR_INVALID;
    L = ctx_learned(h);
    if (needParams && (!L || !L->haveParams)) return ctx_fail(h, V3DM_ERR_STATE, "learned parameters not set");
    return V3DM_OK;
}

}  // namespace

void learned_destroy(LearnedState*& L)
{
    if (!L) return;
    freeDev(L->w.P);
    freeDev(L->w.W1b);
    freeDev(L->w.W2b);
    freeDev(L->w.W2Tb);
    freeDev(L->w.W2p);
    for (float** p : {&L->M, &L->V, &L->grad, &L->pool, &L->Z, &L->dC, &L->dZf, &L->part, &L->lossPart}) freeDev(*p);
    for (__bf16** p : {&L->X, &L->XT, &L->HT, &L->dZ, &L->dZT, &L->dHT}) freeDev(*p);
    freeDev(L->tri);
    freeDev(L->order);
    freeDev(L->hard);
    delete L;
    L = nullptr;
}

MlpWeights const* learned_weights(LearnedState* L) { return (L && L->haveParams) ? &L->w : nullptr; }
float learned_code_scale(LearnedState* L) { return L ? L->codeScale : 0.0f; }

}  // namespace v3dm

using namespace v3dm;

extern "C" {

int v3dm_learned_set_state(v3dm_handle h, const float* params, const float* m, const float* v, int step)
{
    if (!h) return V3DM_ERR_INVALID;
    if (!params || step < 0) return ctx_fail(h, V3DM_ERR_INVALID, "invalid learned state");
    LearnedState*& L = ctx_learned(h);
    int rc = ensure_state(h, L);
    if (rc) return rc;
    hipStream_t const st = h->stream;
    LCHECK(h, hipMemcpyAsync(L->w.P, params, kParams * sizeof(float), hipMemcpyHostToDevice, st));
    if (m) LCHECK(h, hipMemcpyAsync(L->M, m, kParams * sizeof(float), hipMemcpyHostToDevice, st));
    else launch_fill(L->M, kParams, 0.0f, st);
    if (v) LCHECK(h, hipMemcpyAsync(L->V, v, kParams * sizeof(float), hipMemcpyHostToDevice, st));
    else launch_fill(L->V, kParams, 0.0f, st);
    L->step = step;
    // refresh the bf16 copies (no update: apply = 0 with a zero partial set)
    launch_fill(L->grad, kParams, 0.0f, st);
    launch_mlp_adam(L->grad, 1, L->grad, L->w, L->M, L->V, 0.0f, 1, 0, st);
    LCHECK(h, hipGetLastError());
    LCHECK(h, hipStreamSynchronize(st));
    L->haveParams = true;
    return V3DM_OK;
}

int v3dm_learned_get_state(v3dm_handle h, float* params, float* m, float* v, int* step)
{
    LearnedState* L = nullptr;
    int rc = check_learned(h, L, true);
    if (rc) return rc;
    hipStream_t const st = h->stream;
    LCHECK(h, hipStreamSynchronize(st));
    if (params) LCHECK(h, hipMemcpy(params, L->w.P, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (m) LCHECK(h, hipMemcpy(m, L->M, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (v) LCHECK(h, hipMemcpy(v, L->V, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (step) *step = L->step;
    return V3DM_OK;
}

int v3dm_learned_set_code_scale(v3dm_handle h, float scale)
{
    LearnedState* L = nullptr;
    int rc = check_learned(h, L, true);
    if (rc) return rc;
    if (!(scale > 0.0f) || !(scale < 1e6f)) return ctx_fail(h, V3DM_ERR_INVALID, "invalid code scale");
    L->codeScale = scale;
    return V3DM_OK;
}

int v3dm_learned_set_training(v3dm_handle h, int nPool, const float* pool, int nTriplets, const int32_t* triplets)
{
    LearnedState* L = nullptr;
    int rc = check_learned(h, L, true);
    if (rc) return rc;
    if (nPool <= 0 || !pool || nTriplets <= 0 || !triplets) return ctx_fail(h, V3DM_ERR_INVALID, "invalid training set");
    for (int i = 0; i < 3 * nTriplets; ++i)
        if (triplets[i] < 0 || triplets[i] >= nPool) return ctx_fail(h, V3DM_ERR_INVALID, "triplet row out of range");
    freeDev(L->pool);
    freeDev(L->tri);
    freeDev(L->order);
    LCHECK(h, hipMalloc(&L->pool, (size_t)nPool * 128 * sizeof(float)));
    LCHECK(h, hipMalloc(&L->tri, (size_t)nTriplets * 3 * sizeof(int)));
    LCHECK(h, hipMalloc(&L->order, (size_t)nTriplets * sizeof(int)));
    LCHECK(h, hipMemcpy(L->pool, pool, (size_t)nPool * 128 * sizeof(float), hipMemcpyHostToDevice));
    LCHECK(h, hipMemcpy(L->tri, triplets, (size_t)nTriplets * 3 * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> ident(nTriplets);
    for (int i = 0; i < nTriplets; ++i) ident[i] = i;
    LCHECK(h, hipMemcpy(L->order, ident.data(), (size_t)nTriplets * sizeof(int), hipMemcpyHostToDevice));
    L->nPool = nPool;
    L->nTri = nTriplets;
    return V3DM_OK;
}

// steps minibatches of `batch` triplets; every epoch the triplet order is reshuffled
// (seeded).  losses[s] = mean loss of step s (may be NULL).
int v3dm_learned_train(v3dm_handle h, int steps, int batch, float lr, float tau, uint64_t seed, float* losses)
{
    LearnedState* L = nullptr;
    int rc = check_learned(h, L, true);
    if (rc) return rc;
    if (!L->pool) return ctx_fail(h, V3DM_ERR_STATE, "no training set");
    if (steps < 0 || batch < 2 || batch > L->nTri || !(lr >= 0.0f) || !(tau > 0.0f))
        return ctx_fail(h, V3DM_ERR_INVALID, "invalid training parameters");
    hipStream_t const st = h->stream;
    int const perEpoch = L->nTri / batch;
    int const nLoss = (batch + 63) / 64;
    std::vector<float> lossDev;
    float* lossAll = nullptr;
    if (losses && steps > 0) LCHECK(h, hipMalloc(&lossAll, (size_t)steps * nLoss * sizeof(float)));
    std::mt19937_64 rng(seed);
    std::vector<int> order(L->nTri);
    for (int s = 0; s < steps; ++s) {
        int const inEpoch = s % perEpoch;
        if (inEpoch == 0) {
            for (int i = 0; i < L->nTri; ++i) order[i] = i;
            std::shuffle(order.begin(), order.end(), rng);
            LCHECK(h, hipStreamSynchronize(st));  // the previous epoch's order is consumed
            LCHECK(h, hipMemcpy(L->order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        if ((rc = enqueue_step(h, L, inEpoch * batch, batch, lr, 1, tau))) {
            if (lossAll) (void)hipFree(lossAll);
            return rc;
        }
        if (lossAll)
            LCHECK(h, hipMemcpyAsync(lossAll + (size_t)s * nLoss, L->lossPart, nLoss * sizeof(float),
                                     hipMemcpyDeviceToDevice, st));
    }
    LCHECK(h, hipStreamSynchronize(st));
    if (lossAll) {
        std::vector<float> host((size_t)steps * nLoss);
        hipError_t const e = hipMemcpy(host.data(), lossAll, host.size() * sizeof(float), hipMemcpyDeviceToHost);
        (void)hipFree(lossAll);
        if (e != hipSuccess) return ctx_fail(h, V3DM_ERR_HIP, "loss download failed");
        for (int s = 0; s < steps; ++s) {
            double acc = 0.0;
            for (int k = 0; k < nLoss; ++k) acc += host[(size_t)s * nLoss + k];
            losses[s] = (float)(acc / batch);
        }
    }
    return V3DM_OK;
}

// Loss and gradient (FP32, packed parameter layout) of one minibatch of the training set:
// triplets given as indices into the training triplet list.  No update.  hard (may be
// NULL) receives the in-batch hardest positive chosen for each anchor.
int v3dm_learned_gradients(v3dm_handle h, int nBatch, const int32_t* tripletIdx, float tau, float* loss,
                           float* grads, int32_t* hard)
{
    LearnedState* L = nullptr;
    int rc = check_learned(h, L, true);
    if (rc) return rc;
    if (!L->pool) return ctx_fail(h, V3DM_ERR_STATE, "no training set");
    if (nBatch < 2 || nBatch > L->nTri || !tripletIdx || !(tau > 0.0f))
        return ctx_fail(h, V3DM_ERR_INVALID, "invalid batch");
    for (int i = 0; i < nBatch; ++i)
        if (tripletIdx[i] < 0 || tripletIdx[i] >= L->nTri) return ctx_fail(h, V3DM_ERR_INVALID, "triplet index");
    hipStream_t const st = h->stream;
    LCHECK(h, hipStreamSynchronize(st));
    LCHECK(h, hipMemcpy(L->order, tripletIdx, nBatch * sizeof(int), hipMemcpyHostToDevice));
    if ((rc = enqueue_step(h, L, 0, nBatch, 0.0f, 0, tau))) return rc;
    LCHECK(h, hipStreamSynchronize(st));
    if (grads) LCHECK(h, hipMemcpy(grads, L->grad, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (hard) LCHECK(h, hipMemcpy(hard, L->hard, nBatch * sizeof(int), hipMemcpyDeviceToHost));
    if (loss) {
        int const nLoss = (nBatch + 63) / 64;
        std::vector<float> lp(nLoss);
        LCHECK(h, hipMemcpy(lp.data(), L->lossPart, nLoss * sizeof(float), hipMemcpyDeviceToHost));
        double acc = 0.0;
        for (float x : lp) acc += x;
        *loss = (float)(acc / nBatch);
    }
    // the training order is reset by the next v3dm_learned_train epoch
    return V3DM_OK;
}

// Embedding z (before normalisation) of n host descriptors -> codes (n x 32, FP32).
int v3dm_learned_embed(v3dm_handle h, int n, const float* desc, float* codes)
{
    LearnedState* L = nullptr;
    int rc = check_learned(h, L, true);
    if (rc) return rc;
    if (n <= 0 || !desc || !codes) return ctx_fail(h, V3DM_ERR_INVALID, "invalid embedding input");
    int const Rp = roundUp(n, kRowPadMlp);
    if ((rc = ensure_workspace(h, L, Rp))) return rc;
    hipStream_t const st = h->stream;
    float* src = nullptr;
    LCHECK(h, hipMalloc(&src, (size_t)n * 128 * sizeof(float)));
    hipError_t e = hipMemcpy(src, desc, (size_t)n * 128 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_mlp_rows_bf16(src, n, Rp, L->X, st);
        launch_mlp_forward(L->X, Rp, L->w, nullptr, L->Z, st);
        e = hipStreamSynchronize(st);
    }
    if (e == hipSuccess) e = hipMemcpy(codes, L->Z, (size_t)n * kCode * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(src);
    if (e != hipSuccess) return ctx_fail(h, V3DM_ERR_HIP, "embedding failed");
    return V3DM_OK;
}

}  // extern "C"
