// Precision modes other than the exact FP32 path: the BF16, FP8 and learned-code shortlist
// forms.
#include "v3dm_ext.h"

#include "v3dm_context.h"
#include "v3dm_learned.h"
#include "v3dm_short.h"

namespace v3dm {

struct ExtState {
    int capL = 0, capR = 0;  // rows of the bf16 buffers (multiple of kShortRowPad)
    __bf16* bL = nullptr;
    __bf16* bR = nullptr;
    bool bLValid = false, bRValid = false;
    int* secL = nullptr;  // exact second best candidate per query (shortlist forms)
    int* secR = nullptr;
    bool haveSecond = false;
    unsigned char* f8L = nullptr;  // e4m3 descriptors, capL x 128 / capR x 128
    unsigned char* f8R = nullptr;
    unsigned char* eL = nullptr;  // E8M0 scale byte per descriptor
    unsigned char* eR = nullptr;
    bool f8LValid = false, f8RValid = false;
    signed char* cL = nullptr;  // learned codes (int8, 32 per feature)
    signed char* cR = nullptr;
    bool cLValid = false, cRValid = false;
    float cScale = 0.0f;  // scale the codes were quantised with

    // sampled exact check of a learned result (kSampleRows left features)
    float* sL = nullptr;       // sampled rows
    float* sNorm = nullptr;
    int* sIdx = nullptr;
    Top2* sColPart = nullptr;  // kMaxSplits x kSampleRows
    int* sDec = nullptr;       // exact decisions of the samples (+1 unused right slot; mapped host memory)
    int* sDecHost = nullptr;   // the same memory as the host addresses it
    int* sFlags = nullptr;     // open-query list of the samples
    Top2* sRpart = nullptr;    // re-check partials
    int* sCounters = nullptr;  // kNumCounters
    int* sHost = nullptr;      // learned decisions of the samples (pinned, mapped, coherent)
    int* sHostDev = nullptr;   // sHost as the kernels address it
    hipStream_t side = nullptr;  // the check runs beside the learned passes
    hipEvent_t evFork = nullptr;
    bool forkNeeded = true;    // the features were written since the side stream last waited
    bool sPending = false;
    int sSamples = 0;
    Cand2* partL = nullptr;  // kShortMaxSplits x capL x 2
    Cand2* partR = nullptr;
};

namespace {

int roundUpTo(int n, int m) { return (n + m - 1) / m * m; }

// the learned pass's packed lists live in the first half of the partition buffers
CandP* packed(Cand2* p) { return reinterpret_cast<CandP*>(p); }

#define EXTCHECK(h, call)                                                    \
    do {                                                                     \
        hipError_t const e_ = (call);                                        \
        if (e_ != hipSuccess) return ctx_fail((h), V3DM_ERR_HIP, #call);     \
    } while (0)

int ensure(v3dm_context* h, ExtState*& e)
{
    if (!e) e = new ExtState();
    int const cl = roundUpTo(h->capL, kShortRowPad), cr = roundUpTo(h->capR, kShortRowPad);
    if (e->bL && e->capL == cl && e->capR == cr) return V3DM_OK;
    ext_release(e);
    e->capL = cl;
    e->capR = cr;
    EXTCHECK(h, hipMalloc(&e->bL, (size_t)cl * kDim * sizeof(__bf16)));
    EXTCHECK(h, hipMalloc(&e->bR, (size_t)cr * kDim * sizeof(__bf16)));
    EXTCHECK(h, hipMemsetAsync(e->bL, 0, (size_t)cl * kDim * sizeof(__bf16), h->stream));
    EXTCHECK(h, hipMemsetAsync(e->bR, 0, (size_t)cr * kDim * sizeof(__bf16), h->stream));
    EXTCHECK(h, hipMalloc(&e->partL, (size_t)kShortMaxSplits * cl * 2 * sizeof(Cand2)));
    EXTCHECK(h, hipMalloc(&e->partR, (size_t)kShortMaxSplits * cr * 2 * sizeof(Cand2)));
    EXTCHECK(h, hipMalloc(&e->secL, (size_t)cl * sizeof(int)));
    EXTCHECK(h, hipMalloc(&e->secR, (size_t)cr * sizeof(int)));
    e->bLValid = e->bRValid = false;
    return V3DM_OK;
}

int ensure_fp8(v3dm_context* h, ExtState*& e)
{
    int rc = ensure(h, e);
    if (rc) return rc;
    if (e->f8L) return V3DM_OK;
    size_t const bl = (size_t)e->capL * kDim, br = (size_t)e->capR * kDim;
    EXTCHECK(h, hipMalloc(&e->f8L, bl));
    EXTCHECK(h, hipMalloc(&e->f8R, br));
    EXTCHECK(h, hipMalloc(&e->eL, e->capL));
    EXTCHECK(h, hipMalloc(&e->eR, e->capR));
    // padding rows stay zero with scale 2^0 (they are masked out of every decision)
    EXTCHECK(h, hipMemsetAsync(e->f8L, 0, bl, h->stream));
    EXTCHECK(h, hipMemsetAsync(e->f8R, 0, br, h->stream));
    EXTCHECK(h, hipMemsetAsync(e->eL, 127, e->capL, h->stream));
    EXTCHECK(h, hipMemsetAsync(e->eR, 127, e->capR, h->stream));
    e->f8LValid = e->f8RValid = false;
    return V3DM_OK;
}

int enqueue_bf16(v3dm_context* h, float minScore, float minRatio)
{
    ExtState*& e = ctx_ext(h);
    int rc = ensure(h, e);
    if (rc) return rc;
    hipStream_t const st = h->stream;
    int const nL = h->nL, nR = h->nR;
    bool const work = nL > 0 && nR > 0;
    if (work) {
        if (!e->bLValid) launch_to_bf16(h->dL, nL, e->bL, st);
        if (!e->bRValid) launch_to_bf16(h->dR, nR, e->bR, st);
        e->bLValid = e->bRValid = true;
        ShortPlan const pl = plan_short(nL, nR, 512);
        ShortPlan const pr = plan_short(nR, nL, 512);
        // both directions in one launch (the profiled "dominant kernel" covers both)
        if ((rc = prof_begin(h))) return rc;
        launch_pass_bf16_2(e->bL, nL, e->bR, nR, pl, e->partL, e->capL, pr, e->partR, e->capR, st);
        if ((rc = prof_end(h))) return rc;
        launch_rescore(h->dL, h->dR, nL, nR, e->partL, pl.splits, e->capL, e->partR, pr.splits, e->capR, h->normL,
                       h->normR, h->maxNormBits, kWinRelBf16, 0, INFINITY, minScore, minRatio,
                       h->colMax, h->rowMax, e->secL, e->secR, h->stripes, st);
    }
    e->haveSecond = work;
    launch_mutual_compact(h->colMax, h->rowMax, work ? nL : 0, nR, h->mstatus, next_mutual_epoch(h), h->stripes,
                          reinterpret_cast<int2*>(h->results + kNumCounters), h->results, st, {}, h->mutualWait);
    EXTCHECK(h, hipGetLastError());
    return V3DM_OK;
}

int ensure_codes(v3dm_context* h, ExtState*& e)
{
    int rc = ensure(h, e);
    if (rc) return rc;
    if (e->cL) return V3DM_OK;
    EXTCHECK(h, hipMalloc(&e->cL, (size_t)e->capL * kCode));
    EXTCHECK(h, hipMalloc(&e->cR, (size_t)e->capR * kCode));
    e->cLValid = e->cRValid = false;
    return V3DM_OK;
}

int ensure_sample(v3dm_context* h, ExtState* e)
{
    if (e->sL) return V3DM_OK;
    EXTCHECK(h, hipMalloc(&e->sL, (size_t)kSampleRows * kDim * sizeof(float)));
    EXTCHECK(h, hipMalloc(&e->sNorm, kSampleRows * sizeof(float)));
    EXTCHECK(h, hipMalloc(&e->sIdx, kSampleRows * sizeof(int)));
    EXTCHECK(h, hipMalloc(&e->sColPart, (size_t)kSampleMaxSplits * kSampleRows * sizeof(Top2)));
    EXTCHECK(h, hipHostMalloc(&e->sDecHost, (kSampleRows + 1) * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    EXTCHECK(h, hipHostGetDevicePointer((void**)&e->sDec, e->sDecHost, 0));
    EXTCHECK(h, hipMalloc(&e->sFlags, (kSampleRows + 1) * sizeof(int)));
    EXTCHECK(h, hipMalloc(&e->sRpart, (size_t)(kSampleRows + 1) * kRecheckMaxChunks * sizeof(Top2)));
    EXTCHECK(h, hipMalloc(&e->sCounters, kNumCounters * sizeof(int)));
    EXTCHECK(h, hipHostMalloc(&e->sHost, kSampleRows * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    EXTCHECK(h, hipHostGetDevicePointer((void**)&e->sHostDev, e->sHost, 0));
    EXTCHECK(h, hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    EXTCHECK(h, hipEventCreateWithFlags(&e->evFork, hipEventDisableTiming));
    e->forkNeeded = true;
    return V3DM_OK;
}

// The plain exact path on S sampled left features against all right features, on a side
// stream (it reads only the features and its own buffers).  The side stream waits for the
// handle's stream (evFork, recorded at the start of the call) only when the features were
// written since it last did: every writer of the features first waits for the side stream
// (ext_quiesce), so a side stream that has waited once stays ordered after the features.
// Nothing on the handle's stream waits for the side stream: the exact decisions of the
// samples go to mapped host memory, the mutual check writes the learned decisions of the
// same samples there too, and the comparison runs on the host when the result is read
// (ext_finish, after both streams are synchronised).
int enqueue_sample_check(v3dm_context* h, ExtState* e, float minScore, float minRatio)
{
    hipStream_t const st = e->side;
    if (e->forkNeeded) {
        EXTCHECK(h, hipStreamWaitEvent(st, e->evFork, 0));
        e->forkNeeded = false;
    }
    int const nL = h->nL, nR = h->nR;
    int const S = std::min(std::min(h->fallbackSamples, kSampleRows), nL);
    launch_sample_gather(h->dL, h->normL, nL, S, e->sL, e->sNorm, e->sIdx, e->sCounters, st);
    FusedPlan const p = plan_fused_fp32(S, nR, kSampleMaxSplits, kSampleMaxSplits);
    launch_fused_fp32(e->sL, h->dR, S, nR, p, e->sColPart, kSampleRows, nullptr, 0, st);  // left decisions only
    launch_sample_decide(e->sColPart, p.splits, kSampleRows, S, e->sNorm, h->maxNormBits, h->epsRel, minScore,
                         minRatio, e->sDec, e->sFlags, e->sCounters, st);
    launch_recheck_exact(e->sL, h->dR, S, nR, e->sFlags, kSampleRows, e->sCounters, minScore, minRatio, e->sDec,
                         e->sDec + kSampleRows, e->sRpart, kRecheckGrid, st);
    e->sSamples = S;
    return V3DM_OK;
}

int enqueue_learned(v3dm_context* h, float minScore, float minRatio)
{
    MlpWeights const* w = learned_weights(ctx_learned(h));
    float const scale = learned_code_scale(ctx_learned(h));
    if (!w || !(scale > 0.0f)) return ctx_fail(h, V3DM_ERR_STATE, "learned network or code scale not set");
    ExtState*& e = ctx_ext(h);
    int rc = ensure_codes(h, e);
    if (rc) return rc;
    hipStream_t const st = h->stream;
    int const nL = h->nL, nR = h->nR;
    bool const work = nL > 0 && nR > 0;
    if (work) {
        if (e->cScale != scale) e->cLValid = e->cRValid = false;
        e->cScale = scale;
        if (h->fallbackSamples > 0) {
            if ((rc = ensure_sample(h, e))) return rc;
            if (e->forkNeeded) EXTCHECK(h, hipEventRecord(e->evFork, st));  // the check needs only the features
        }
        // embedding: network forward on the FP32 rows (read up to the next multiple of 128,
        // inside the allocation) with the quantisation fused; code rows past n are never
        // used (candidates past n are masked, queries past n not written)
        launch_mlp_codes2_f32(h->dL, e->cLValid ? 0 : (nL + 127) / 128 * 128, e->cL, h->dR,
                              e->cRValid ? 0 : (nR + 127) / 128 * 128, e->cR, *w, scale, st);
        e->cLValid = e->cRValid = true;
        bool const check = h->fallbackSamples > 0;
        // the main stream's kernels are enqueued first: the check's stream operations cost
        // more host time than its kernels take, and the main stream should not wait for them
        // 1024 work-groups: the code pass runs four per CU
        ShortPlan const pl = plan_short_packed(nL, nR, 1024);
        ShortPlan const pr = plan_short_packed(nR, nL, 1024);
        // both directions in one launch (the profiled "dominant kernel" covers both)
        if ((rc = prof_begin(h))) return rc;
        launch_pass_code2(e->cL, nL, e->cR, nR, pl, packed(e->partL), e->capL, pr, packed(e->partR), e->capR, st);
        if ((rc = prof_end(h))) return rc;
        launch_rescore_packed(h->dL, h->dR, nL, nR, packed(e->partL), pl.splits, e->capL, pl.tilesPerSplit * 64,
                              packed(e->partR), pr.splits, e->capR, pr.tilesPerSplit * 64, h->normL, h->normR,
                              h->maxNormBits, h->learnedTopK, learned_margin(h), minScore, minRatio, h->colMax,
                              h->rowMax, e->secL, e->secR, h->stripes, st);
        e->haveSecond = true;
        SampleOut so;
        if (check) {
            if ((rc = enqueue_sample_check(h, e, minScore, minRatio))) return rc;
            so = SampleOut{e->sSamples, e->sHostDev};
            e->sPending = true;
        }
        launch_mutual_compact(h->colMax, h->rowMax, nL, nR, h->mstatus, next_mutual_epoch(h), h->stripes,
                              reinterpret_cast<int2*>(h->results + kNumCounters), h->results, st, so, h->mutualWait);
    } else {
        e->haveSecond = false;
        launch_mutual_compact(h->colMax, h->rowMax, 0, nR, h->mstatus, next_mutual_epoch(h), h->stripes,
                              reinterpret_cast<int2*>(h->results + kNumCounters), h->results, st, {}, h->mutualWait);
    }
    EXTCHECK(h, hipGetLastError());
    return V3DM_OK;
}

int enqueue_fp8(v3dm_context* h, float minScore, float minRatio)
{
    ExtState*& e = ctx_ext(h);
    int rc = ensure_fp8(h, e);
    if (rc) return rc;
    hipStream_t const st = h->stream;
    int const nL = h->nL, nR = h->nR;
    bool const work = nL > 0 && nR > 0;
    if (work) {
        if (!e->f8LValid) launch_to_fp8(h->dL, nL, e->f8L, e->eL, st);
        if (!e->f8RValid) launch_to_fp8(h->dR, nR, e->f8R, e->eR, st);
        e->f8LValid = e->f8RValid = true;
        ShortPlan const pl = plan_short(nL, nR, 512);
        ShortPlan const pr = plan_short(nR, nL, 512);
        if ((rc = prof_begin(h))) return rc;
        launch_pass_fp8_2(e->f8L, e->eL, nL, e->f8R, e->eR, nR, pl, e->partL, e->capL, pr, e->partR, e->capR, st);
        if ((rc = prof_end(h))) return rc;
        launch_rescore(h->dL, h->dR, nL, nR, e->partL, pl.splits, e->capL, e->partR, pr.splits, e->capR, h->normL,
                       h->normR, h->maxNormBits, kWinRelFp8, 0, INFINITY, minScore, minRatio,
                       h->colMax, h->rowMax, e->secL, e->secR, h->stripes, st);
    }
    e->haveSecond = work;
    launch_mutual_compact(h->colMax, h->rowMax, work ? nL : 0, nR, h->mstatus, next_mutual_epoch(h), h->stripes,
                          reinterpret_cast<int2*>(h->results + kNumCounters), h->results, st, {}, h->mutualWait);
    EXTCHECK(h, hipGetLastError());
    return V3DM_OK;
}

}  // namespace

bool ext_supports(v3dm_context*, int mode)
{
    return mode == V3DM_PREC_FP32 || mode == V3DM_PREC_BF16 || mode == V3DM_PREC_FP8 || mode == V3DM_PREC_LEARNED;
}

// The dominant pass of a shortlist form, launched `reps` times back to back between two
// events (the inputs of the last call: its derived copies are still in place).
int ext_time_pass(v3dm_context* h, int mode, int reps, hipEvent_t start, hipEvent_t stop, int* launches)
{
    ExtState* e = ctx_ext(h);
    int const nL = h->nL, nR = h->nR;
    if (!e || nL <= 0 || nR <= 0) return ctx_fail(h, V3DM_ERR_STATE, "no shortlist result to time");
    hipStream_t const st = h->stream;
    EXTCHECK(h, hipEventRecord(start, st));
    if (mode == V3DM_PREC_LEARNED) {
        if (!e->cL || !e->cR) return ctx_fail(h, V3DM_ERR_STATE, "no learned codes");
        ShortPlan const pl = plan_short_packed(nL, nR, 1024);
        ShortPlan const pr = plan_short_packed(nR, nL, 1024);
        for (int r = 0; r < reps; ++r)
            launch_pass_code2(e->cL, nL, e->cR, nR, pl, packed(e->partL), e->capL, pr, packed(e->partR), e->capR, st);
        *launches = reps;
    } else {
        ShortPlan const pl = plan_short(nL, nR, 512);
        ShortPlan const pr = plan_short(nR, nL, 512);
        if (mode == V3DM_PREC_BF16) {
            if (!e->bL || !e->bR) return ctx_fail(h, V3DM_ERR_STATE, "no bf16 copies");
            for (int r = 0; r < reps; ++r)
                launch_pass_bf16_2(e->bL, nL, e->bR, nR, pl, e->partL, e->capL, pr, e->partR, e->capR, st);
        } else {
            if (!e->f8L || !e->f8R) return ctx_fail(h, V3DM_ERR_STATE, "no fp8 copies");
            for (int r = 0; r < reps; ++r)
                launch_pass_fp8_2(e->f8L, e->eL, nL, e->f8R, e->eR, nR, pl, e->partL, e->capL, pr, e->partR,
                                  e->capR, st);
        }
        *launches = reps;
    }
    EXTCHECK(h, hipEventRecord(stop, st));
    EXTCHECK(h, hipGetLastError());
    return V3DM_OK;
}

int ext_enqueue(v3dm_context* h, int mode, float minScore, float minRatio)
{
    if (mode == V3DM_PREC_BF16) return enqueue_bf16(h, minScore, minRatio);
    if (mode == V3DM_PREC_FP8) return enqueue_fp8(h, minScore, minRatio);
    if (mode == V3DM_PREC_LEARNED) return enqueue_learned(h, minScore, minRatio);
    return ctx_fail(h, V3DM_ERR_UNSUPPORTED, "precision mode not available");
}

// Called before a result is read: a learned result whose sampled check shows a recall below
// the handle's threshold is replaced by the plain form's (stats.nFallback = 1).
int ext_finish(v3dm_context* h)
{
    ExtState* e = ctx_ext(h);
    if (!e || !e->sPending) return V3DM_OK;
    e->sPending = false;
    EXTCHECK(h, hipStreamSynchronize(h->stream));
    EXTCHECK(h, hipStreamSynchronize(e->side));
    // exactly accepted samples, and those whose learned decision differs
    int accepted = 0, missed = 0;
    for (int k = 0; k < e->sSamples; ++k) {
        int const x = e->sDecHost[k], d = e->sHost[k];
        accepted += x >= 0;
        missed += (x >= 0 && d != x);
    }
    h->stats.nSampled = e->sSamples;
    h->stats.nSampleMissed = missed;
    if (missed > fallback_allowed(h, accepted)) {
        int const rc = enqueue_plain(h, h->lastMinScore, h->lastMinRatio);
        if (rc) return rc;
        h->stats.nFallback = 1;
    }
    return V3DM_OK;
}

int ext_second_best(v3dm_context* h, int* colSecond, int* rowSecond)
{
    ExtState* e = ctx_ext(h);
    if (!e || !e->haveSecond) return ctx_fail(h, V3DM_ERR_STATE, "no shortlist result");
    EXTCHECK(h, hipStreamSynchronize(h->stream));
    if (colSecond && h->nL > 0)
        EXTCHECK(h, hipMemcpy(colSecond, e->secL, (size_t)h->nL * sizeof(int), hipMemcpyDeviceToHost));
    if (rowSecond && h->nR > 0)
        EXTCHECK(h, hipMemcpy(rowSecond, e->secR, (size_t)h->nR * sizeof(int), hipMemcpyDeviceToHost));
    return V3DM_OK;
}

void ext_quiesce(v3dm_context* h)
{
    ExtState* e = ctx_ext(h);
    if (!e || !e->side) return;
    (void)hipStreamSynchronize(e->side);
    e->forkNeeded = true;
}

void ext_synchronize(v3dm_context* h)
{
    ExtState* e = ctx_ext(h);
    if (e && e->side) (void)hipStreamSynchronize(e->side);
}

void ext_discard_result(v3dm_context* h)
{
    ExtState* e = ctx_ext(h);
    if (!e) return;
    e->sPending = false;
    e->haveSecond = false;
}

void ext_features_changed(v3dm_context* h, bool left)
{
    ExtState* e = ctx_ext(h);
    if (!e) return;
    e->sPending = false;  // the last result is of the old features: its check is moot
    if (left) e->bLValid = e->f8LValid = e->cLValid = false;
    else e->bRValid = e->f8RValid = e->cRValid = false;
}

void ext_release(ExtState* e)
{
    if (!e) return;
    if (e->bL) (void)hipFree(e->bL);
    if (e->bR) (void)hipFree(e->bR);
    if (e->partL) (void)hipFree(e->partL);
    if (e->partR) (void)hipFree(e->partR);
    if (e->secL) (void)hipFree(e->secL);
    if (e->secR) (void)hipFree(e->secR);
    e->secL = e->secR = nullptr;
    e->haveSecond = false;
    for (unsigned char* p : {e->f8L, e->f8R, e->eL, e->eR})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)e->cL, (void*)e->cR})
        if (p) (void)hipFree(p);
    e->cL = e->cR = nullptr;
    e->cLValid = e->cRValid = false;
    if (e->side) (void)hipStreamSynchronize(e->side);
    for (void* p : {(void*)e->sL, (void*)e->sNorm, (void*)e->sIdx, (void*)e->sColPart, (void*)e->sFlags,
                    (void*)e->sRpart, (void*)e->sCounters})
        if (p) (void)hipFree(p);
    if (e->sHost) (void)hipHostFree(e->sHost);
    if (e->sDecHost) (void)hipHostFree(e->sDecHost);
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->evFork) (void)hipEventDestroy(e->evFork);
    e->side = nullptr;
    e->evFork = nullptr;
    e->forkNeeded = true;
    e->sL = e->sNorm = nullptr;
    e->sIdx = e->sDec = e->sDecHost = e->sFlags = e->sCounters = e->sHost = e->sHostDev = nullptr;
    e->sColPart = e->sRpart = nullptr;
    e->sPending = false;
    e->f8L = e->f8R = e->eL = e->eR = nullptr;
    e->f8LValid = e->f8RValid = false;
    e->bL = e->bR = nullptr;
    e->partL = e->partR = nullptr;
    e->capL = e->capR = 0;
    e->bLValid = e->bRValid = false;
}

void ext_destroy(ExtState*& e)
{
    ext_release(e);
    delete e;
    e = nullptr;
}

}  // namespace v3dm
