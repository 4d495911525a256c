// C ABI of the gfx950 putative matcher: handle state, buffer management, orchestration.
//
// Mirrors V3D_CUDA::SIFT_Matching (V3D/CUDA/v3d_cudamatching.h:15-67,
// v3d_cudamatching.cu:439-694).  Device buffers are owned by the handle; descriptor
// buffers hold a multiple of 128 rows so tile loads never leave the allocation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/v3dm.h"
#include "v3dm_device.h"
#include "v3dm_kernels.h"
#include "v3dm_ext.h"
#include "v3dm_context.h"

using namespace v3dm;

namespace {

thread_local std::string g_createError;

int roundUp(int n, int m) { return (n + m - 1) / m * m; }

}  // namespace


namespace {

int fail(v3dm_context* h, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    else g_createError = buf;
    return code;
}

#define HIPCHECK(h, call)                                                                           \
    do {                                                                                            \
        hipError_t const e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                       \
            return fail((h), (e_ == hipErrorOutOfMemory) ? V3DM_ERR_NOMEM : V3DM_ERR_HIP,          \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

template <typename T>
void freeDev(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void releaseBuffers(v3dm_context* h)
{
    freeDev(h->dL);
    freeDev(h->dR);
    freeDev(h->normL);
    freeDev(h->normR);
    freeDev(h->maxNormBits);
    freeDev(h->colPart);
    freeDev(h->rowPart);
    freeDev(h->colMax);
    freeDev(h->rowMax);
    freeDev(h->flagList);
    freeDev(h->rpart);
    freeDev(h->mstatus);
    freeDev(h->stripes);
    freeDev(h->results);
    freeDev(h->dQ);
    freeDev(h->dLeftXY);
    freeDev(h->dRightXY);
    freeDev(h->guidedBits);
    if (h->hostStage) (void)hipHostFree(h->hostStage);
    h->hostStage = nullptr;
    h->dQBytes = 0;
    h->allocated = false;
    h->nL = h->nR = 0;
    h->haveResult = false;
    h->masks.clear();
    ext_release(h->ext);
}

// The ring of event pairs is harvested when it is full and in v3dm_get_stats.
int harvestEvents(v3dm_context* h)
{
    for (int e = 0; e < h->evPending; ++e) {
        HIPCHECK(h, hipEventSynchronize(h->evStop[e]));
        float ms = 0.0f;
        HIPCHECK(h, hipEventElapsedTime(&ms, h->evStart[e], h->evStop[e]));
        h->kernelMs += ms;
        h->kernelLaunches += 1;
    }
    h->evPending = 0;
    return V3DM_OK;
}

int roundWarn(v3dm_context* h, int n, const char* who, const char* what)
{
    if ((n % 32) != 0) {
        if (h->warnOnRounding) {
            std::cerr << "SIFT_Matching::" << who << " warning: " << what << " should be a multiple of 32"
                      << std::endl;
            std::cerr << "Rounding " << what << " down." << std::endl;
        }
        n = n - (n & 31);
    }
    return n;
}

}  // namespace

// ---------------------------------------------------------------------------------------- internal API
namespace v3dm {

hipStream_t ctx_stream(v3dm_context* h) { return h->stream; }
ExtState*& ctx_ext(v3dm_context* h) { return h->ext; }
BatchState*& ctx_batch(v3dm_context* h) { return h->batch; }
LearnedState*& ctx_learned(v3dm_context* h) { return h->learned; }
int ctx_fail(v3dm_context* h, int code, const char* msg) { return fail(h, code, "%s", msg); }

// Enqueue the plain FP32 exact path for the current features; leaves colMax/rowMax
// decided and the compacted pairs + counters in h->results.
int prof_begin(v3dm_context* h)
{
    if (!h->profiling) return V3DM_OK;
    if (h->evPending == (int)h->evStart.size()) {
        int const rc = harvestEvents(h);
        if (rc) return rc;
    }
    HIPCHECK(h, hipEventRecord(h->evStart[h->evPending], h->stream));
    return V3DM_OK;
}

int prof_end(v3dm_context* h)
{
    if (!h->profiling) return V3DM_OK;
    HIPCHECK(h, hipEventRecord(h->evStop[h->evPending], h->stream));
    h->evPending += 1;
    return V3DM_OK;
}

int enqueue_plain(v3dm_context* h, float minScore, float minRatio)
{
    hipStream_t const st = h->stream;
    int const nL = h->nL, nR = h->nR;
    bool const work = nL > 0 && nR > 0;
    int rc;
    if (work) {
        FusedPlan const p = plan_fused_fp32(nL, nR, 512);
        if ((rc = prof_begin(h))) return rc;
        launch_fused_fp32(h->dL, h->dR, nL, nR, p, h->colPart, h->capL, h->rowPart, h->capR, st);
        if ((rc = prof_end(h))) return rc;
        launch_merge_decide(h->colPart, p.splits, h->capL, h->rowPart, p.leftBlocks, h->capR, nL, nR, h->normL,
                            h->normR, h->maxNormBits, h->epsRel, minScore, minRatio, h->colMax, h->rowMax,
                            h->flagList, h->capL, h->results, st);
        launch_recheck_exact(h->dL, h->dR, nL, nR, h->flagList, h->capL, h->results, minScore, minRatio,
                             h->colMax, h->rowMax, h->rpart, kRecheckGrid, st);
    }
    launch_mutual_compact(h->colMax, h->rowMax, work ? nL : 0, nR, h->mstatus, next_mutual_epoch(h), h->stripes,
                          reinterpret_cast<int2*>(h->results + kNumCounters), h->results, st, {}, h->mutualWait);
    HIPCHECK(h, hipGetLastError());
    return V3DM_OK;
}

}  // namespace v3dm

// ---------------------------------------------------------------------------------------- C ABI
extern "C" {

int v3dm_create(v3dm_handle* out, int device)
{
    if (!out) return fail(nullptr, V3DM_ERR_INVALID, "v3dm_create: null handle pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(nullptr, V3DM_ERR_HIP, "v3dm_create: no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(nullptr, V3DM_ERR_INVALID, "v3dm_create: device %d out of range [0, %d)", device, count);
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, V3DM_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    v3dm_context* h = new v3dm_context();
    h->device = device;
    e = hipStreamCreateWithFlags(&h->ownStream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return fail(nullptr, V3DM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    h->stream = h->ownStream;
    *out = h;
    return V3DM_OK;
}

void v3dm_destroy(v3dm_handle h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    ext_quiesce(h);
    releaseBuffers(h);
    ext_destroy(h->ext);
    batch_destroy(h->batch);
    learned_destroy(h->learned);
    for (auto e : h->evStart) (void)hipEventDestroy(e);
    for (auto e : h->evStop) (void)hipEventDestroy(e);
    if (h->ownStream) (void)hipStreamDestroy(h->ownStream);
    delete h;
}

const char* v3dm_last_error(v3dm_handle h) { return h ? h->err.c_str() : g_createError.c_str(); }

int v3dm_allocate(v3dm_handle h, int nMaxLeft, int nMaxRight, int useGuidedMatching)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (nMaxLeft < 0 || nMaxRight < 0) return fail(h, V3DM_ERR_INVALID, "allocate: negative maximum");
    if (nMaxLeft > (1 << 20) || nMaxRight > (1 << 20))
        return fail(h, V3DM_ERR_INVALID, "allocate: maximum above 2^20 features");
    HIPCHECK(h, hipSetDevice(h->device));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    ext_quiesce(h);
    releaseBuffers(h);
    h->useGuided = useGuidedMatching != 0;
    nMaxLeft = roundWarn(h, nMaxLeft, "allocate()", "nMaxLeftFeatures");
    nMaxRight = roundWarn(h, nMaxRight, "allocate()", "nMaxRightFeatures");
    h->nMaxL = nMaxLeft;
    h->nMaxR = nMaxRight;
    h->capL = std::max(kRowPad, roundUp(nMaxLeft, kRowPad));
    h->capR = std::max(kRowPad, roundUp(nMaxRight, kRowPad));
    size_t const cl = (size_t)h->capL, cr = (size_t)h->capR;
    HIPCHECK(h, hipMalloc(&h->dL, cl * kDim * sizeof(float)));
    HIPCHECK(h, hipMalloc(&h->dR, cr * kDim * sizeof(float)));
    HIPCHECK(h, hipMemsetAsync(h->dL, 0, cl * kDim * sizeof(float), h->stream));
    HIPCHECK(h, hipMemsetAsync(h->dR, 0, cr * kDim * sizeof(float), h->stream));
    HIPCHECK(h, hipMalloc(&h->normL, cl * sizeof(float)));
    HIPCHECK(h, hipMalloc(&h->normR, cr * sizeof(float)));
    HIPCHECK(h, hipMalloc(&h->maxNormBits, 4 * sizeof(unsigned int)));
    HIPCHECK(h, hipMemsetAsync(h->maxNormBits, 0, 4 * sizeof(unsigned int), h->stream));
    HIPCHECK(h, hipMalloc(&h->colPart, (size_t)kMaxSplits * cl * sizeof(Top2)));
    HIPCHECK(h, hipMalloc(&h->rowPart, (cl / 128) * cr * sizeof(Top2)));
    HIPCHECK(h, hipMalloc(&h->colMax, cl * sizeof(int)));
    HIPCHECK(h, hipMalloc(&h->rowMax, cr * sizeof(int)));
    HIPCHECK(h, hipMalloc(&h->flagList, (cl + cr) * sizeof(int)));
    HIPCHECK(h, hipMalloc(&h->rpart, (cl + cr) * kRecheckMaxChunks * sizeof(Top2)));
    size_t const statusBytes = (cl / kMutualSlice + 1) * sizeof(unsigned long long);
    HIPCHECK(h, hipMalloc(&h->mstatus, statusBytes));
    HIPCHECK(h, hipMemsetAsync(h->mstatus, 0, statusBytes, h->stream));
    HIPCHECK(h, hipMalloc(&h->stripes, kStripes * kStripeStride * sizeof(int)));
    HIPCHECK(h, hipMemsetAsync(h->stripes, 0, kStripes * kStripeStride * sizeof(int), h->stream));
    h->pairCap = (int)std::min(cl, cr);
    HIPCHECK(h, hipMalloc(&h->results, (kNumCounters + 2 * (size_t)h->pairCap) * sizeof(int)));
    HIPCHECK(h, hipMemsetAsync(h->results, 0, (kNumCounters + 2 * (size_t)h->pairCap) * sizeof(int), h->stream));
    HIPCHECK(h, hipHostMalloc(&h->hostStage, (kNumCounters + 2 * (size_t)h->pairCap) * sizeof(int)));
    if (h->useGuided) {
        HIPCHECK(h, hipMalloc(&h->dLeftXY, cl * 2 * sizeof(float)));
        HIPCHECK(h, hipMalloc(&h->dRightXY, cr * 2 * sizeof(float)));
        // one bit per (left, right) pair: 1/32 of the reference's score matrix (cu:467)
        size_t const bitBytes = cl * (cr / 32) * sizeof(unsigned);
        HIPCHECK(h, hipMalloc(&h->guidedBits, bitBytes));
        HIPCHECK(h, hipMemsetAsync(h->guidedBits, 0, bitBytes, h->stream));
    }
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    h->allocated = true;
    return V3DM_OK;
}

int v3dm_deallocate(v3dm_handle h)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    HIPCHECK(h, hipSetDevice(h->device));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    ext_quiesce(h);
    releaseBuffers(h);
    h->nMaxL = h->nMaxR = 0;
    return V3DM_OK;
}

int v3dm_set_flags(v3dm_handle h, int warnOnRounding, int useReductionPath)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    h->warnOnRounding = warnOnRounding != 0;
    h->useReductionPath = useReductionPath != 0;
    return V3DM_OK;
}

}  // extern "C"

namespace {

enum Source { kHostF32, kDeviceF32, kHostU8 };

int setFeatures(v3dm_context* h, bool left, int n, const void* src, Source kind)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!h->allocated) return fail(h, V3DM_ERR_STATE, "set%sFeatures before allocate", left ? "Left" : "Right");
    if (n < 0) return fail(h, V3DM_ERR_INVALID, "set%sFeatures: negative count %d", left ? "Left" : "Right", n);
    n = roundWarn(h, n, left ? "setLeftFeatures()" : "setRightFeatures()", "nFeatures");
    n = std::min(n, left ? h->nMaxL : h->nMaxR);
    if (n > 0 && !src) return fail(h, V3DM_ERR_INVALID, "null descriptor pointer");
    HIPCHECK(h, hipSetDevice(h->device));
    ext_quiesce(h);  // the sampled check's side stream reads the features
    float* dst = left ? h->dL : h->dR;
    size_t const bytes = (size_t)n * kDim * sizeof(float);
    if (n > 0) {
        if (kind == kHostF32) {
            HIPCHECK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
        } else if (kind == kDeviceF32) {
            HIPCHECK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
        } else {
            size_t const qb = (size_t)(left ? h->capL : h->capR) * kDim;
            if (h->dQBytes < qb) {
                freeDev(h->dQ);
                HIPCHECK(h, hipMalloc(&h->dQ, qb));
                h->dQBytes = qb;
            }
            HIPCHECK(h, hipMemcpyAsync(h->dQ, src, (size_t)n * kDim, hipMemcpyHostToDevice, h->stream));
            launch_dequant_u8(h->dQ, n, dst, h->stream);
        }
    }
    unsigned int* mx = h->maxNormBits + (left ? 0 : 1);
    HIPCHECK(h, hipMemsetAsync(mx, 0, sizeof(unsigned int), h->stream));
    launch_norms(dst, n, left ? h->normL : h->normR, mx, h->stream);
    HIPCHECK(h, hipGetLastError());
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (left) h->nL = n;
    else h->nR = n;
    h->haveResult = false;
    h->masks.clear();
    ext_features_changed(h, left);
    return V3DM_OK;
}

int copyResults(v3dm_context* h, int32_t* pairs, int cap, int* nPairs)
{
    if (!h->haveResult) return fail(h, V3DM_ERR_STATE, "no result: call findPutativeMatches first");
    // One copy of the counters and of as many pairs as the previous result held (plus a
    // margin), enqueued before the wait for the sampled check so that both end in one round
    // trip; the rest only when the result is larger.
    size_t const headBytes = kNumCounters * sizeof(int);
    size_t const prev = (size_t)std::max(h->stats.nMatches, 0);
    size_t const guess = std::min((size_t)h->pairCap, prev + prev / 4 + 1024);
    HIPCHECK(h, hipMemcpyAsync(h->hostStage, h->results, headBytes + guess * 2 * sizeof(int), hipMemcpyDeviceToHost,
                               h->stream));
    int const rcf = ext_finish(h);
    if (rcf) return rcf;
    if (h->stats.nFallback)  // the plain form's result replaced the learned one
        HIPCHECK(h, hipMemcpyAsync(h->hostStage, h->results, headBytes + guess * 2 * sizeof(int),
                                   hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (h->hostStage[kCntMutualError] == (int)h->mutualEpoch)
        return fail(h, V3DM_ERR_HIP, "mutual check: ordering wait between work-groups ran out");
    int const n = h->hostStage[kCntPairs];
    if ((size_t)n > guess) {
        HIPCHECK(h, hipMemcpyAsync(h->hostStage + kNumCounters + 2 * guess, h->results + kNumCounters + 2 * guess,
                                   ((size_t)n - guess) * 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(h, hipStreamSynchronize(h->stream));
    }
    h->stats.nMatches = n;
    h->stats.nRechecked = h->hostStage[kCntFlagDone];
    h->stats.nRescored = h->hostStage[kCntRescoredDone];
    if (nPairs) *nPairs = n;
    int const ncopy = std::min(n, std::max(cap, 0));
    if (ncopy > 0) {
        if (!pairs) return fail(h, V3DM_ERR_INVALID, "null pairs buffer");
        memcpy(pairs, h->hostStage + kNumCounters, (size_t)ncopy * 2 * sizeof(int32_t));
    }
    if (n > cap) return fail(h, V3DM_ERR_CAPACITY, "pairs buffer holds %d pairs, %d matches found", cap, n);
    return V3DM_OK;
}

// runGuidedMatching / runGuidedHomographyMatching (cu:554-694).  The positions are
// uploaded (AoS x, y per feature), the call's mask is added to the bitmap of removed pairs
// (set on the first guided call after findPutativeMatches, so masks compound as the
// reference's in-place overwrites of its score matrix do), then the exact FP32 matcher runs
// over the current descriptors with removed pairs scored -1e20.  Guided calls always score
// in FP32, whatever the precision mode.
int runGuided(v3dm_context* h, int kind, const double M[9], float distanceThreshold, const float* leftXY,
              const float* rightXY, float minScore, float minRatio, int32_t* pairs, int cap, int* nPairs)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!h->allocated) return fail(h, V3DM_ERR_STATE, "guided matching before allocate");
    if (!h->useGuided || !h->guidedBits)
        return fail(h, V3DM_ERR_STATE, "guided matching needs allocate(..., useGuidedMatching = true)");
    if (!M) return fail(h, V3DM_ERR_INVALID, "null matrix");
    int const nL = h->nL, nR = h->nR;
    if ((nL > 0 && !leftXY) || (nR > 0 && !rightXY)) return fail(h, V3DM_ERR_INVALID, "null position array");
    HIPCHECK(h, hipSetDevice(h->device));
    hipStream_t const st = h->stream;
    // a result still pending (e.g. the sampled check of a learned result) is superseded
    ext_discard_result(h);
    float m[9];
    for (int t = 0; t < 9; ++t) m[t] = (float)M[t];  // cu:592-595 (double -> float3)
    float const sqrThreshold = distanceThreshold * distanceThreshold;
    bool const work = nL > 0 && nR > 0;
    if (work) {
        HIPCHECK(h, hipMemcpyAsync(h->dLeftXY, leftXY, (size_t)nL * 2 * sizeof(float), hipMemcpyHostToDevice, st));
        HIPCHECK(h, hipMemcpyAsync(h->dRightXY, rightXY, (size_t)nR * 2 * sizeof(float), hipMemcpyHostToDevice,
                                   st));
        launch_guided_mask(kind, m, sqrThreshold, h->dLeftXY, h->dRightXY, nL, nR, h->guidedBits, h->capL,
                           h->masks.empty(), st);
    }
    GuidedMask gm;
    gm.kind = kind;
    for (int t = 0; t < 9; ++t) gm.m[t] = m[t];
    gm.sqrThreshold = sqrThreshold;
    h->masks.push_back(gm);
    int rc;
    if (work) {
        FusedPlan const p = plan_fused_fp32(nL, nR, 512);
        if ((rc = prof_begin(h))) return rc;
        launch_guided_fp32(h->dL, h->dR, nL, nR, p, h->colPart, h->capL, h->rowPart, h->capR, h->guidedBits,
                           h->capL, st);
        if ((rc = prof_end(h))) return rc;
        launch_merge_decide(h->colPart, p.splits, h->capL, h->rowPart, p.leftBlocks, h->capR, nL, nR, h->normL,
                            h->normR, h->maxNormBits, h->epsRel, minScore, minRatio, h->colMax, h->rowMax,
                            h->flagList, h->capL, h->results, st);
        launch_recheck_exact(h->dL, h->dR, nL, nR, h->flagList, h->capL, h->results, minScore, minRatio, h->colMax,
                             h->rowMax, h->rpart, kRecheckGrid, st, h->guidedBits, h->capL);
    }
    launch_mutual_compact(h->colMax, h->rowMax, work ? nL : 0, nR, h->mstatus, next_mutual_epoch(h), h->stripes,
                          reinterpret_cast<int2*>(h->results + kNumCounters), h->results, st, {}, h->mutualWait);
    HIPCHECK(h, hipGetLastError());
    h->lastMinScore = minScore;
    h->lastMinRatio = minRatio;
    h->haveResult = true;
    h->lastNL = nL;
    h->lastNR = nR;
    h->stats.nLeft = nL;
    h->stats.nRight = nR;
    h->stats.precision = V3DM_PREC_FP32;
    h->stats.nFallback = 0;
    h->stats.nSampled = 0;
    h->stats.nSampleMissed = 0;
    return copyResults(h, pairs, cap, nPairs);
}

}  // namespace

extern "C" {

int v3dm_set_left(v3dm_handle h, int n, const float* desc) { return setFeatures(h, true, n, desc, kHostF32); }
int v3dm_set_right(v3dm_handle h, int n, const float* desc) { return setFeatures(h, false, n, desc, kHostF32); }
int v3dm_set_left_device(v3dm_handle h, int n, const float* d) { return setFeatures(h, true, n, d, kDeviceF32); }
int v3dm_set_right_device(v3dm_handle h, int n, const float* d) { return setFeatures(h, false, n, d, kDeviceF32); }
int v3dm_set_left_u8(v3dm_handle h, int n, const uint8_t* q) { return setFeatures(h, true, n, q, kHostU8); }
int v3dm_set_right_u8(v3dm_handle h, int n, const uint8_t* q) { return setFeatures(h, false, n, q, kHostU8); }

int v3dm_set_precision(v3dm_handle h, int mode)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (mode < V3DM_PREC_FP32 || mode > V3DM_PREC_LEARNED)
        return fail(h, V3DM_ERR_INVALID, "unknown precision mode %d", mode);
    if (!ext_supports(h, mode)) return fail(h, V3DM_ERR_UNSUPPORTED, "precision mode %d is not available", mode);
    h->precision = mode;
    return V3DM_OK;
}

int v3dm_set_certification(v3dm_handle h, float epsRel)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!(epsRel >= 0.0f) || !(epsRel < 1.0f)) return fail(h, V3DM_ERR_INVALID, "epsRel must be in [0, 1)");
    h->epsRel = epsRel;
    return V3DM_OK;
}

int v3dm_set_learned_topk(v3dm_handle h, int k)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (k < 1 || k > 32) return fail(h, V3DM_ERR_INVALID, "topK must be in [1, 32]");
    h->learnedTopK = k;
    return V3DM_OK;
}

int v3dm_set_learned_margin(v3dm_handle h, float marginRel)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (marginRel != marginRel) return fail(h, V3DM_ERR_INVALID, "margin must be a number");
    h->learnedMargin = marginRel;
    return V3DM_OK;
}

int v3dm_set_fallback(v3dm_handle h, int samples, float minRecall)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (samples < 0 || samples > kSampleRows || !(minRecall >= 0.0f) || !(minRecall <= 1.0f))
        return fail(h, V3DM_ERR_INVALID, "samples must be in [0, %d] and minRecall in [0, 1]", kSampleRows);
    h->fallbackSamples = samples;
    h->fallbackRecall = minRecall;
    return V3DM_OK;
}

int v3dm_set_fallback_rule(v3dm_handle h, int rule)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (rule != V3DM_FALLBACK_THRESHOLD && rule != V3DM_FALLBACK_BINOMIAL)
        return fail(h, V3DM_ERR_INVALID, "unknown fallback rule %d", rule);
    h->fallbackRule = rule;
    return V3DM_OK;
}

int v3dm_debug_mutual_wait(v3dm_handle h, int spinLimit, int neverReady)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    h->mutualWait.spinLimit = spinLimit < 0 ? kMutualSpinLimit : (unsigned)spinLimit;
    h->mutualWait.readXor = neverReady ? 0x80000000u : 0u;
    return V3DM_OK;
}

int v3dm_fallback_allowed_misses(int accepted, float minRecall)
{
    if (accepted < 0 || accepted > kFallbackMaxAccepted || !(minRecall >= 0.0f) || !(minRecall <= 1.0f)) return -1;
    // the largest m whose sampled recall (a - m) / a, formed in float, is not below minRecall
    int m = 0;
    while (m < accepted && !((float)(accepted - m - 1) / (float)accepted < minRecall)) ++m;
    return m;
}

int v3dm_fallback_allowed_misses_binomial(int accepted, float minRecall)
{
    if (accepted < 0 || accepted > kFallbackMaxAccepted || !(minRecall >= 0.0f) || !(minRecall <= 1.0f)) return -1;
    if (accepted == 0 || minRecall == 1.0f) return 0;
    if (minRecall == 0.0f) return accepted;
    // binomial tail by the pmf recurrence P[X = k + 1] = P[X = k] (a - k) / (k + 1) q / r,
    // in logarithms (r^a underflows for large a)
    double const r = minRecall, q = 1.0 - r;
    double const lqr = std::log(q / r);
    double lpmf = accepted * std::log(r);
    double cdf = std::exp(lpmf);
    int k = 0;
    while (k < accepted && 1.0 - cdf > kFallbackAlpha) {
        lpmf += std::log((double)(accepted - k) / (double)(k + 1)) + lqr;
        cdf += std::exp(lpmf);
        ++k;
    }
    return k;
}

int v3dm_set_derived_cache(v3dm_handle h, int enable)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    h->cacheDerived = enable != 0;
    return V3DM_OK;
}

int v3dm_set_stream(v3dm_handle h, void* stream)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    ext_quiesce(h);
    h->stream = stream ? (hipStream_t)stream : h->ownStream;
    return V3DM_OK;
}

int v3dm_find_putative_matches_async(v3dm_handle h, float minScore, float minRatio)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!h->allocated) return fail(h, V3DM_ERR_STATE, "findPutativeMatches before allocate");
    HIPCHECK(h, hipSetDevice(h->device));
    // a pending sampled check of the previous (learned) call is evaluated before this call
    // replaces its result -- a failing check reruns the plain form for it -- so that every
    // learned call pays for its check and any fallback, read or not
    int const rcf = ext_finish(h);
    if (rcf) return rcf;
    h->lastMinScore = minScore;
    h->lastMinRatio = minRatio;
    h->stats.nFallback = 0;
    h->stats.nSampled = 0;
    h->stats.nSampleMissed = 0;
    h->masks.clear();
    if (!h->cacheDerived) {
        ext_features_changed(h, true);
        ext_features_changed(h, false);
    }
    int rc;
    if (h->precision == V3DM_PREC_FP32) rc = enqueue_plain(h, minScore, minRatio);
    else rc = ext_enqueue(h, h->precision, minScore, minRatio);
    if (rc) return rc;
    h->haveResult = true;
    h->lastNL = h->nL;
    h->lastNR = h->nR;
    h->stats.nLeft = h->nL;
    h->stats.nRight = h->nR;
    h->stats.precision = h->precision;
    return V3DM_OK;
}

int v3dm_get_matches(v3dm_handle h, int32_t* pairs, int cap, int* nPairs)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    HIPCHECK(h, hipSetDevice(h->device));
    return copyResults(h, pairs, cap, nPairs);
}

int v3dm_find_putative_matches(v3dm_handle h, float minScore, float minRatio, int32_t* pairs, int cap, int* nPairs)
{
    int const rc = v3dm_find_putative_matches_async(h, minScore, minRatio);
    if (rc) return rc;
    return copyResults(h, pairs, cap, nPairs);
}

int v3dm_device_results(v3dm_handle h, const int32_t** d_pairs, const int32_t** d_counters)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!h->allocated) return fail(h, V3DM_ERR_STATE, "not allocated");
    if (d_pairs) *d_pairs = h->results + kNumCounters;
    if (d_counters) *d_counters = h->results;
    return V3DM_OK;
}

int v3dm_get_maxima(v3dm_handle h, int32_t* colMax, int32_t* rowMax)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!h->haveResult) return fail(h, V3DM_ERR_STATE, "no result: call findPutativeMatches first");
    HIPCHECK(h, hipSetDevice(h->device));
    int const rcf = ext_finish(h);
    if (rcf) return rcf;
    if (colMax && h->lastNL > 0)
        HIPCHECK(h, hipMemcpyAsync(colMax, h->colMax, (size_t)h->lastNL * sizeof(int), hipMemcpyDeviceToHost,
                                   h->stream));
    if (rowMax && h->lastNR > 0)
        HIPCHECK(h, hipMemcpyAsync(rowMax, h->rowMax, (size_t)h->lastNR * sizeof(int), hipMemcpyDeviceToHost,
                                   h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return V3DM_OK;
}

int v3dm_synchronize(v3dm_handle h)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    HIPCHECK(h, hipSetDevice(h->device));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    ext_synchronize(h);
    return V3DM_OK;
}

int v3dm_set_profiling(v3dm_handle h, int enable)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    HIPCHECK(h, hipSetDevice(h->device));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (enable && h->evStart.empty()) {
        h->evStart.resize(256);
        h->evStop.resize(256);
        for (int i = 0; i < 256; ++i) {
            HIPCHECK(h, hipEventCreate(&h->evStart[i]));
            HIPCHECK(h, hipEventCreate(&h->evStop[i]));
        }
    }
    h->profiling = enable != 0;
    h->evPending = 0;
    h->kernelMs = 0.0;
    h->kernelLaunches = 0;
    return V3DM_OK;
}

int v3dm_score_block(v3dm_handle h, int mode, int nRows, int nCols, float* scores)
{
    if (!h) return fail(nullptr, V3DM_ERR_INVALID, "null handle");
    if (!h->allocated) return fail(h, V3DM_ERR_STATE, "allocate first");
    if (mode != V3DM_PREC_FP32 && mode != V3DM_PREC_BF16 && mode != V3DM_PREC_FP8)
        return fail(h, V3DM_ERR_INVALID, "score blocks exist for the FP32, BF16 and FP8 passes");
    if (nRows < 0 || nRows > h->nL || nCols < 0 || nCols > h->nR)
        return fail(h, V3DM_ERR_INVALID, "block %d x %d outside the %d x %d feature sets", nRows, nCols, h->nL, h->nR);
    if (nRows == 0 || nCols == 0) return V3DM_OK;
    if (!scores) return fail(h, V3DM_ERR_INVALID, "null output");
    HIPCHECK(h, hipSetDevice(h->device));
    float* d = nullptr;
    size_t const bytes = (size_t)nRows * nCols * sizeof(float);
    HIPCHECK(h, hipMalloc(&d, bytes));
    launch_score_block(mode, h->dL, h->dR, nRows, nCols, d, h->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(scores, d, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(h, V3DM_ERR_HIP, "score block: %s", hipGetErrorString(e));
    return V3DM_OK;
}

int v3dm_time_dominant_kernel(v3dm_handle h, int reps, float* msPerLaunch)
{
    if (!h || !msPerLaunch || reps < 1) return fail(h, V3DM_ERR_INVALID, "bad argument");
    if (!h->haveResult) return fail(h, V3DM_ERR_STATE, "no result: call findPutativeMatches first");
    HIPCHECK(h, hipSetDevice(h->device));
    int const rcf = ext_finish(h);
    if (rcf) return rcf;
    hipEvent_t a = nullptr, b = nullptr;
    HIPCHECK(h, hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(h, V3DM_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
    }
    int launches = 0, rc = V3DM_OK;
    if (h->precision == V3DM_PREC_FP32) {
        if (h->nL > 0 && h->nR > 0) {
            FusedPlan const p = plan_fused_fp32(h->nL, h->nR, 512);
            e = hipEventRecord(a, h->stream);
            for (int r = 0; r < reps && e == hipSuccess; ++r)
                launch_fused_fp32(h->dL, h->dR, h->nL, h->nR, p, h->colPart, h->capL, h->rowPart, h->capR, h->stream);
            if (e == hipSuccess) e = hipEventRecord(b, h->stream);
            launches = reps;
        }
    } else {
        rc = ext_time_pass(h, h->precision, reps, a, b, &launches);
    }
    float ms = 0.0f;
    if (rc == V3DM_OK && e == hipSuccess && launches > 0) {
        e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(h, V3DM_ERR_HIP, "timing failed: %s", hipGetErrorString(e));
    if (launches == 0) return fail(h, V3DM_ERR_STATE, "no features to time");
    *msPerLaunch = ms / (float)launches;
    // the timing runs overwrote the partial results: the result of the last call stays in
    // h->results, but its per-feature decisions are recomputed only by the next call
    return V3DM_OK;
}

int v3dm_get_second_best(v3dm_handle h, int32_t* colSecond, int32_t* rowSecond)
{
    if (!h) return V3DM_ERR_INVALID;
    return ext_second_best(h, colSecond, rowSecond);
}

int v3dm_get_stats(v3dm_handle h, v3dm_stats* out)
{
    if (!h || !out) return fail(h, V3DM_ERR_INVALID, "null argument");
    HIPCHECK(h, hipSetDevice(h->device));
    if (h->profiling) {
        int const rc = harvestEvents(h);
        if (rc) return rc;
    }
    if (h->haveResult) {
        int const rcf = ext_finish(h);
        if (rcf) return rcf;
        int cnt[kNumCounters];
        HIPCHECK(h, hipMemcpyAsync(cnt, h->results, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        h->stats.nMatches = cnt[kCntPairs];
        h->stats.nRechecked = cnt[kCntFlagDone];
        h->stats.nRescored = cnt[kCntRescoredDone];
    }
    *out = h->stats;
    out->kernelMs = h->kernelMs;
    out->kernelLaunches = h->kernelLaunches;
    return V3DM_OK;
}

int v3dm_run_guided_F(v3dm_handle h, const double F[9], float distanceThreshold, const float* leftXY,
                      const float* rightXY, float minScore, float minRatio, int32_t* pairs, int cap, int* nPairs)
{
    return runGuided(h, 0, F, distanceThreshold, leftXY, rightXY, minScore, minRatio, pairs, cap, nPairs);
}

int v3dm_run_guided_H(v3dm_handle h, const double H[9], float distanceThreshold, const float* leftXY,
                      const float* rightXY, float minScore, float minRatio, int32_t* pairs, int cap, int* nPairs)
{
    return runGuided(h, 1, H, distanceThreshold, leftXY, rightXY, minScore, minRatio, pairs, cap, nPairs);
}

}  // extern "C"

namespace v3dm {

int fallback_allowed(v3dm_context const* h, int accepted)
{
    return h->fallbackRule == V3DM_FALLBACK_BINOMIAL ? v3dm_fallback_allowed_misses_binomial(accepted, h->fallbackRecall)
                                                     : v3dm_fallback_allowed_misses(accepted, h->fallbackRecall);
}

}  // namespace v3dm

