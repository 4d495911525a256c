// Internal state of one matcher handle (shared by the C ABI and the precision modes).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/v3dm.h"
#include "v3dm_device.h"
#include "v3dm_ext.h"
#include "v3dm_kernels.h"

struct v3dm_context {
    int device = 0;
    hipStream_t ownStream = nullptr;
    hipStream_t stream = nullptr;

    bool warnOnRounding = true;     // h:18
    bool useReductionPath = false;  // h:19
    bool useGuided = false;
    bool allocated = false;
    int precision = V3DM_PREC_FP32;
    float epsRel = v3dm::kEpsRelFp32;  // 0: every query goes to the exact re-check
    int learnedTopK = 6;               // candidates per query rescored in the learned mode
    float learnedMargin = 0.07f;       // ... and no more than this * scale^2 below the best (< 0: off)
    int fallbackSamples = v3dm::kSampleRows;  // sampled check of learned results (0: off)
    float fallbackRecall = 0.995f;     // sampled recall below this -> the plain form is run
    int fallbackRule = V3DM_FALLBACK_THRESHOLD;  // or V3DM_FALLBACK_BINOMIAL
    float lastMinScore = 0.9f, lastMinRatio = 0.8f;
    bool cacheDerived = true;          // keep bf16 / fp8 / code copies while features are unchanged

    int nMaxL = 0, nMaxR = 0;  // rounded maxima (cu:444-463)
    int nL = 0, nR = 0;        // current counts (cu:502, cu:518)
    int capL = 0, capR = 0;    // allocated rows (multiple of kRowPad)

    float* dL = nullptr;
    float* dR = nullptr;
    float* normL = nullptr;
    float* normR = nullptr;
    unsigned int* maxNormBits = nullptr;  // [0] = left, [1] = right
    v3dm::Top2* colPart = nullptr;
    v3dm::Top2* rowPart = nullptr;
    int* colMax = nullptr;
    int* rowMax = nullptr;
    int* flagList = nullptr;
    v3dm::Top2* rpart = nullptr;       // re-check chunk partials: (capL + capR) x kRecheckMaxChunks
    unsigned long long* mstatus = nullptr;  // mutual-check status words: capL / kMutualSlice + 1
    unsigned mutualEpoch = 0;               // epoch of the last mutual check (never 0 once run)
    v3dm::MutualWait mutualWait;            // the look-back's wait (v3dm_debug_mutual_wait)
    int* stripes = nullptr;      // striped counts (kStripes slots), cleared by the gather kernel
    int* results = nullptr;  // counters (kNumCounters ints) followed by int2 pairs
    int* hostStage = nullptr;
    int pairCap = 0;
    uint8_t* dQ = nullptr;  // uint8 upload staging (capL or capR rows)
    size_t dQBytes = 0;

    // guided matching state: masks applied since the last findPutativeMatches
    std::vector<v3dm::GuidedMask> masks;
    float* dLeftXY = nullptr;
    float* dRightXY = nullptr;
    unsigned* guidedBits = nullptr;  // removed pairs: bit j%32 of word (j/32) * capL + i

    // last result
    bool haveResult = false;
    v3dm_stats stats{};
    int lastNL = 0, lastNR = 0;

    // profiling of the dominant kernel
    bool profiling = false;
    std::vector<hipEvent_t> evStart, evStop;
    int evPending = 0;
    double kernelMs = 0.0;
    int kernelLaunches = 0;

    v3dm::ExtState* ext = nullptr;
    v3dm::BatchState* batch = nullptr;
    v3dm::LearnedState* learned = nullptr;

    std::string err;
};

namespace v3dm {

// Code-score margin of the learned mode's candidate selection: learnedMargin times the
// square of the code scale (the code score of a code with itself is about scale^2).
inline float learned_margin(v3dm_context* h)
{
    float const s = learned_code_scale(ctx_learned(h));
    return (h->learnedMargin < 0.0f || !(s > 0.0f)) ? INFINITY : h->learnedMargin * s * s;
}

// Misses a sampled check of `accepted` exact matches tolerates under the handle's rule.
int fallback_allowed(v3dm_context const* h, int accepted);

// The mutual check of the handle's next call: its status words and a fresh epoch.
inline unsigned next_mutual_epoch(v3dm_context* h)
{
    if (++h->mutualEpoch == 0) ++h->mutualEpoch;
    return h->mutualEpoch;
}

}  // namespace v3dm
