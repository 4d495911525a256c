// Host-side launchers of the matcher kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace v3dm {

struct Top2;

constexpr int kMaxSplits = 64;    // right-feature splits of the fused kernel
constexpr int kCntPairs = 0;      // counters[0]: number of matches
constexpr int kCntFlag = 1;       // counters[1]: open left queries of the running call (cleared at its end)
constexpr int kCntFlagDone = 2;   // counters[2]: queries re-checked exactly by the last call
constexpr int kCntRescoredDone = 4;  // counters[4]: ... by the last call
constexpr int kCntFlagR = 5;      // counters[5]: open right queries of the running call
constexpr int kCntMutualError = 6;  // counters[6]: epoch of a mutual check whose ordering wait ran out
constexpr int kNumCounters = 8;
// Counts summed by many work-groups go to kStripes slots one cache line apart (atomics on a
// single address serialise in L2); the gather kernel sums and clears them.
constexpr int kSampleRows = 128;  // sampled check of a shortlist result: at most one left block
constexpr int kSampleMaxSplits = 256;
constexpr double kFallbackAlpha = 0.05;  // level of the sampled check's one-sided test
constexpr int kFallbackMaxAccepted = 1 << 24;  // largest count v3dm_fallback_allowed_misses takes  // ... its right features split into up to 256 work-groups
constexpr int kStripes = 64;
constexpr int kStripeStride = 32;  // ints (128 bytes)
constexpr int kRecheckMaxChunks = 64;  // candidate chunks per re-checked query
constexpr int kRowPad = 128;      // device descriptor buffers hold a multiple of 128 rows
constexpr int kRecheckGrid = 1024;
// |approximate score - exact score| <= epsRel * ||q|| * max||c|| + kEpsAbs for the FP32
// MFMA path: 128-term FP32 dot products err by at most ~129 * 2^-23 * sum|q_k c_k|, and
// sum|q_k c_k| <= ||q|| ||c||; epsRel = 2^-15 leaves a factor 2 on top.  epsRel <= 0
// sends every query to the exact re-check.
constexpr float kEpsRelFp32 = 3.0517578125e-05f;  // 2^-15
constexpr float kEpsAbs = 1e-30f;

struct RecheckPlan {
    int nChunks = 1;
    int per = 1;  // candidates per thread of one chunk
};

struct FusedPlan {
    int leftBlocks = 0;
    int splits = 0;
    int tilesPerSplit = 0;
};

// selection stage (v3dm_select.hip)
void launch_norms(float const* X, int n, float* norm, unsigned int* maxNormBits, hipStream_t st);
void launch_merge_decide(Top2 const* colPart, int nColParts, int ldc, Top2 const* rowPart, int nRowParts, int ldr,
                         int nL, int nR, float const* normL, float const* normR, unsigned int const* maxNormBits,
                         float epsRel, float minScore, float minRatio, int* colMax, int* rowMax, int* flagList,
                         int flagOffR, int* counters, hipStream_t st);
RecheckPlan plan_recheck(int nC);
// flagList: open left queries at [0, ...), open right queries at [flagOffR, ...).  gbits:
// guided-matching bitmap (pairs scored -1e20), or null.
void launch_recheck_exact(float const* L, float const* R, int nL, int nR, int const* flagList, int flagOffR,
                          int const* counters, float minScore, float minRatio, int* colMax, int* rowMax, Top2* part,
                          int grid, hipStream_t st, unsigned const* gbits = nullptr, int ldi = 0);
// The learned decisions of the S sampled left features of the sampled exact check:
// out[k] = colMax[sample_row(k, nL, S)], k < S.  out may be mapped host memory.  S = 0:
// nothing to write.
struct SampleOut {
    int S = 0;
    int* out = nullptr;
};
// nL = 0 publishes an empty result.  status: ceil(nL / kMutualSlice) words (zero before
// the first call); epoch: nonzero and different from the previous call's.  The sampled
// learned decisions, when asked for, are written by work-group 0 (colMax is final by then).
constexpr int kMutualSlice = 1024;  // left features per work-group of the mutual check
// The look-back's bounded wait: at most spinLimit sleeps; a predecessor's status word counts
// when its epoch equals epoch ^ readXor (0 in every call but the test of the time-out path,
// v3dm_debug_mutual_wait, which makes no word count so that the wait runs out).
constexpr unsigned kMutualSpinLimit = 1u << 20;
struct MutualWait {
    unsigned spinLimit = kMutualSpinLimit;
    unsigned readXor = 0;
};
void launch_mutual_compact(int const* colMax, int const* rowMax, int nL, int nR, unsigned long long* status,
                           unsigned epoch, int* stripes, int2* pairs, int* counters, hipStream_t st,
                           SampleOut const& so = {}, MutualWait const& mw = {});
// uint8 SIFT descriptors -> floats, as SIFT_Feature::normalizeDescriptor
// (V3D/Base/v3d_feature.h:30-38): var = sum q^2, d = q * (1/sqrtf(var)), IEEE rounding.
void launch_dequant_u8(uint8_t const* q, int n, float* out, hipStream_t st);

// plain FP32 fused score + top-2 kernel (v3dm_fp32.hip)
FusedPlan plan_fused_fp32(int nL, int nR, int targetGroups, int maxSplits = kMaxSplits);
void launch_fused_fp32(float const* L, float const* R, int nL, int nR, FusedPlan const& p, Top2* colPart, int ldc,
                       Top2* rowPart, int ldr, hipStream_t st);

// guided matching (v3dm_guided.hip, v3dm_fp32.hip).  Bitmap of removed (left i, right j)
// pairs: bit (j & 31) of bits[(j >> 5) * ldi + i]; ldi >= the left rows, a multiple of 128.
// kind 0: Sampson distance with F (cu:277-311), 1: transfer distance with H (cu:313-343);
// m: the matrix rounded to float, row-major.  first: set the bitmap, else add to it.
void launch_guided_mask(int kind, float const (&m)[9], float sqrThreshold, float const* leftXY,
                        float const* rightXY, int nL, int nR, unsigned* bits, int ldi, bool first, hipStream_t st);
void launch_guided_fp32(float const* L, float const* R, int nL, int nR, FusedPlan const& p, Top2* colPart, int ldc,
                        Top2* rowPart, int ldr, unsigned const* bits, int ldi, hipStream_t st);

// measurement: the approximate scores of a form's pass, out[i * nCols + j] (v3dm_scores.hip;
// mode 0 FP32, 1 BF16, 2 FP8)
void launch_score_block(int mode, float const* L, float const* R, int nRows, int nCols, float* out, hipStream_t st);

// sampled exact check of a shortlist result (v3dm_select.hip)
// (also zeroes the check's counters)
void launch_sample_gather(float const* L, float const* normL, int nL, int S, float* Ls, float* normS, int* idx,
                          int* counters, hipStream_t st);
// merge + certified decision of the S sampled rows (maxNormBits[1]: max right norm)
void launch_sample_decide(Top2 const* colPart, int nParts, int ldc, int S, float const* normS,
                          unsigned int const* maxNormBits, float epsRel, float minScore, float minRatio, int* dec,
                          int* flagList, int* counters, hipStream_t st);

}  // namespace v3dm
