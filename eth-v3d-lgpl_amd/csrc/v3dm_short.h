// Shortlist forms: BF16 one-direction passes and exact rescoring.
#pragma once
#include <hip/hip_runtime.h>

namespace v3dm {

// Two best (approximate score, candidate index) of one query in one partition.
struct Cand2 {
    float v1;
    int j1;
    float v2;
    int j2;
};

// Packed lists of the single-pair learned code pass (8 bytes instead of 16): per group, the
// code score (|score| <= 32 * 31 * 31 < 2^15) in the high 16 bits and the candidate's index
// inside its split (< 65536: at most 1024 tiles of 64 per split) in the low 16 bits; score
// kPackedEmpty marks a group without a candidate.
struct CandP {
    unsigned int a, b;
};
constexpr int kPackedEmpty = -32768;
constexpr int kPackedMaxTiles = 1024;  // 64-candidate tiles per split of a packed list

struct ShortPlan {
    int queryBlocks = 0;
    int splits = 0;
    int tilesPerSplit = 0;
};

constexpr int kShortMaxSplits = 16;
constexpr int kRescoreSlots = 2 * 2 * kShortMaxSplits;  // candidates in every list of a query
constexpr int kRescoreTopKMax = 32;
constexpr int kShortRowPad = 256;  // bf16 descriptor buffers hold a multiple of 256 rows
// Per-score error bound of the BF16 scores: operands rounded to bf16 err by <= 2^-9
// relative, so |s_bf16 - s| <= (2^-8 + 2^-16) * ||q|| * ||c|| including FP32 accumulation
// and the 6 truncated key bits.  Candidates within twice the bound of the approximate
// second best are rescored.
constexpr float kWinRelBf16 = 2.0f * 1.01f * 0.00390625f;

// FP8 form: e4m3 operands with per-descriptor power-of-two scales err by <= 2^-4 relative
// each, so the worst case |s_fp8 - s| <= (2^-3 + 2^-8) * sum_k |q_k c_k| would rescore every
// listed candidate.  On normalised SIFT descriptors the error of a 128-term dot product has
// a standard deviation near 0.006 * ||q|| * ||c||, so the window is 2^-5 * ||q|| * max||c||
// (about five standard deviations); the recall it gives is measured against the exact form.
constexpr float kWinRelFp8 = 0.03125f;

void launch_to_bf16(float const* X, int rows, __bf16* Y, hipStream_t st);
ShortPlan plan_short(int nQ, int nC, int targetGroups);
// plan_short with at most kPackedMaxTiles tiles per split (every index inside a split fits
// 16 bits; at most 2^20 candidates need 16 splits)
ShortPlan plan_short_packed(int nQ, int nC, int targetGroups);
// both directions of a pair in one launch: A as queries against B (plan pa -> partA), B
// against A (pb -> partB)
void launch_pass_bf16_2(__bf16 const* A, int nA, __bf16 const* B, int nB, ShortPlan const& pa, Cand2* partA,
                        int ldA, ShortPlan const& pb, Cand2* partB, int ldB, hipStream_t st);
void launch_to_fp8(float const* X, int rows, unsigned char* Y, unsigned char* E, hipStream_t st);
void launch_pass_fp8_2(unsigned char const* A, unsigned char const* EA, int nA, unsigned char const* B,
                       unsigned char const* EB, int nB, ShortPlan const& pa, Cand2* partA, int ldA, ShortPlan const& pb,
                       Cand2* partB, int ldB, hipStream_t st);
// Exact rescoring (v3dm_rescore.hip); colSecond / rowSecond may be NULL.  topK = 0: the
// candidates within winRel * |q| * max|c| of the second best list value (certified forms);
// topK > 0: the topK best by list value, stopping at the first one more than `margin`
// below the best (INFINITY: no margin).
void launch_rescore(float const* L, float const* R, int nL, int nR, Cand2 const* partL, int nPartL, int ldqL,
                    Cand2 const* partR, int nPartR, int ldqR, float const* normL, float const* normR,
                    unsigned int const* maxNormBits, float winRel, int topK, float margin, float minScore,
                    float minRatio, int* colMax, int* rowMax, int* colSecond, int* rowSecond, int* stripes,
                    hipStream_t st);
// The same over packed lists (the learned code pass): spanL / spanR = candidates per split
// of the left / right queries' lists (the split's first candidate is split * span).
void launch_rescore_packed(float const* L, float const* R, int nL, int nR, CandP const* partL, int nPartL, int ldqL,
                           int spanL, CandP const* partR, int nPartR, int ldqR, int spanR, float const* normL,
                           float const* normR, unsigned int const* maxNormBits, int topK, float margin,
                           float minScore, float minRatio, int* colMax, int* rowMax, int* colSecond, int* rowSecond,
                           int* stripes, hipStream_t st);
// learned code pass (v3dm_code.hip), both directions in one launch: A as queries against B
// (plan pa -> partA), B against A; packed lists (CandP); plans with at most 1024 tiles per
// split
void launch_pass_code2(signed char const* A, int nA, signed char const* B, int nB, ShortPlan const& pa,
                       CandP* partA, int ldA, ShortPlan const& pb, CandP* partB, int ldB, hipStream_t st);

}  // namespace v3dm
