// Batched matching: many image pairs of uneven sizes in one pass over a device work queue.
//
// Device layout of a batch (all buffers owned by the handle, sized for the batch):
//  * images: the descriptors of every image, FP32 rows of 128 floats; image m starts at
//    row imgBase[m], a multiple of kBatchRowAlign, and its padding rows are zero.  A BF16
//    copy uses the same row numbering.
//  * query space: every (pair, side) is a segment of query rows; segment p (< nPairs) holds
//    the left queries of pair p, segment nPairs + p the right queries.  Segments start at
//    multiples of kBatchRowAlign; segOfBlock[q / kBatchRowAlign] names the segment of query
//    row q.  The per-query decisions (sel) and flags live in this space.
//  * work items: int4 records in a device array, taken in order by persistent work-groups
//    through an atomic counter; the host sorts them by decreasing cost.
#pragma once
#include <hip/hip_runtime.h>

#include "v3dm_device.h"
#include "v3dm_kernels.h"
#include "v3dm_short.h"

namespace v3dm {

constexpr int kBatchRowAlign = 256;
constexpr int kCodeSplitTiles = 32;  // learned lists: candidates per split, in 64-row tiles
constexpr int kBatchQueues = 4;  // work counters: 0 plain FP32 pass, 1 BF16 / learned pass
// Counters of a batch (an int array of kBatchCounters): queue heads, open-query count,
// rescored-candidate count, totals published by the gather kernel.
constexpr int kBCntFlag = 4;
constexpr int kBCntPairs = 6;
constexpr int kBCntFlagDone = 7;
constexpr int kBCntRescoredDone = 8;
constexpr int kBCntFbItems = 10;  // plain items of the pairs a learned run hands to the plain form
constexpr int kBQueueCheck0 = 3;  // work counter of the sampled check of FP32 batches
constexpr int kBCntOpenRescore = 12;  // open decisions of the shortlist rescoring
constexpr int kBatchCounters = 16;

struct BatchSeg {
    int qOff;     // first query row of the segment
    int n;        // queries (feature count of the query image, rounded down to 32)
    int img;      // query image
    int oImg;     // candidate image
    int nO;       // candidates
    int pair;
    int partOff;  // first partial of the segment (Top2 units for the FP32 form, Cand2 for shortlists)
    int nParts;   // partials per query
    int ld;       // stride between the partials of one query (padded n)
    int pad0, pad1, pad2;
};

// Plain FP32 item: x = pair, y = left block (128 rows), z = first tile, w = last tile (excl.)
// Shortlist item:  x = segment, y = query block (256 rows), z = split, w = tiles per split
// (64-row tiles of candidates); the learned code pass runs splits z .. nParts - 1.

// Misses the sampled check tolerates per number of exactly accepted samples (passed by
// value: allowed[a] = v3dm_fallback_allowed_misses(a, minRecall)).
struct MissLimits {
    unsigned char allowed[kSampleRows + 1];
};

// One 256-row query block of a shortlist plan, in pair-by-pair order (both sides of a pair
// in turn): everything the learned rescoring needs in one scalar load.
struct BatchBlock {
    int blk;      // the block: queries blk * 256 ..
    int q0;       // first query of the block within its segment
    int n, nO;    // the segment's queries and candidates
    int nParts, partOff, ld;
    int rowBase;  // image row of query q0
    int cBase;    // image row of candidate 0
    int pad0, pad1, pad2;
};

struct BatchDev {
    float const* img;
    __bf16 const* imgB;
    signed char const* codes;     // learned codes, 32 bytes per image row
    uint8_t const* u8;            // uint8 rows (images uploaded as uint8) or null
    float const* denom;           // their scales: element k = (float)u8[k] * denom
    int const* imgBase;
    float const* norm;            // per image row
    unsigned int const* maxNorm;  // per image (float bits)
    BatchSeg const* seg;
    int const* segOfBlock;
    BatchBlock const* blkInfo;    // shortlist plans: the 256-row query blocks, pair by pair
    int nPairs;
    int qRows;                    // size of the query space
};

// Sampled check of uint8 batches (v3dm_check.hip).  Per image row: rowCS = (128 * sum of
// the uint8 components, bits of the row's 1/sqrt scale); rows that are no candidate
// (all-zero rows, padding) have rowCS = (kDegenerateC, bits of 1).
constexpr int kDegenerateC = -(1 << 30);
constexpr int kChkMaxSplits = 1024;  // splits of one check query (64 Ki candidates at 8 tiles per split)

struct CheckDev {
    uint8_t const* u8;   // the batch's uint8 rows
    int2 const* rowCS;
    int4* part;          // keys per (phase, pair, split, query slot)
    int maxSplits;
    int tilesPerSplit[2];  // 32-row candidate tiles per split, per phase
    int const* nSplits;  // per (phase, pair)
    int* qB;             // phase 1 queries: matched features of B, per pair kSampleRows slots
    int* slotB;          // ... and the sample slot that matched each
    int* cntB;           // ... and their number
    int* decA;           // exact decisions of the phase 0 samples
    int samples;
};

// The side whose features the sampled check of pair p draws (0: left, 1: right): the one
// with more features, so that each sample scans the smaller image (left on a tie).
__device__ __forceinline__ int sample_side(BatchDev const& b, int p)
{
    return b.seg[p].n >= b.seg[b.nPairs + p].n ? 0 : 1;
}


// the BF16 copy of a uint8 batch from its uint8 rows and row scales (v3dm_batch.hip)
void launch_batch_u8_to_bf16(uint8_t const* u8, float const* denom, int rows, __bf16* Y, hipStream_t st);
// kernels of the plain FP32 form (v3dm_fp32.hip); leftRows: left features per work-group,
// 128 (4 waves) or 256 (8 waves), the plan's left blocks; nItemsDev (may be null): the item
// count in device memory, at most nItems
void launch_batch_fused_fp32(BatchDev const& b, int4 const* items, int nItems, int* queue, Top2* part, int grid,
                             int leftRows, hipStream_t st, int const* nItemsDev = nullptr);
// shortlist pass (v3dm_short.hip); queryRows: queries per work-group, 128 (2 waves) or 256
// (4 waves), the plan's query blocks
void launch_batch_pass_bf16(BatchDev const& b, int4 const* items, int nItems, int* queue, Cand2* part, int grid,
                            int queryRows, hipStream_t st);
// uint8 images: certified integer scores, the open decisions (listed in `open`, count in
// *nOpen, which must be zero) scored exactly (certify = false: every decision scored
// exactly); FP32 images: exact scores
void launch_batch_rescore(BatchDev const& b, Cand2 const* part, float winRel, int topK, float margin, float minScore,
                          float minRatio, bool certify, int* sel, int* open, int* nOpen, int* stripes, hipStream_t st);
// learned code pass (v3dm_code.hip); queryRows: queries per work-group, 128 or 256 (the
// plan's query blocks)
void launch_batch_pass_code(BatchDev const& b, int4 const* items, int nItems, int* queue, Cand2* part, int grid,
                            int queryRows, hipStream_t st);
// selection and bookkeeping (v3dm_batch.hip)
void launch_batch_reset(int* counters, int* pairCount, int nPairs, hipStream_t st);
void launch_batch_norms(float const* img, int const* rowImg, int rows, float* norm, unsigned int* maxNorm,
                        hipStream_t st);
void launch_batch_merge_decide(BatchDev const& b, Top2 const* part, float epsRel, float minScore, float minRatio,
                               int* sel, int* flagList, int* counters, hipStream_t st);
// the segments with segMask set only (the others keep their decisions); persistent grid
void launch_batch_merge_masked(BatchDev const& b, Top2 const* part, float epsRel, float minScore, float minRatio,
                               int* sel, int* flagList, int* counters, int const* segMask, int grid, hipStream_t st);
// sampled exact check of a shortlist run (v3dm_batch.hip, v3dm_fp32.hip)
void launch_batch_sample_gather(BatchDev const& b, int samples, float* sRows, float* sNorm, int* sIdx, int* sCount,
                                hipStream_t st);
void launch_batch_sample_fp32(BatchDev const& b, float const* sRows, int const* sCount, int4 const* items,
                              int nItems, int* queue, Top2* sPart, int maxSplits, int grid, hipStream_t st);
void launch_batch_sample_decide(BatchDev const& b, Top2 const* sPart, int maxSplits, int const* nSplits,
                                float const* sNorm, int const* sIdx, int const* sCount, float epsRel, float minScore,
                                float minRatio, MissLimits const& lim, int const* sel, int* pairFlag, int* counts,
                                hipStream_t st);
// uint8 batches: phase 0 (samples of the larger image) and phase 1 (their exact partners);
// the phase 1 decide sets pairFlag[p] and counts[4p..] = exact matches among the samples,
// missed, decisions taken exactly, samples
void launch_batch_check_probe(BatchDev const& b, CheckDev const& c, int phase, int4 const* items, int nItems,
                              int cus, hipStream_t st);
void launch_batch_check_decide(BatchDev const& b, CheckDev const& c, int phase, float minScore, float minRatio,
                               MissLimits const& lim, int const* sel, int* pairFlag, int* counts, hipStream_t st);
void launch_batch_recheck(BatchDev const& b, int const* flagList, int* counters, float minScore, float minRatio,
                          int* sel, int grid, hipStream_t st);
// the plain items of the pairs with pairFlag set -> fbItems (count in *fbCount, which must
// be zero), segMask[s] = flag of segment s's pair (2 * nPairs segments)
void launch_batch_fb_filter(int4 const* items, int nItems, int const* pairFlag, int nPairs, int4* fbItems,
                            int* fbCount, int* segMask, hipStream_t st);
void launch_batch_mutual(BatchDev const& b, int const* sel, int leftRows, int2* tmp, int* blockCount, int2* pairs,
                         int* pairCount, int* stripes, int* counters, hipStream_t st);

}  // namespace v3dm
