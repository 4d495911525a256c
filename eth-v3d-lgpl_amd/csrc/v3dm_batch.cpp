// Batched matching host side: image upload, pair list, work-item plans, autotuning of the
// item granularity and persistent grid, hipGraph capture of a batch run, result download.
//
// Replaces the caller loop of the reference (V3D/Apps/SfM/match_sequence.cpp:165-382,
// V3D/Apps/SfM/match_using_voctree.cpp:264-609), which sets the two feature sets of each
// pair and calls findPutativeMatches (V3D/CUDA/v3d_cudamatching.cu:490-552) once per pair.
// Per pair the result is the same: counts rounded down to a multiple of 32 (cu:493-502),
// matches sorted by left index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "v3dm_batch.h"
#include "v3dm_context.h"
#include "v3dm_learned.h"

namespace v3dm {

void launch_batch_load_u8(uint8_t const* src, int const* srcRow, int const* nImg, int const* blkImg,
                          int const* imgBase, int rows, float* img, uint8_t* imgU8, float* denom, int2* rowCS,
                          hipStream_t st);
void launch_batch_load_f32(float const* src, int const* srcRow, int const* nImg, int const* blkImg,
                           int const* imgBase, int rows, float* img, hipStream_t st);

namespace {

MissLimits miss_limits(v3dm_context const* h)
{
    MissLimits lim;
    for (int a = 0; a <= kSampleRows; ++a) lim.allowed[a] = (unsigned char)fallback_allowed(h, a);
    return lim;
}

int roundUp(int n, int m) { return (n + m - 1) / m * m; }

// 32-row candidate tiles per item of the sampled check's phases (phase 1 has few queries per
// pair: longer items)
constexpr int kCheckTiles[2] = {8, 32};

struct Plan {
    std::vector<int4> items;
    std::vector<BatchSeg> seg;
    std::vector<BatchBlock> blocks;  // shortlist plans: query blocks pair by pair
    size_t partSlots = 0;
    int tiles = 0;  // tiles per item asked for
    int grid = 0;
    int qRows = kBatchRowAlign;  // queries per work-group (shortlist plans: the query blocks)
    float passMs = -1.0f;  // autotuner measurement
};

}  // namespace

struct BatchState {
    // images
    int nImages = 0;
    std::vector<int> imgN;     // rounded counts
    std::vector<int> imgBaseH;
    int imgRows = 0, imgCapRows = 0;
    float* img = nullptr;
    __bf16* imgB = nullptr;
    bool bfValid = false;
    signed char* codes = nullptr;  // learned codes of every image row
    uint8_t* imgU8 = nullptr;      // uint8 rows (uint8 uploads), aligned layout
    float* denom = nullptr;
    int2* rowCS = nullptr;          // sampled check of uint8 batches: row constants
    bool haveU8 = false;
    bool codesValid = false;
    float codesScale = 0.0f;
    int* imgBase = nullptr;
    int* blkImg = nullptr;
    int* nImgD = nullptr;
    int* srcRow = nullptr;
    float* norm = nullptr;
    unsigned int* maxNorm = nullptr;
    void* upload = nullptr;  // device staging of the raw input
    size_t uploadCap = 0;
    void* pinned = nullptr;  // host staging
    size_t pinnedCap = 0;
    int imgCap = 0;

    // pairs
    int nPairs = 0;
    std::vector<int> pa, pb;
    int qRows = 0, leftRows = 0;
    std::vector<int> segOfBlockH;
    int* segOfBlock = nullptr;
    int segBlkCap = 0;

    // plans (0 = plain FP32, 1 = BF16 shortlist, 2 = learned codes) and their device copies
    // (3 = the plain plan of the pairs a learned run hands to the plain form: short items,
    // so that a few pairs still spread over the whole GPU)
    Plan plan[4];
    bool planValid[4] = {false, false, false, false};
    int4* items[4] = {nullptr, nullptr, nullptr, nullptr};
    int itemCap[4] = {0, 0, 0, 0};
    BatchSeg* segD[4] = {nullptr, nullptr, nullptr, nullptr};
    int segCap[4] = {0, 0, 0, 0};
    BatchBlock* blkInfo[4] = {nullptr, nullptr, nullptr, nullptr};
    int blkInfoCap[4] = {0, 0, 0, 0};
    Top2* partF = nullptr;
    size_t partFCap = 0;
    Cand2* partB = nullptr;
    size_t partBCap = 0;

    // selection / results
    int* sel = nullptr;
    int* flagList = nullptr;
    int* openList = nullptr;  // open decisions of the shortlist rescoring
    int selCap = 0;
    int* counters = nullptr;  // kBatchCounters ints
    int* stripes = nullptr;   // striped counts, cleared by the gather kernel
    int* pairCount = nullptr;
    int2* tmp = nullptr;
    int* blockCount = nullptr;
    int2* out = nullptr;
    int outCap = 0, pairCap = 0;
    int* hostRes = nullptr;  // pinned: counters + pair counts
    int hostResCap = 0;

    // sampled exact check of learned runs (kSampleRows left features per pair) and the
    // per-pair fallback to the plain form
    float* sRows = nullptr;
    float* sNorm = nullptr;
    int* sIdx = nullptr;
    int* sCount = nullptr;
    int* sNSplits = nullptr;
    Top2* sPart = nullptr;
    int4* sItems = nullptr;
    int nSItems = 0, sMaxSplits = 1;
    // uint8 batches: the two phases of the mutual sampled check (v3dm_check.hip)
    int4* cItems[2] = {nullptr, nullptr};
    int nCItems[2] = {0, 0};
    int* cNSplits = nullptr;
    int4* cPart = nullptr;
    int cMaxSplits = 1;
    int* cQB = nullptr;
    int* cSlotB = nullptr;
    int* cCntB = nullptr;
    int* cDecA = nullptr;
    int* pairFlag = nullptr;  // nPairs flags followed by 4 counts per pair
    int* hostChk = nullptr;   // pinned copy
    int4* fbItems = nullptr;
    int fbItemCap = 0;
    int* segMask = nullptr;
    bool checkPending = false;

    // autotuning and graphs
    bool autotune = true;
    int forceTiles[3] = {0, 0, 0}, forceGrid[3] = {0, 0, 0}, forceQRows[3] = {0, 0, 0};
    struct GraphKey {
        int mode;
        float minScore, minRatio, epsRel;
        int version;
        int topK;
        float margin;
        int samples;
        float minRecall;
        bool operator<(GraphKey const& o) const
        {
            return std::tie(mode, minScore, minRatio, epsRel, version, topK, margin, samples, minRecall) <
                   std::tie(o.mode, o.minScore, o.minRatio, o.epsRel, o.version, o.topK, o.margin, o.samples,
                            o.minRecall);
        }
    };
    std::map<GraphKey, hipGraphExec_t> graphs;
    int version = 0;  // bumped when the layout or a plan changes
    bool useGraph = true;

    // last run
    bool pending = false;
    int lastMode = 0;
    float lastMinScore = 0.9f, lastMinRatio = 0.8f;
    hipEvent_t evA = nullptr, evB = nullptr;
    float lastMs = 0.0f;
};

namespace {

#define BCHECK(h, call)                                                       \
    do {                                                                      \
        hipError_t const e_ = (call);                                         \
        if (e_ != hipSuccess) return ctx_fail((h), V3DM_ERR_HIP, #call);      \
    } while (0)

template <class T>
int grow(v3dm_context* h, T*& p, size_t& cap, size_t need)
{
    if (p && cap >= need) return V3DM_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t const n = std::max<size_t>(need, 1);
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return ctx_fail(h, V3DM_ERR_NOMEM, "device allocation failed");
    cap = n;
    return V3DM_OK;
}

template <class T>
int growi(v3dm_context* h, T*& p, int& cap, size_t need)
{
    size_t c = (size_t)cap;
    int const rc = grow(h, p, c, need);
    cap = (int)c;
    return rc;
}

void drop_graphs(BatchState* b)
{
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

BatchDev dev_view(BatchState* b, int form)
{
    BatchDev d;
    d.img = b->img;
    d.imgB = b->imgB;
    d.codes = b->codes;
    d.u8 = b->haveU8 ? b->imgU8 : nullptr;
    d.denom = b->denom;
    d.imgBase = b->imgBase;
    d.norm = b->norm;
    d.maxNorm = b->maxNorm;
    d.seg = b->segD[form];
    d.segOfBlock = b->segOfBlock;
    d.blkInfo = b->blkInfo[form];
    d.nPairs = b->nPairs;
    d.qRows = b->qRows;
    return d;
}

CheckDev check_view(BatchState* b, int samples)
{
    CheckDev c;
    c.u8 = b->imgU8;
    c.rowCS = b->rowCS;
    c.part = b->cPart;
    c.maxSplits = b->cMaxSplits;
    c.tilesPerSplit[0] = kCheckTiles[0];
    c.tilesPerSplit[1] = kCheckTiles[1];
    c.nSplits = b->cNSplits;
    c.qB = b->cQB;
    c.slotB = b->cSlotB;
    c.cntB = b->cCntB;
    c.decA = b->cDecA;
    c.samples = samples;
    return c;
}

// Query space and the segment skeleton (partials filled in by the plans).
void layout_segments(BatchState* b, std::vector<BatchSeg>& seg)
{
    int const P = b->nPairs;
    seg.assign(2 * P, BatchSeg{});
    int q = 0;
    for (int side = 0; side < 2; ++side) {
        if (side == 1) b->leftRows = q;
        for (int p = 0; p < P; ++p) {
            BatchSeg& s = seg[side * P + p];
            s.img = side == 0 ? b->pa[p] : b->pb[p];
            s.oImg = side == 0 ? b->pb[p] : b->pa[p];
            s.n = b->imgN[s.img];
            s.nO = b->imgN[s.oImg];
            s.pair = p;
            s.qOff = q;
            q += roundUp(std::max(s.n, 1), kBatchRowAlign);
        }
    }
    b->qRows = q;
}

// Plain FP32 plan: items (pair, left block, split, tiles) with `tiles` 64-row right tiles
// per item; partials: left queries one per split, right queries one per left block.
// Plain plan: items (pair, left block of leftRows features, split, tiles per split).
Plan plan_plain(BatchState* b, int tiles, int grid, int leftRows = 128)
{
    Plan pl;
    pl.tiles = tiles;
    pl.grid = grid;
    pl.qRows = leftRows;
    layout_segments(b, pl.seg);
    int const P = b->nPairs;
    size_t off = 0;
    std::vector<std::pair<int, int4>> work;
    for (int p = 0; p < P; ++p) {
        BatchSeg& sl = pl.seg[p];
        BatchSeg& sr = pl.seg[P + p];
        int const nL = sl.n, nR = sr.n;
        int const padL = roundUp(std::max(nL, 1), kBatchRowAlign), padR = roundUp(std::max(nR, 1), kBatchRowAlign);
        sl.ld = padL;
        sr.ld = padR;
        if (nL <= 0 || nR <= 0) {
            sl.nParts = sr.nParts = 0;
            sl.partOff = sr.partOff = 0;
            continue;
        }
        int const nTiles = (nR + 63) / 64;
        int const splits = (nTiles + tiles - 1) / tiles;
        int const blocks = (nL + leftRows - 1) / leftRows;  // leftRows divides kBatchRowAlign
        sl.nParts = splits;
        sl.partOff = (int)off;
        off += (size_t)splits * padL;
        sr.nParts = blocks;
        sr.partOff = (int)off;
        off += (size_t)blocks * padR;
        for (int lb = 0; lb < blocks; ++lb)
            for (int s = 0; s < splits; ++s) {
                int const cost = std::min(tiles, nTiles - s * tiles);
                work.push_back({cost, make_int4(p, lb, s, tiles)});
            }
    }
    std::stable_sort(work.begin(), work.end(),
                     [](std::pair<int, int4> const& x, std::pair<int, int4> const& y) { return x.first > y.first; });
    pl.items.reserve(work.size());
    for (auto& w : work) pl.items.push_back(w.second);
    pl.partSlots = off;
    return pl;
}

// Shortlist plan: items (segment, query block, split, tiles) with at least `tiles` 64-row
// candidate tiles per item and at most maxSplits (<= kShortMaxSplits) splits per query.
// wholeRows: one item per (segment, query block) that runs every split of the row in turn
// (the learned code pass); else one item per split.
Plan plan_short_batch(BatchState* b, int tiles, int grid, int maxSplits, bool wholeRows, int qRows)
{
    Plan pl;
    pl.tiles = tiles;
    pl.grid = grid;
    pl.qRows = qRows;
    layout_segments(b, pl.seg);
    size_t off = 0;
    std::vector<std::pair<int, int4>> work;
    for (int si = 0; si < (int)pl.seg.size(); ++si) {
        BatchSeg& s = pl.seg[si];
        s.ld = roundUp(std::max(s.n, 1), kBatchRowAlign);
        if (s.n <= 0 || s.nO <= 0) {
            s.nParts = 0;
            s.partOff = 0;
            continue;
        }
        int const nTiles = (s.nO + 63) / 64;
        int const tps = std::max(tiles, (nTiles + maxSplits - 1) / maxSplits);
        int const splits = (nTiles + tps - 1) / tps;
        s.nParts = splits;
        s.partOff = (int)off;
        off += (size_t)splits * s.ld;
        int const qBlocks = (s.n + qRows - 1) / qRows;  // qRows divides kBatchRowAlign: blocks stay in ld
        for (int qb = 0; qb < qBlocks; ++qb) {
            if (wholeRows) {
                work.push_back({nTiles, make_int4(si, qb, 0, tps)});
                continue;
            }
            for (int sp = 0; sp < splits; ++sp) {
                int const cost = std::min(tps, nTiles - sp * tps);
                work.push_back({cost, make_int4(si, qb, sp, tps)});
            }
        }
    }
    std::stable_sort(work.begin(), work.end(),
                     [](std::pair<int, int4> const& x, std::pair<int, int4> const& y) { return x.first > y.first; });
    pl.items.reserve(work.size());
    for (auto& w : work) pl.items.push_back(w.second);
    pl.partSlots = off;
    // the query blocks pair by pair, both sides of a pair next to each other: the learned
    // rescoring's queries of one side are the other side's candidates, so a pair's rows are
    // fetched into an XCD's L2 once
    int const P = b->nPairs;
    for (int p = 0; p < P; ++p)
        for (int side = 0; side < 2; ++side) {
            BatchSeg const& sg = pl.seg[side * P + p];
            int const nb = roundUp(std::max(sg.n, 1), kBatchRowAlign) / kBatchRowAlign;
            for (int k = 0; k < nb; ++k) {
                BatchBlock bb{};
                bb.blk = sg.qOff / kBatchRowAlign + k;
                bb.q0 = k * kBatchRowAlign;
                bb.n = sg.n;
                bb.nO = sg.nO;
                bb.nParts = sg.nParts;
                bb.partOff = sg.partOff;
                bb.ld = sg.ld;
                bb.rowBase = b->imgBaseH[sg.img] + bb.q0;
                bb.cBase = b->imgBaseH[sg.oImg];
                pl.blocks.push_back(bb);
            }
        }
    return pl;
}

int upload_plan(v3dm_context* h, BatchState* b, int form, Plan const& pl)
{
    int rc;
    if ((rc = growi(h, b->items[form], b->itemCap[form], pl.items.size()))) return rc;
    if ((rc = growi(h, b->segD[form], b->segCap[form], pl.seg.size()))) return rc;
    if ((rc = growi(h, b->blkInfo[form], b->blkInfoCap[form], pl.blocks.size()))) return rc;
    if (form == 0 || form == 3) {
        if ((rc = grow(h, b->partF, b->partFCap, pl.partSlots))) return rc;
    } else {
        if ((rc = grow(h, b->partB, b->partBCap, pl.partSlots * 2))) return rc;
    }
    hipStream_t const st = h->stream;
    if (!pl.items.empty())
        BCHECK(h, hipMemcpyAsync(b->items[form], pl.items.data(), pl.items.size() * sizeof(int4),
                                 hipMemcpyHostToDevice, st));
    if (!pl.seg.empty())
        BCHECK(h, hipMemcpyAsync(b->segD[form], pl.seg.data(), pl.seg.size() * sizeof(BatchSeg),
                                 hipMemcpyHostToDevice, st));
    if (!pl.blocks.empty())
        BCHECK(h, hipMemcpyAsync(b->blkInfo[form], pl.blocks.data(), pl.blocks.size() * sizeof(BatchBlock),
                                 hipMemcpyHostToDevice, st));
    BCHECK(h, hipStreamSynchronize(st));  // host vectors may change after return
    return V3DM_OK;
}

// Items of the sampled check.  uint8 images (the mutual check of v3dm_check.hip): per phase
// up to kCheckTiles 32-row tiles of the image the phase scans (phase 0: the pair's smaller
// image, phase 1: the larger), each item as two int4 with every row offset resolved here.
// FP32 images: (pair, 0, split, 8 tiles of 64 features of the smaller image) for the
// samples' own decisions.
int ensure_sample_plan(v3dm_context* h, BatchState* b)
{
    int const P = b->nPairs;
    std::vector<BatchSeg> seg;
    layout_segments(b, seg);
    for (void* ptr : {(void*)b->sRows, (void*)b->sNorm, (void*)b->sIdx, (void*)b->sCount, (void*)b->sNSplits,
                      (void*)b->sPart, (void*)b->sItems, (void*)b->pairFlag, (void*)b->segMask, (void*)b->cItems[0],
                      (void*)b->cItems[1], (void*)b->cNSplits, (void*)b->cPart, (void*)b->cQB, (void*)b->cSlotB,
                      (void*)b->cCntB, (void*)b->cDecA})
        if (ptr) (void)hipFree(ptr);
    b->sRows = b->sNorm = nullptr;
    b->sIdx = b->sCount = b->sNSplits = nullptr;
    b->sPart = nullptr;
    b->sItems = nullptr;
    b->cItems[0] = b->cItems[1] = nullptr;
    b->cNSplits = b->cQB = b->cSlotB = b->cCntB = b->cDecA = nullptr;
    b->cPart = nullptr;
    b->nSItems = b->nCItems[0] = b->nCItems[1] = 0;
    if (b->hostChk) (void)hipHostFree(b->hostChk);
    b->hostChk = nullptr;
    auto sorted_items = [](std::vector<std::pair<int, int4>>& work) {
        std::stable_sort(work.begin(), work.end(),
                         [](std::pair<int, int4> const& x, std::pair<int, int4> const& y) { return x.first > y.first; });
        std::vector<int4> items;
        for (auto& w : work) items.push_back(w.second);
        return items;
    };
    if (b->haveU8) {
        std::vector<int> nSplits(2 * P, 0);
        int maxS = 1;
        std::vector<int4> items[2];
        for (int phase = 0; phase < 2; ++phase) {
            // two int4 per item (the probe reads them in one go): (pair, split, first tile,
            // tiles) and (row of the first candidate tile, first row of the query image, nA, 0)
            std::vector<std::pair<int, int4>> work, rec;
            for (int p = 0; p < P; ++p) {
                int const nL = seg[p].n, nR = seg[P + p].n;
                if (nL <= 0 || nR <= 0) continue;
                int const nA = std::max(nL, nR), nB = std::min(nL, nR);
                int const sa = nL >= nR ? 0 : 1;  // the device's sample_side
                int const imgA = seg[sa * P + p].img, imgB = seg[(1 - sa) * P + p].img;
                int const cBase = b->imgBaseH[phase == 0 ? imgB : imgA];
                int const qBase = b->imgBaseH[phase == 0 ? imgA : imgB];
                int const nTiles = (phase == 0 ? nB : nA) / 32;
                int const tps = kCheckTiles[phase];
                int const splits = (nTiles + tps - 1) / tps;
                nSplits[phase * P + p] = splits;
                maxS = std::max(maxS, splits);
                for (int sp = 0; sp < splits; ++sp) {
                    int const t0 = sp * tps, nT = std::min(tps, nTiles - t0);
                    work.push_back({nT, make_int4(p, sp, t0, nT)});
                    rec.push_back({nT, make_int4(cBase + 32 * t0, qBase, nA, 0)});
                }
            }
            std::vector<int4> const first = sorted_items(work), second = sorted_items(rec);
            for (size_t i = 0; i < first.size(); ++i) {
                items[phase].push_back(first[i]);
                items[phase].push_back(second[i]);
            }
        }
        size_t const slots = (size_t)P * kSampleRows;
        for (int phase = 0; phase < 2; ++phase) {
            BCHECK(h, hipMalloc(&b->cItems[phase], std::max<size_t>(items[phase].size(), 1) * sizeof(int4)));
            if (!items[phase].empty())
                BCHECK(h, hipMemcpy(b->cItems[phase], items[phase].data(), items[phase].size() * sizeof(int4),
                                    hipMemcpyHostToDevice));
            b->nCItems[phase] = (int)items[phase].size() / 2;
        }
        BCHECK(h, hipMalloc(&b->cNSplits, 2 * P * sizeof(int)));
        BCHECK(h, hipMemcpy(b->cNSplits, nSplits.data(), 2 * P * sizeof(int), hipMemcpyHostToDevice));
        BCHECK(h, hipMalloc(&b->cPart, 2 * slots * maxS * sizeof(int4)));
        BCHECK(h, hipMalloc(&b->cQB, slots * sizeof(int)));
        BCHECK(h, hipMalloc(&b->cSlotB, slots * sizeof(int)));
        BCHECK(h, hipMalloc(&b->cDecA, slots * sizeof(int)));
        BCHECK(h, hipMalloc(&b->cCntB, P * sizeof(int)));
        b->cMaxSplits = maxS;
    } else {
        std::vector<std::pair<int, int4>> work;
        std::vector<int> nSplits(P, 0);
        int const tps = 8;
        int maxS = 1;
        for (int p = 0; p < P; ++p) {
            int const nL = seg[p].n, nR = seg[P + p].n;
            if (nL <= 0 || nR <= 0) continue;
            int const nTiles = ((nL >= nR ? nR : nL) + 63) / 64;
            int const splits = (nTiles + tps - 1) / tps;
            nSplits[p] = splits;
            maxS = std::max(maxS, splits);
            for (int sp = 0; sp < splits; ++sp)
                work.push_back({std::min(tps, nTiles - sp * tps), make_int4(p, 0, sp, tps)});
        }
        std::vector<int4> items = sorted_items(work);
        size_t const rows = (size_t)P * kSampleRows;
        BCHECK(h, hipMalloc(&b->sRows, rows * kDim * sizeof(float)));
        BCHECK(h, hipMalloc(&b->sNorm, rows * sizeof(float)));
        BCHECK(h, hipMalloc(&b->sIdx, rows * sizeof(int)));
        BCHECK(h, hipMalloc(&b->sCount, P * sizeof(int)));
        BCHECK(h, hipMalloc(&b->sNSplits, P * sizeof(int)));
        BCHECK(h, hipMalloc(&b->sPart, rows * maxS * sizeof(Top2)));
        BCHECK(h, hipMalloc(&b->sItems, std::max<size_t>(items.size(), 1) * sizeof(int4)));
        BCHECK(h, hipMemcpy(b->sNSplits, nSplits.data(), P * sizeof(int), hipMemcpyHostToDevice));
        if (!items.empty())
            BCHECK(h, hipMemcpy(b->sItems, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
        b->nSItems = (int)items.size();
        b->sMaxSplits = maxS;
    }
    BCHECK(h, hipMalloc(&b->pairFlag, 5 * P * sizeof(int)));
    BCHECK(h, hipMalloc(&b->segMask, 2 * P * sizeof(int)));
    BCHECK(h, hipMemset(b->pairFlag, 0, 5 * P * sizeof(int)));
    BCHECK(h, hipHostMalloc(&b->hostChk, 5 * P * sizeof(int), hipHostMallocDefault));
    std::memset(b->hostChk, 0, 5 * P * sizeof(int));
    b->checkPending = false;
    return V3DM_OK;
}

int ensure_query_space(v3dm_context* h, BatchState* b)
{
    std::vector<BatchSeg> seg;
    layout_segments(b, seg);
    int const nBlk = b->qRows / kBatchRowAlign;
    b->segOfBlockH.assign(nBlk, 0);
    for (int si = 0; si < (int)seg.size(); ++si) {
        int const nb = roundUp(std::max(seg[si].n, 1), kBatchRowAlign) / kBatchRowAlign;
        for (int k = 0; k < nb; ++k) b->segOfBlockH[seg[si].qOff / kBatchRowAlign + k] = si;
    }
    int rc;
    if ((rc = growi(h, b->segOfBlock, b->segBlkCap, nBlk))) return rc;
    if ((rc = growi(h, b->sel, b->selCap, b->qRows))) return rc;
    if (b->flagList) (void)hipFree(b->flagList);
    if (b->openList) (void)hipFree(b->openList);
    b->flagList = b->openList = nullptr;
    BCHECK(h, hipMalloc(&b->flagList, (size_t)std::max(b->qRows, 1) * sizeof(int)));
    BCHECK(h, hipMalloc(&b->openList, (size_t)std::max(b->qRows, 1) * sizeof(int)));
    int const nbMutual = (b->leftRows + 1023) / 1024;
    if (b->tmp) (void)hipFree(b->tmp);
    if (b->blockCount) (void)hipFree(b->blockCount);
    b->tmp = nullptr;
    b->blockCount = nullptr;
    BCHECK(h, hipMalloc(&b->tmp, (size_t)std::max(nbMutual, 1) * 1024 * sizeof(int2)));
    BCHECK(h, hipMalloc(&b->blockCount, (size_t)std::max(nbMutual, 1) * 2 * sizeof(int)));  // counts, offsets
    // every left query gives at most one match
    if ((rc = growi(h, b->out, b->outCap, std::max(b->leftRows, 1)))) return rc;
    if ((rc = growi(h, b->pairCount, b->pairCap, std::max(b->nPairs, 1)))) return rc;
    if (!b->counters) BCHECK(h, hipMalloc(&b->counters, kBatchCounters * sizeof(int)));
    if (!b->stripes) {
        BCHECK(h, hipMalloc(&b->stripes, kStripes * kStripeStride * sizeof(int)));
        BCHECK(h, hipMemsetAsync(b->stripes, 0, kStripes * kStripeStride * sizeof(int), h->stream));
    }
    BCHECK(h, hipMemsetAsync(b->counters, 0, kBatchCounters * sizeof(int), h->stream));
    if (b->hostRes && b->hostResCap < kBatchCounters + b->nPairs) {
        (void)hipHostFree(b->hostRes);
        b->hostRes = nullptr;
    }
    if (!b->hostRes) {
        b->hostResCap = kBatchCounters + std::max(b->nPairs, 1);
        BCHECK(h, hipHostMalloc(&b->hostRes, b->hostResCap * sizeof(int), hipHostMallocDefault));
    }
    if (nBlk > 0) {
        BCHECK(h, hipMemcpyAsync(b->segOfBlock, b->segOfBlockH.data(), nBlk * sizeof(int), hipMemcpyHostToDevice,
                                 h->stream));
    }
    BCHECK(h, hipStreamSynchronize(h->stream));
    return V3DM_OK;
}

int form_of(int mode) { return mode == V3DM_PREC_FP32 ? 0 : (mode == V3DM_PREC_LEARNED ? 2 : 1); }

// Derived copies of the images for a form: bf16 rows (BF16 form) or learned codes.
void enqueue_derived(v3dm_context* h, BatchState* b, int form)
{
    hipStream_t const st = h->stream;
    if (form == 1) {
        if (b->haveU8) launch_batch_u8_to_bf16(b->imgU8, b->denom, b->imgRows, b->imgB, st);
        else launch_to_bf16(b->img, b->imgRows, b->imgB, st);
    } else if (form == 2) {
        MlpWeights const& w = *learned_weights(ctx_learned(h));
        float const scale = learned_code_scale(ctx_learned(h));
        if (b->haveU8) launch_mlp_codes_u8(b->imgU8, b->denom, b->imgRows, w, scale, b->codes, st);
        else launch_mlp_codes_f32(b->img, b->imgRows, w, scale, b->codes, st);
    }
}

// Enqueue one batch run of `mode` with the current plan of its form.
int enqueue_run(v3dm_context* h, BatchState* b, int mode, float minScore, float minRatio)
{
    int const form = form_of(mode);
    hipStream_t const st = h->stream;
    BatchDev const d = dev_view(b, form);
    Plan const& pl = b->plan[form];
    launch_batch_reset(b->counters, b->pairCount, b->nPairs, st);
    if (!h->cacheDerived) enqueue_derived(h, b, form);
    if (form == 0) {
        launch_batch_fused_fp32(d, b->items[0], (int)pl.items.size(), b->counters + 0, b->partF, pl.grid, pl.qRows,
                                st);
        launch_batch_merge_decide(d, b->partF, h->epsRel, minScore, minRatio, b->sel, b->flagList, b->counters, st);
        launch_batch_recheck(d, b->flagList, b->counters, minScore, minRatio, b->sel, 1024, st);
    } else if (form == 1) {
        launch_batch_pass_bf16(d, b->items[1], (int)pl.items.size(), b->counters + 1, b->partB, pl.grid, pl.qRows,
                               st);
        launch_batch_rescore(d, b->partB, kWinRelBf16, 0, INFINITY, minScore, minRatio, h->epsRel > 0.0f, b->sel,
                             b->openList, b->counters + kBCntOpenRescore, b->stripes, st);
    } else {
        launch_batch_pass_code(d, b->items[2], (int)pl.items.size(), b->counters + 2, b->partB, pl.grid, pl.qRows,
                               st);
        launch_batch_rescore(d, b->partB, 0.0f, h->learnedTopK, learned_margin(h), minScore, minRatio,
                             h->epsRel > 0.0f, b->sel, b->openList, b->counters + kBCntOpenRescore, b->stripes, st);
        if (h->fallbackSamples > 0) {
            int dev = 0, cus = 256;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            MissLimits const lim = miss_limits(h);
            if (b->haveU8) {
                // exact mutual matches among the samples against the learned result
                CheckDev const c = check_view(b, h->fallbackSamples);
                launch_batch_check_probe(d, c, 0, b->cItems[0], b->nCItems[0], cus, st);
                launch_batch_check_decide(d, c, 0, minScore, minRatio, lim, b->sel, b->pairFlag,
                                          b->pairFlag + b->nPairs, st);
                launch_batch_check_probe(d, c, 1, b->cItems[1], b->nCItems[1], cus, st);
                launch_batch_check_decide(d, c, 1, minScore, minRatio, lim, b->sel, b->pairFlag,
                                          b->pairFlag + b->nPairs, st);
            } else {
                // FP32 images: the samples' own decisions against the learned ones
                launch_batch_sample_gather(d, h->fallbackSamples, b->sRows, b->sNorm, b->sIdx, b->sCount, st);
                launch_batch_sample_fp32(d, b->sRows, b->sCount, b->sItems, b->nSItems, b->counters + kBQueueCheck0, b->sPart,
                                         b->sMaxSplits, 2 * cus, st);
                launch_batch_sample_decide(d, b->sPart, b->sMaxSplits, b->sNSplits, b->sNorm, b->sIdx, b->sCount,
                                           h->epsRel, minScore, minRatio, lim, b->sel, b->pairFlag,
                                           b->pairFlag + b->nPairs, st);
            }
            // pairs whose sampled recall fell short get the plain form's decisions, on the
            // device: their plain items, merge masked to their segments, exact re-check
            // (kernels with nothing to do when no pair is flagged)
            Plan const& pp = b->plan[3];
            int const np = (int)pp.items.size();
            BatchDev const d0 = dev_view(b, 3);
            launch_batch_fb_filter(b->items[3], np, b->pairFlag, b->nPairs, b->fbItems, b->counters + kBCntFbItems,
                                   b->segMask, st);
            launch_batch_fused_fp32(d0, b->fbItems, np, b->counters + 0, b->partF, pp.grid, pp.qRows, st,
                                    b->counters + kBCntFbItems);
            launch_batch_merge_masked(d0, b->partF, h->epsRel, minScore, minRatio, b->sel, b->flagList, b->counters,
                                      b->segMask, 2 * cus, st);
            launch_batch_recheck(d0, b->flagList, b->counters, minScore, minRatio, b->sel, 1024, st);
        }
    }
    launch_batch_mutual(d, b->sel, b->leftRows, b->tmp, b->blockCount, b->out, b->pairCount, b->stripes,
                        b->counters, st);
    BCHECK(h, hipGetLastError());
    return V3DM_OK;
}

// Time the pass kernel of `form` under plan `pl` (best of two runs), in ms.
int time_pass(v3dm_context* h, BatchState* b, int form, Plan const& pl, float& ms)
{
    int rc = upload_plan(h, b, form, pl);
    if (rc) return rc;
    b->plan[form] = pl;
    hipStream_t const st = h->stream;
    BatchDev const d = dev_view(b, form);
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        launch_batch_reset(b->counters, b->pairCount, b->nPairs, st);
        BCHECK(h, hipEventRecord(b->evA, st));
        if (form == 0)
            launch_batch_fused_fp32(d, b->items[0], (int)pl.items.size(), b->counters + 0, b->partF, pl.grid,
                                    pl.qRows, st);
        else if (form == 1)
            launch_batch_pass_bf16(d, b->items[1], (int)pl.items.size(), b->counters + 1, b->partB, pl.grid,
                                   pl.qRows, st);
        else
            launch_batch_pass_code(d, b->items[2], (int)pl.items.size(), b->counters + 2, b->partB, pl.grid, pl.qRows,
                               st);
        BCHECK(h, hipEventRecord(b->evB, st));
        BCHECK(h, hipEventSynchronize(b->evB));
        float t = 0.0f;
        BCHECK(h, hipEventElapsedTime(&t, b->evA, b->evB));
        if (rep > 0) best = std::min(best, t);  // the first run warms caches and code
    }
    ms = best;
    return V3DM_OK;
}

// Choose (tiles per item, persistent grid, and the work-group's tile shape: 128 or 256 query
// rows against 64-candidate tiles, i.e. 4 or 8 waves for the plain form, 2 or 4 waves for
// the BF16 and code passes) for a form: the given override, or the fastest candidate
// measured on this batch, or the default when tuning is off.
int ensure_plan(v3dm_context* h, BatchState* b, int form)
{
    if (b->planValid[form]) return V3DM_OK;
    int dev = 0;
    (void)hipGetDevice(&dev);
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    auto make = [&](int tiles, int grid, int qRows) {
        return form == 0 ? plan_plain(b, tiles, grid, qRows)
                         : plan_short_batch(b, tiles, grid, form == 2 ? 8 : 16, form == 2, qRows);
    };
    // the learned lists' split size (at least kCodeSplitTiles, or the size set with
    // v3dm_batch_set_tuning) is not tuned: it sets which candidates are listed, so the
    // result would depend on tuning; the code pass's tile shape and grid are
    std::vector<int> tileCands = form == 0   ? std::vector<int>{8, 16, 32, 64}
                                 : form == 1 ? std::vector<int>{4, 8, 16, 32}
                                             : std::vector<int>{kCodeSplitTiles};
    std::vector<int> shapeCands = form == 0 ? std::vector<int>{128, 256} : std::vector<int>{256, 128};
    // persistent grids: the code pass runs 4 work-groups of 4 waves per CU (8 of 2 waves),
    // the BF16 pass 2 of 4 waves (4 of 2), the plain form 2 of 4 waves (1 of 8)
    auto gridCands = [&](int qRows) {
        int const perCU = form == 2 ? (qRows == 128 ? 4 : 2) : form == 1 ? (qRows == 128 ? 2 : 1) : 1;
        return std::vector<int>{perCU * cus, 2 * perCU * cus};
    };
    int bestT = form == 0 ? 16 : form == 1 ? 8 : kCodeSplitTiles, bestG = 2 * cus;
    int bestQ = form == 0 ? 128 : kBatchRowAlign;
    float bestMs = -1.0f;
    if (b->forceTiles[form] > 0 || b->forceGrid[form] > 0 || b->forceQRows[form] > 0) {
        bestT = b->forceTiles[form] > 0 ? b->forceTiles[form] : bestT;
        bestQ = b->forceQRows[form] > 0 ? b->forceQRows[form] : bestQ;
        bestG = b->forceGrid[form] > 0 ? b->forceGrid[form] : gridCands(bestQ)[0];
    } else if (b->autotune) {
        for (int q : shapeCands)
            for (int t : tileCands)
                for (int g : gridCands(q)) {
                    float ms = 0.0f;
                    int const rc = time_pass(h, b, form, make(t, g, q), ms);
                    if (rc) return rc;
                    if (bestMs < 0.0f || ms < bestMs) {
                        bestMs = ms;
                        bestT = t;
                        bestG = g;
                        bestQ = q;
                    }
                }
    }
    drop_graphs(b);
    Plan pl = make(bestT, bestG, bestQ);
    pl.passMs = bestMs;
    int const rc = upload_plan(h, b, form, pl);
    if (rc) return rc;
    b->plan[form] = std::move(pl);
    b->planValid[form] = true;
    b->version += 1;
    return V3DM_OK;
}

// The plain form for the pairs of a learned run that fail the sampled check runs inside the
// run (enqueue_run) on its own plan: items of kFallbackTiles 64-row tiles, so that even a
// single small pair spreads over the GPU; and room for the flagged items.
constexpr int kFallbackTiles = 2;

int ensure_fallback_plan(v3dm_context* h, BatchState* b)
{
    int rc;
    if (!b->planValid[3]) {
        int dev = 0, cus = 256;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        drop_graphs(b);
        Plan pl = plan_plain(b, kFallbackTiles, 2 * cus);
        if ((rc = upload_plan(h, b, 3, pl))) return rc;
        b->plan[3] = std::move(pl);
        b->planValid[3] = true;
        b->version += 1;
    }
    size_t const need = std::max<size_t>(b->plan[3].items.size(), 1);
    if ((size_t)b->fbItemCap < need) {
        drop_graphs(b);  // captured runs hold the old buffer
        if ((rc = growi(h, b->fbItems, b->fbItemCap, need))) return rc;
        b->version += 1;
    }
    return V3DM_OK;
}

// Statistics of the sampled check of the last learned run (host copy of the flags).
void fallback_stats(v3dm_context* h, BatchState* b)
{
    int const P = b->nPairs;
    int const* flag = b->hostChk;
    int const* cnt = b->hostChk + P;
    int nFall = 0, sampled = 0, missed = 0;
    for (int p = 0; p < P; ++p) {
        sampled += cnt[4 * p + 3];
        missed += cnt[4 * p + 1];
        nFall += flag[p] ? 1 : 0;
    }
    h->stats.nSampled = sampled;
    h->stats.nSampleMissed = missed;
    h->stats.nFallback = nFall;
}

int check_batch(v3dm_context* h, BatchState*& b)
{
    if (!h) return V3DM_ERR_INVALID;
    b = ctx_batch(h);
    if (!b || b->nImages <= 0) return ctx_fail(h, V3DM_ERR_STATE, "no batch images set");
    if (b->nPairs <= 0) return ctx_fail(h, V3DM_ERR_STATE, "no batch pairs set");
    return V3DM_OK;
}

int set_images(v3dm_context* h, int nImages, int32_t const* counts, void const* desc, bool u8)
{
    if (!h) return V3DM_ERR_INVALID;
    if (nImages <= 0 || !counts || !desc) return ctx_fail(h, V3DM_ERR_INVALID, "invalid batch images");
    BatchState*& b = ctx_batch(h);
    if (!b) {
        b = new BatchState();
        BCHECK(h, hipEventCreate(&b->evA));
        BCHECK(h, hipEventCreate(&b->evB));
    }
    size_t totalRaw = 0;
    std::vector<int> srcRowH(nImages), nH(nImages), baseH(nImages);
    int rows = 0;
    for (int m = 0; m < nImages; ++m) {
        if (counts[m] < 0) return ctx_fail(h, V3DM_ERR_INVALID, "negative feature count");
        srcRowH[m] = (int)totalRaw;
        totalRaw += (size_t)counts[m];
        nH[m] = counts[m] / 32 * 32;  // cu:493-502
        baseH[m] = rows;
        rows += roundUp(std::max(nH[m], 1), kBatchRowAlign);
    }
    if (totalRaw > (size_t)0x7fffffff / kDim) return ctx_fail(h, V3DM_ERR_INVALID, "batch too large");
    drop_graphs(b);
    b->planValid[0] = b->planValid[1] = b->planValid[2] = b->planValid[3] = false;
    b->nPairs = 0;
    b->pending = false;
    b->nImages = nImages;
    b->imgN = nH;
    b->imgBaseH = baseH;
    b->imgRows = rows;
    int rc;
    size_t imgCap = (size_t)b->imgCapRows;
    if (imgCap < (size_t)rows) {
        if (b->img) (void)hipFree(b->img);
        if (b->imgB) (void)hipFree(b->imgB);
        if (b->norm) (void)hipFree(b->norm);
        if (b->codes) (void)hipFree(b->codes);  // sized with the image rows: allocated again on use
        b->codes = nullptr;
        for (void* ptr : {(void*)b->imgU8, (void*)b->denom, (void*)b->rowCS})
            if (ptr) (void)hipFree(ptr);
        b->imgU8 = nullptr;
        b->denom = nullptr;
        b->rowCS = nullptr;
        b->img = nullptr;
        b->imgB = nullptr;
        b->norm = nullptr;
        BCHECK(h, hipMalloc(&b->img, (size_t)rows * kDim * sizeof(float)));
        BCHECK(h, hipMalloc(&b->imgB, (size_t)rows * kDim * sizeof(__bf16)));
        BCHECK(h, hipMalloc(&b->norm, (size_t)rows * sizeof(float)));
        b->imgCapRows = rows;
    }
    if ((rc = growi(h, b->imgBase, b->imgCap, nImages))) return rc;
    if (b->nImgD) (void)hipFree(b->nImgD);
    if (b->srcRow) (void)hipFree(b->srcRow);
    if (b->maxNorm) (void)hipFree(b->maxNorm);
    if (b->blkImg) (void)hipFree(b->blkImg);
    b->nImgD = b->srcRow = b->blkImg = nullptr;
    b->maxNorm = nullptr;
    BCHECK(h, hipMalloc(&b->nImgD, nImages * sizeof(int)));
    BCHECK(h, hipMalloc(&b->srcRow, nImages * sizeof(int)));
    BCHECK(h, hipMalloc(&b->maxNorm, nImages * sizeof(unsigned int)));
    int const nBlk = rows / kBatchRowAlign;
    std::vector<int> blkImgH(nBlk);
    for (int m = 0; m < nImages; ++m)
        for (int k = 0; k < roundUp(std::max(nH[m], 1), kBatchRowAlign) / kBatchRowAlign; ++k)
            blkImgH[baseH[m] / kBatchRowAlign + k] = m;
    BCHECK(h, hipMalloc(&b->blkImg, nBlk * sizeof(int)));

    size_t const elt = u8 ? 1 : sizeof(float);
    size_t const bytes = totalRaw * kDim * elt;
    if (b->pinnedCap < bytes) {
        if (b->pinned) (void)hipHostFree(b->pinned);
        b->pinned = nullptr;
        b->pinnedCap = 0;
        BCHECK(h, hipHostMalloc(&b->pinned, std::max<size_t>(bytes, 1), hipHostMallocDefault));
        b->pinnedCap = std::max<size_t>(bytes, 1);
    }
    if (b->uploadCap < bytes) {
        if (b->upload) (void)hipFree(b->upload);
        b->upload = nullptr;
        b->uploadCap = 0;
        BCHECK(h, hipMalloc(&b->upload, std::max<size_t>(bytes, 1)));
        b->uploadCap = std::max<size_t>(bytes, 1);
    }
    hipStream_t const st = h->stream;
    BCHECK(h, hipStreamSynchronize(st));  // the staging buffers may still feed an earlier upload
    std::memcpy(b->pinned, desc, bytes);
    BCHECK(h, hipMemcpyAsync(b->upload, b->pinned, bytes, hipMemcpyHostToDevice, st));
    BCHECK(h, hipMemcpyAsync(b->imgBase, baseH.data(), nImages * sizeof(int), hipMemcpyHostToDevice, st));
    BCHECK(h, hipMemcpyAsync(b->nImgD, nH.data(), nImages * sizeof(int), hipMemcpyHostToDevice, st));
    BCHECK(h, hipMemcpyAsync(b->srcRow, srcRowH.data(), nImages * sizeof(int), hipMemcpyHostToDevice, st));
    BCHECK(h, hipMemcpyAsync(b->blkImg, blkImgH.data(), nBlk * sizeof(int), hipMemcpyHostToDevice, st));
    BCHECK(h, hipMemsetAsync(b->maxNorm, 0, nImages * sizeof(unsigned int), st));
    if (u8 && !b->imgU8) {
        BCHECK(h, hipMalloc(&b->imgU8, (size_t)b->imgCapRows * kDim));
        BCHECK(h, hipMalloc(&b->denom, (size_t)b->imgCapRows * sizeof(float)));
        BCHECK(h, hipMalloc(&b->rowCS, (size_t)b->imgCapRows * sizeof(int2)));
    }
    b->haveU8 = u8;
    if (u8)
        launch_batch_load_u8(static_cast<uint8_t const*>(b->upload), b->srcRow, b->nImgD, b->blkImg, b->imgBase, rows,
                             b->img, b->imgU8, b->denom, b->rowCS, st);
    else
        launch_batch_load_f32(static_cast<float const*>(b->upload), b->srcRow, b->nImgD, b->blkImg, b->imgBase, rows,
                              b->img, st);
    launch_batch_norms(b->img, b->blkImg, rows, b->norm, b->maxNorm, st);
    b->bfValid = false;
    b->codesValid = false;
    BCHECK(h, hipGetLastError());
    BCHECK(h, hipStreamSynchronize(st));  // host vectors above go out of scope
    return V3DM_OK;
}

}  // namespace

void batch_destroy(BatchState*& b)
{
    if (!b) return;
    drop_graphs(b);
    for (void* p : {(void*)b->img, (void*)b->imgB, (void*)b->codes, (void*)b->imgU8, (void*)b->denom, (void*)b->imgBase, (void*)b->blkImg, (void*)b->nImgD,
                    (void*)b->srcRow, (void*)b->norm, (void*)b->maxNorm, b->upload, (void*)b->segOfBlock,
                    (void*)b->items[0], (void*)b->items[1], (void*)b->items[2], (void*)b->items[3],
                    (void*)b->segD[0], (void*)b->segD[1], (void*)b->segD[2], (void*)b->segD[3], (void*)b->blkInfo[0],
                    (void*)b->blkInfo[1], (void*)b->blkInfo[2], (void*)b->blkInfo[3], (void*)b->partF,
                    (void*)b->partB, (void*)b->sel, (void*)b->flagList, (void*)b->openList, (void*)b->sRows, (void*)b->sNorm, (void*)b->sIdx, (void*)b->sCount,
                    (void*)b->sNSplits, (void*)b->sPart, (void*)b->sItems, (void*)b->pairFlag, (void*)b->fbItems,
                    (void*)b->segMask, (void*)b->rowCS, (void*)b->cItems[0],
                    (void*)b->cItems[1], (void*)b->cNSplits, (void*)b->cPart, (void*)b->cQB, (void*)b->cSlotB,
                    (void*)b->cCntB, (void*)b->cDecA, (void*)b->counters, (void*)b->stripes, (void*)b->pairCount,
                    (void*)b->tmp, (void*)b->blockCount, (void*)b->out})
        if (p) (void)hipFree(p);
    if (b->pinned) (void)hipHostFree(b->pinned);
    if (b->hostRes) (void)hipHostFree(b->hostRes);
    if (b->hostChk) (void)hipHostFree(b->hostChk);
    if (b->evA) (void)hipEventDestroy(b->evA);
    if (b->evB) (void)hipEventDestroy(b->evB);
    delete b;
    b = nullptr;
}

}  // namespace v3dm

using namespace v3dm;

extern "C" {

int v3dm_batch_set_images_u8(v3dm_handle h, int nImages, const int32_t* counts, const uint8_t* desc)
{
    return set_images(h, nImages, counts, desc, true);
}

int v3dm_batch_set_images(v3dm_handle h, int nImages, const int32_t* counts, const float* desc)
{
    return set_images(h, nImages, counts, desc, false);
}

int v3dm_batch_set_pairs(v3dm_handle h, int nPairs, const int32_t* left, const int32_t* right)
{
    if (!h) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b || b->nImages <= 0) return ctx_fail(h, V3DM_ERR_STATE, "set the batch images first");
    if (nPairs <= 0 || !left || !right) return ctx_fail(h, V3DM_ERR_INVALID, "invalid pair list");
    for (int p = 0; p < nPairs; ++p)
        if (left[p] < 0 || left[p] >= b->nImages || right[p] < 0 || right[p] >= b->nImages)
            return ctx_fail(h, V3DM_ERR_INVALID, "pair refers to an image outside the batch");
    drop_graphs(b);
    b->planValid[0] = b->planValid[1] = b->planValid[2] = b->planValid[3] = false;
    b->pending = false;
    b->nPairs = nPairs;
    b->pa.assign(left, left + nPairs);
    b->pb.assign(right, right + nPairs);
    int const rc = ensure_query_space(h, b);
    if (rc) return rc;
    return ensure_sample_plan(h, b);
}

int v3dm_batch_set_tuning(v3dm_handle h, int autotune, int mode, int tilesPerItem, int grid)
{
    if (!h) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b) return ctx_fail(h, V3DM_ERR_STATE, "set the batch images first");
    if (mode < 0 || mode > V3DM_PREC_LEARNED || tilesPerItem < 0 || grid < 0)
        return ctx_fail(h, V3DM_ERR_INVALID, "invalid tuning");
    int const form = form_of(mode);
    b->autotune = autotune != 0;
    b->forceTiles[form] = tilesPerItem;
    b->forceGrid[form] = grid;
    b->planValid[form] = false;
    drop_graphs(b);
    return V3DM_OK;
}

int v3dm_batch_set_tile_shape(v3dm_handle h, int mode, int queryRows)
{
    if (!h) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b) return ctx_fail(h, V3DM_ERR_STATE, "set the batch images first");
    if (mode < 0 || mode > V3DM_PREC_LEARNED) return ctx_fail(h, V3DM_ERR_INVALID, "invalid mode");
    int const form = form_of(mode);
    // every form: 128 or 256 query rows per work-group (plain: left features)
    bool const ok = queryRows == 0 || queryRows == 128 || queryRows == 256;
    if (!ok) return ctx_fail(h, V3DM_ERR_INVALID, "tile shape not available for this mode");
    b->forceQRows[form] = queryRows;
    b->planValid[form] = false;
    drop_graphs(b);
    return V3DM_OK;
}

int v3dm_batch_get_tile_shape(v3dm_handle h, int mode, int* queryRows, int* candidateRows)
{
    if (!h) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b) return ctx_fail(h, V3DM_ERR_STATE, "no batch");
    int const form = form_of(mode);
    if (!b->planValid[form]) return ctx_fail(h, V3DM_ERR_STATE, "no plan for this mode yet");
    // plain: left block x 64-row right tiles; shortlists: query block x 64 candidates
    if (queryRows) *queryRows = b->plan[form].qRows;
    if (candidateRows) *candidateRows = 64;
    return V3DM_OK;
}

int v3dm_batch_get_tuning(v3dm_handle h, int mode, int* tilesPerItem, int* grid, int* nItems, float* passMs)
{
    if (!h) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b) return ctx_fail(h, V3DM_ERR_STATE, "no batch");
    int const form = form_of(mode);
    if (!b->planValid[form]) return ctx_fail(h, V3DM_ERR_STATE, "no plan for this mode yet");
    Plan const& pl = b->plan[form];
    if (tilesPerItem) *tilesPerItem = pl.tiles;
    if (grid) *grid = pl.grid;
    if (nItems) *nItems = (int)pl.items.size();
    if (passMs) *passMs = pl.passMs;
    return V3DM_OK;
}

int v3dm_batch_use_graph(v3dm_handle h, int enable)
{
    if (!h) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b) return ctx_fail(h, V3DM_ERR_STATE, "no batch");
    b->useGraph = enable != 0;
    return V3DM_OK;
}

int v3dm_batch_match_async(v3dm_handle h, float minScore, float minRatio)
{
    BatchState* b = nullptr;
    int rc = check_batch(h, b);
    if (rc) return rc;
    int const mode = h->precision;
    if (mode != V3DM_PREC_FP32 && mode != V3DM_PREC_BF16 && mode != V3DM_PREC_LEARNED)
        return ctx_fail(h, V3DM_ERR_UNSUPPORTED, "batched matching supports the FP32, BF16 and learned modes");
    int const form = form_of(mode);
    hipStream_t const st = h->stream;
    if (form == 2) {
        LearnedState* L = ctx_learned(h);
        if (!learned_weights(L) || !(learned_code_scale(L) > 0.0f))
            return ctx_fail(h, V3DM_ERR_STATE, "learned network or code scale not set");
        if (!b->codes) BCHECK(h, hipMalloc(&b->codes, (size_t)b->imgCapRows * kCode));
        if (b->codesScale != learned_code_scale(L)) b->codesValid = false;
        b->codesScale = learned_code_scale(L);
    }
    // derived copies outside the run when they are cached (else every run derives them)
    if (h->cacheDerived && form == 1 && !b->bfValid) {
        enqueue_derived(h, b, 1);
        b->bfValid = true;
    }
    if (h->cacheDerived && form == 2 && !b->codesValid) {
        enqueue_derived(h, b, 2);
        b->codesValid = true;
    }
    if (!h->cacheDerived && form != 0 && !b->planValid[form]) enqueue_derived(h, b, form);  // for the tuner
    if ((rc = ensure_plan(h, b, form))) return rc;
    if (form == 2 && h->fallbackSamples > 0 && (rc = ensure_fallback_plan(h, b))) return rc;
    BCHECK(h, hipEventRecord(b->evA, st));
    if (b->useGraph) {
        BatchState::GraphKey const key{mode, minScore, minRatio, h->epsRel, b->version * 2 + (h->cacheDerived ? 1 : 0),
                                       h->learnedTopK, form == 2 ? learned_margin(h) : 0.0f,
                                       form == 2 ? h->fallbackSamples : 0, form == 2 ? h->fallbackRecall : 0.0f};
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            hipGraph_t g = nullptr;
            BCHECK(h, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int const rc2 = enqueue_run(h, b, mode, minScore, minRatio);
            hipError_t const ec = hipStreamEndCapture(st, &g);
            if (rc2) {
                if (g) (void)hipGraphDestroy(g);
                return rc2;
            }
            if (ec != hipSuccess) return ctx_fail(h, V3DM_ERR_HIP, "hipStreamEndCapture failed");
            hipGraphExec_t ex = nullptr;
            hipError_t const ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return ctx_fail(h, V3DM_ERR_HIP, "hipGraphInstantiate failed");
            it = b->graphs.emplace(key, ex).first;
        }
        BCHECK(h, hipGraphLaunch(it->second, st));
    } else if ((rc = enqueue_run(h, b, mode, minScore, minRatio))) {
        return rc;
    }
    BCHECK(h, hipEventRecord(b->evB, st));
    BCHECK(h, hipMemcpyAsync(b->hostRes, b->counters, kBatchCounters * sizeof(int), hipMemcpyDeviceToHost, st));
    BCHECK(h, hipMemcpyAsync(b->hostRes + kBatchCounters, b->pairCount, b->nPairs * sizeof(int),
                             hipMemcpyDeviceToHost, st));
    b->checkPending = form == 2 && h->fallbackSamples > 0;
    if (b->checkPending)
        BCHECK(h, hipMemcpyAsync(b->hostChk, b->pairFlag, 5 * b->nPairs * sizeof(int), hipMemcpyDeviceToHost, st));
    b->lastMinScore = minScore;
    b->lastMinRatio = minRatio;
    b->pending = true;
    b->lastMode = mode;
    return V3DM_OK;
}

int v3dm_batch_get_matches(v3dm_handle h, int32_t* offsets, int32_t* pairs, int cap, int* nTotal)
{
    BatchState* b = nullptr;
    int rc = check_batch(h, b);
    if (rc) return rc;
    if (!b->pending) return ctx_fail(h, V3DM_ERR_STATE, "no batch run pending");
    if (!nTotal || cap < 0 || (cap > 0 && !pairs)) return ctx_fail(h, V3DM_ERR_INVALID, "invalid output");
    BCHECK(h, hipStreamSynchronize(h->stream));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->evA, b->evB) == hipSuccess) b->lastMs = ms;
    h->stats.nFallback = h->stats.nSampled = h->stats.nSampleMissed = 0;
    if (b->checkPending) {
        b->checkPending = false;
        fallback_stats(h, b);
    }
    int const total = b->hostRes[kBCntPairs];
    *nTotal = total;
    if (offsets) {
        int acc = 0;
        for (int p = 0; p < b->nPairs; ++p) {
            offsets[p] = acc;
            acc += b->hostRes[kBatchCounters + p];
        }
        offsets[b->nPairs] = acc;
    }
    h->stats.nMatches = total;
    h->stats.nRechecked = b->hostRes[kBCntFlagDone];
    h->stats.nRescored = b->hostRes[kBCntRescoredDone];
    h->stats.precision = b->lastMode;
    h->stats.nLeft = h->stats.nRight = 0;
    if (total > cap) return ctx_fail(h, V3DM_ERR_CAPACITY, "pair buffer too small");
    if (total > 0) BCHECK(h, hipMemcpy(pairs, b->out, (size_t)total * sizeof(int2), hipMemcpyDeviceToHost));
    return V3DM_OK;
}

int v3dm_batch_match(v3dm_handle h, float minScore, float minRatio, int32_t* offsets, int32_t* pairs, int cap,
                     int* nTotal)
{
    int const rc = v3dm_batch_match_async(h, minScore, minRatio);
    if (rc) return rc;
    return v3dm_batch_get_matches(h, offsets, pairs, cap, nTotal);
}

int v3dm_batch_get_fallback(v3dm_handle h, int32_t* flags)
{
    if (!h || !flags) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b || b->nPairs <= 0 || !b->hostChk) return ctx_fail(h, V3DM_ERR_STATE, "no batch");
    for (int p = 0; p < b->nPairs; ++p) flags[p] = b->hostChk[p];
    return V3DM_OK;
}

int v3dm_batch_time_dominant_kernel(v3dm_handle h, int reps, float* msPerLaunch)
{
    BatchState* b = nullptr;
    int rc = check_batch(h, b);
    if (rc) return rc;
    if (reps < 1 || !msPerLaunch) return ctx_fail(h, V3DM_ERR_INVALID, "bad argument");
    if (!b->pending) return ctx_fail(h, V3DM_ERR_STATE, "no batch run yet");
    int const form = form_of(b->lastMode);
    if (!b->planValid[form]) return ctx_fail(h, V3DM_ERR_STATE, "no plan for the last run's mode");
    Plan const& pl = b->plan[form];
    hipStream_t const st = h->stream;
    BatchDev const d = dev_view(b, form);
    int* q = nullptr;  // one zeroed work counter per launch
    BCHECK(h, hipMalloc(&q, (size_t)reps * sizeof(int)));
    hipError_t e = hipMemsetAsync(q, 0, (size_t)reps * sizeof(int), st);
    if (e == hipSuccess) e = hipEventRecord(b->evA, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        int const n = (int)pl.items.size();
        if (form == 0) launch_batch_fused_fp32(d, b->items[0], n, q + r, b->partF, pl.grid, pl.qRows, st);
        else if (form == 1) launch_batch_pass_bf16(d, b->items[1], n, q + r, b->partB, pl.grid, pl.qRows, st);
        else launch_batch_pass_code(d, b->items[2], n, q + r, b->partB, pl.grid, pl.qRows, st);
    }
    if (e == hipSuccess) e = hipEventRecord(b->evB, st);
    if (e == hipSuccess) e = hipEventSynchronize(b->evB);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->evA, b->evB);
    (void)hipFree(q);
    if (e != hipSuccess) return ctx_fail(h, V3DM_ERR_HIP, "timing the batch pass failed");
    *msPerLaunch = ms / reps;
    return V3DM_OK;
}

int v3dm_batch_get_check(v3dm_handle h, int32_t* counts)
{
    if (!h || !counts) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b || b->nPairs <= 0 || !b->hostChk) return ctx_fail(h, V3DM_ERR_STATE, "no batch");
    for (int k = 0; k < 4 * b->nPairs; ++k) counts[k] = b->hostChk[b->nPairs + k];
    return V3DM_OK;
}

int v3dm_batch_last_ms(v3dm_handle h, float* ms)
{
    if (!h || !ms) return V3DM_ERR_INVALID;
    BatchState* b = ctx_batch(h);
    if (!b) return ctx_fail(h, V3DM_ERR_STATE, "no batch");
    *ms = b->lastMs;
    return V3DM_OK;
}

}  // extern "C"
