// Sampled exact check of a learned batch run over uint8 images: the sampled recall of the
// pair's exact match set.
//
// Per pair p, side A is the image with more features (left on a tie) and B the other.
// Phase 0 decides S = min(samples, nA) evenly spaced features i_k of A exactly against B
// (the per-feature rule of V3D/CUDA/v3d_cudamatching.cu:133-210); phase 1 decides, against
// A, every feature j_k of B that phase 0 matched.  Sample k is an exact match when phase 1
// returns i_k (the mutual check of cu:546-551 restricted to the samples), and the learned
// result holds it when its decisions of i_k and j_k are j_k and i_k.  The pair falls back
// to the plain form when its misses exceed the limit for its number of exact matches.
//
// Scores: the uint8 rows are taken as int8 (u ^ 0x80 = u - 128), so one v_mfma_i32_32x32x32_i8
// chain over 128 components gives D = sum_k (qa_k - 128)(qb_k - 128), and with the MFMA's
// accumulator input set to 128 * sum(qb) + 128 * sum(qa) - 2^21 its result is the exact
// integer I = sum_k qa_k qb_k.  t = fl32(rb * I) and s~ = fl32(ra * t) (ra, rb the rows' 1/sqrt
// scales) differ from the exact score fl32(sum_k fl32(qa_k ra) fl32(qb_k rb)) by at most
// 5.01 * 2^-24 * s~ (every term is nonnegative); the candidates' keys keep t with its five
// low mantissa bits replaced by the row inside the 32-row block (2^-18 relative), so a
// decision from the keys is certified with eps = 2^-17 * s~1.  A decision the bound leaves
// open is taken exactly: the candidates whose score could be in the exact top-2 (s~ >=
// s~2 - 2 eps) are found again from the integer scores and scored with the oracle's
// arithmetic (FP64 sum of the FP32 products, ascending k).
#include <climits>

#include "v3dm_batch.h"
#include "v3dm_device.h"

namespace v3dm {

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kChkTile = 32;       // candidate rows per MFMA block
constexpr int kChkCand = 256;      // candidates an open decision scores exactly (more: all of them)
constexpr float kCertChk = 7.62939453125e-06f;  // 2^-17
constexpr float kKeySlack = 1.0000076293945312f;  // 1 + 2^-17

__device__ __forceinline__ int chk_nq(BatchDev const& b, CheckDev const& c, int phase, int p, BatchSeg const& sA)
{
    return phase == 0 ? min(c.samples, sA.n) : c.cntB[p];
}

// image-local row of query slot k (phase 0: the k-th evenly spaced feature of A)
__device__ __forceinline__ int chk_query(CheckDev const& c, int phase, int p, int k, int nQ, int nA)
{
    return phase == 0 ? (int)(((long long)k * nA) / nQ) : c.qB[p * kSampleRows + k];
}

__device__ __forceinline__ float key_value(int key) { return __int_as_float(key & ~31); }

// The second of the two largest after `key` joins (k1, k2), k1 >= k2: max(k2, min(k1, key)),
// which is the median of the three, in one v_med3_i32.
__device__ __forceinline__ int top2_second(int k1, int k2, int key)
{
    int r;
    __asm__("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(k1), "v"(k2), "v"(key));
    return r;
}

// Waves of a probe work-group: query block (32 rows) w % qbp of the ceil(nQ / 32) blocks,
// tile part w / qbp of the item's tiles (qbp = the block count rounded up to 1, 2 or 4), so
// that a phase with few queries still keeps the four waves busy; parts = 4 / qbp.
__device__ __forceinline__ int probe_qbp(int nQ)
{
    int const nqb = (nQ + 31) / 32;
    return nqb <= 1 ? 1 : (nqb <= 2 ? 2 : 4);
}

// exact score of image rows q and c (the oracle's arithmetic on the dequantised floats)
__device__ __forceinline__ float exact_score(float const* __restrict__ q, float const* __restrict__ c)
{
    float4 const* qp = reinterpret_cast<float4 const*>(q);
    float4 const* cp = reinterpret_cast<float4 const*>(c);
    double acc = 0.0;
#pragma unroll 8
    for (int u = 0; u < 32; ++u) {
        float4 const x = qp[u], y = cp[u];
        acc += (double)x.x * (double)y.x;
        acc += (double)x.y * (double)y.y;
        acc += (double)x.z * (double)y.z;
        acc += (double)x.w * (double)y.w;
    }
    return (float)acc;
}

// Top2 of a work-group of 128 threads (two waves), the order-independent merge
__device__ __forceinline__ Top2 wg_top2(Top2 s, Top2* red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s = top2_merge(s, top2_shfl_xor(s, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    return top2_merge(red[0], red[1]);
}

}  // namespace

// ---------------------------------------------------------------------------- probe
// Each work-group takes the items blockIdx.x, blockIdx.x + gridDim.x, ... (largest first; two
// int4 each, every row offset resolved by the host: pair, split, first tile, tiles; row of
// the first tile, first row of the query image, nA): the pair's (up to 128) query rows of
// this phase against the item's 32-row candidate tiles.  Wave w takes query block w % qbp
// (qbp = the block count rounded up to 1, 2 or 4) and, of every group of four consecutive
// tiles, tiles w / qbp, w / qbp + parts, ... (parts = 4 / qbp: with few queries the waves
// split the tiles instead).  The work-group stages each group through LDS (rows padded to
// 144 bytes, so the lanes' 16-byte reads of 32 rows spread over the banks), loading the
// next group into registers while the current one is multiplied.  An item's loads depend
// only on its record, so they start together: the first group's tiles, and the query rows
// (phase 1: after the pair's list of matched features).  Per lane (query l % 32, candidate
// rows 8 (r / 4) + 4 (l / 32) + r % 4 of each block) the two best keys and the tile of the
// best; the lane halves merge by a shuffle, the waves of a query block through LDS (equal
// keys keep the smaller tile).
// part: int4 (key1, key2, tile of key1, 0) per (phase, pair, split, query slot).
__device__ __forceinline__ void key_merge(int& k1, int& k2, int& b1, int o1, int o2, int ob)
{
    bool const take = o1 > k1 || (o1 == k1 && ob >= 0 && (b1 < 0 || ob < b1));
    k2 = max(min(k1, o1), max(k2, o2));
    b1 = take ? ob : b1;
    k1 = max(k1, o1);
}

constexpr int kRowPadB = 144;                 // LDS bytes per staged 128-byte row
constexpr int kTileB = kChkTile * kRowPadB;   // 4608
constexpr int kGroupTiles = 4;                // tiles per stage
constexpr int kProbeGroupsPerCU = 3;

// group `grp` of the item's tiles [t0, t1) -> registers: thread tid holds chunk tid % 8 of
// row tid / 8 of each of the group's tiles, and rowCS of row tid % 32 when tid < 32
struct StageRegs {
    uint4 v[kGroupTiles];
    int2 cs[kGroupTiles];
};

__device__ __forceinline__ void stage_load(CheckDev const& c, size_t cBase, int t0, int t1, int grp, StageRegs& r)
{
    int const tid = threadIdx.x;
#pragma unroll
    for (int g = 0; g < kGroupTiles; ++g) {
        int const t = t0 + grp * kGroupTiles + g;
        if (t < t1) {
            size_t const row = cBase + (size_t)t * kChkTile;
            r.v[g] = *reinterpret_cast<uint4 const*>(c.u8 + (row + (tid >> 3)) * kDim + 16 * (tid & 7));
            if (tid < kChkTile) r.cs[g] = c.rowCS[row + tid];
        }
    }
}

__device__ __forceinline__ void stage_store(unsigned char* st, int2* stCS, int t0, int t1, int grp,
                                            StageRegs const& r)
{
    int const tid = threadIdx.x;
#pragma unroll
    for (int g = 0; g < kGroupTiles; ++g) {
        if (t0 + grp * kGroupTiles + g < t1) {
            uint4 const v = r.v[g];  // uint8 -> int8 u - 128
            *reinterpret_cast<uint4*>(st + g * kTileB + (tid >> 3) * kRowPadB + 16 * (tid & 7)) =
                make_uint4(v.x ^ 0x80808080u, v.y ^ 0x80808080u, v.z ^ 0x80808080u, v.w ^ 0x80808080u);
            if (tid < kChkTile) stCS[g * kChkTile + tid] = r.cs[g];
        }
    }
}

// One staged 32-row tile against this wave's query block: keys of the 32 x 32 block into
// (k1, k2) and tile t as the tile of k1 when it holds it
__device__ __forceinline__ void probe_tile(unsigned char const* tp, int2 const* cs, v4i const (&bq)[4], int cq, int t,
                                           int h, int c32, int& k1, int& k2, int& blk1)
{
    v4i a[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) a[m] = *reinterpret_cast<v4i const*>(tp + c32 * kRowPadB + 32 * m + 16 * h);
    v16i cv;
    float sv[16];
#pragma unroll
    for (int g = 0; g < 4; ++g) {  // accumulator input: 128 sum(qb) + 128 sum(qa) - 2^21
        int4 const x0 = *reinterpret_cast<int4 const*>(cs + 8 * g + 4 * h);
        int4 const x1 = *reinterpret_cast<int4 const*>(cs + 8 * g + 4 * h + 2);
        cv[4 * g + 0] = x0.x + cq;
        sv[4 * g + 0] = __int_as_float(x0.y);
        cv[4 * g + 1] = x0.z + cq;
        sv[4 * g + 1] = __int_as_float(x0.w);
        cv[4 * g + 2] = x1.x + cq;
        sv[4 * g + 2] = __int_as_float(x1.y);
        cv[4 * g + 3] = x1.z + cq;
        sv[4 * g + 3] = __int_as_float(x1.w);
    }
    v16i d = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[0], bq[0], cv, 0, 0, 0);
    d = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[1], bq[1], d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[2], bq[2], d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[3], bq[3], d, 0, 0, 0);
    int const o = k1;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        float const tv = sv[q] * (float)d[q];  // degenerate rows: negative
        int const key = (__float_as_int(tv) & ~31) | (8 * (q >> 2) + 4 * h + (q & 3));
        k2 = top2_second(k1, k2, key);
        k1 = max(k1, key);
    }
    blk1 = (k1 != o) ? t : blk1;
}

__global__ __launch_bounds__(256) void k_batch_check_probe(BatchDev b, CheckDev c, int phase,
                                                           int4 const* __restrict__ items, int nItems)
{
    __shared__ int4 red[4][32];
    __shared__ __attribute__((aligned(16))) unsigned char st[2][kGroupTiles * kTileB];
    __shared__ __attribute__((aligned(16))) int2 stCS[2][kGroupTiles * kChkTile];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    for (int it = blockIdx.x; it < nItems; it += gridDim.x) {
        int4 const w = items[2 * it];       // (pair, split, first tile, tiles)
        int4 const wr = items[2 * it + 1];  // (row of the first tile, first row of the query image, nA, 0)
        int const p = w.x;
        int const t0 = w.z, t1 = w.z + w.w;
        size_t const cBase = (size_t)wr.x - (size_t)kChkTile * t0;  // tile t starts at row cBase + 32 t
        int const nGroups = (t1 - t0 + kGroupTiles - 1) / kGroupTiles;
        StageRegs r;
        stage_load(c, cBase, t0, t1, 0, r);
        // phase 1: the pair's matched features of B, every slot a wave may take (c32 + 32 j)
        int qs[4] = {0, 0, 0, 0};
        int nQ;
        if (phase == 0) {
            nQ = min(c.samples, wr.z);
        } else {
            nQ = c.cntB[p];
#pragma unroll
            for (int j = 0; j < 4; ++j) qs[j] = c.qB[p * kSampleRows + 32 * j + c32];
        }
        int const qbp = probe_qbp(nQ), parts = 4 / qbp;
        int const qb = wave % qbp, part = wave / qbp;
        bool const active = qb * 32 < nQ;  // uniform over the wave
        int const k = qb * 32 + c32;
        // image-local row of query slot kk (phase 0: the kk-th evenly spaced feature of A),
        // formed by every wave so that the slot loads above are always consumed here
        int const kk = max(min(k, nQ - 1), 0);
        int qLocal = 0;
        if (phase == 0)
            qLocal = nQ > 0 ? (int)(((long long)kk * wr.z) / nQ) : 0;
        else
            qLocal = kk == k ? (qb == 0 ? qs[0] : qb == 1 ? qs[1] : qb == 2 ? qs[2] : qs[3])
                             : c.qB[p * kSampleRows + kk];
        v4i bq[4];
        int cq = 0;
        if (active) {
            size_t const qRow = (size_t)wr.y + qLocal;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                v4i const u = *reinterpret_cast<v4i const*>(c.u8 + qRow * kDim + 32 * m + 16 * h);
                bq[m] = u ^ (int)0x80808080;
            }
            cq = c.rowCS[qRow].x - (1 << 21);
        }
        stage_store(st[0], stCS[0], t0, t1, 0, r);
        __syncthreads();
        int k1 = INT_MIN, k2 = INT_MIN, blk1 = -1;
        for (int gi = 0; gi < nGroups; ++gi) {
            int const buf = gi & 1;
            if (gi + 1 < nGroups) stage_load(c, cBase, t0, t1, gi + 1, r);
            if (active)
                for (int j = part; j < kGroupTiles; j += parts) {
                    int const t = t0 + gi * kGroupTiles + j;
                    if (t < t1) probe_tile(st[buf] + j * kTileB, stCS[buf] + j * kChkTile, bq, cq, t, h, c32, k1, k2, blk1);
                }
            if (gi + 1 < nGroups) stage_store(st[buf ^ 1], stCS[buf ^ 1], t0, t1, gi + 1, r);
            __syncthreads();
        }
        if (active) key_merge(k1, k2, blk1, __shfl_xor(k1, 32), __shfl_xor(k2, 32), __shfl_xor(blk1, 32));
        if (parts > 1) {  // uniform over the work-group
            if (active && h == 0) red[wave][c32] = make_int4(k1, k2, blk1, 0);
            __syncthreads();
            if (active && part == 0 && h == 0)
                for (int q = 1; q < parts; ++q) {
                    int4 const e = red[qb + q * qbp][c32];
                    key_merge(k1, k2, blk1, e.x, e.y, e.z);
                }
        }
        if (active && part == 0 && h == 0 && k < nQ)
            c.part[(((size_t)phase * b.nPairs + p) * c.maxSplits + w.y) * kSampleRows + k] =
                make_int4(k1, k2, blk1, 0);
        __syncthreads();  // the stages and the LDS merge are done
    }
}

// ---------------------------------------------------------------------------- decide
// One work-group (128 threads, thread = query slot) per pair: merge the splits' keys, the
// certified decision, the exact decision of open queries (the whole work-group per open
// query), then phase 0: the decisions and the list of matched B features for phase 1;
// phase 1: the comparison with the learned decisions and the pair's fallback flag.
__global__ __launch_bounds__(128) void k_batch_check_decide(BatchDev b, CheckDev c, int phase, float minScore,
                                                            float minRatio, MissLimits lim,
                                                            int const* __restrict__ sel,
                                                            int* __restrict__ pairFlag, int* __restrict__ counts)
{
    __shared__ int openSlot[kSampleRows];
    __shared__ float openThr[kSampleRows];
    __shared__ int openDec[kSampleRows];
    __shared__ int cand[kChkCand];
    __shared__ int scanSplit[kChkMaxSplits];
    __shared__ int nOpen, nCand, nScan;
    __shared__ Top2 red[2];
    __shared__ int wcnt[2][3];
    int const p = blockIdx.x;
    int const k = threadIdx.x;
    int const sa = sample_side(b, p);
    BatchSeg const sA = b.seg[sa * b.nPairs + p];
    BatchSeg const sB = b.seg[(1 - sa) * b.nPairs + p];
    int const nQ = (sA.n > 0 && sB.n > 0) ? chk_nq(b, c, phase, p, sA) : 0;
    int const qImg = phase == 0 ? sA.img : sB.img;
    int const cImg = phase == 0 ? sB.img : sA.img;
    int const nC = phase == 0 ? sB.n : sA.n;
    size_t const cBase = (size_t)b.imgBase[cImg];
    if (k == 0) nOpen = 0;
    __syncthreads();
    int d = -1;
    int qLocal = -1;
    if (k < nQ) {
        qLocal = chk_query(c, phase, p, k, nQ, sA.n);
        float const ra = b.denom[(size_t)b.imgBase[qImg] + qLocal];
        if (ra > 0.0f && ra < INFINITY) {  // else an all-zero row: every exact score is NaN
            int K1 = INT_MIN, K2 = INT_MIN, B1 = -1;
            int const ns = c.nSplits[phase * b.nPairs + p];
            int4 const* pp = c.part + ((size_t)phase * b.nPairs + p) * c.maxSplits * kSampleRows + k;
#pragma unroll 8
            for (int s = 0; s < ns; ++s) {
                int4 const e = pp[(size_t)s * kSampleRows];
                bool const take = e.x > K1;  // splits in ascending order: equal keys keep the first
                K2 = max(max(K2, e.y), min(K1, e.x));
                B1 = take ? e.z : B1;
                K1 = max(K1, e.x);
            }
            float const s1 = K1 >= 0 ? ra * key_value(K1) : kInit;
            float const s2 = K2 >= 0 ? ra * key_value(K2) : kInit;
            int const a1 = K1 >= 0 ? B1 * kChkTile + (K1 & 31) : -1;
            float const eps = kCertChk * fmaxf(s1, 0.0f) + kEpsAbs;
            d = decide_certified(Top2{s1, s2, a1}, eps, minScore, minRatio);
            if (d == kFlag) {
                int const o = atomicAdd(&nOpen, 1);
                openSlot[o] = k;
                openThr[o] = s2 - 2.0f * eps;
            }
        }
    }
    __syncthreads();
    // exact decisions of the open queries.  The exact top-2 lies among the candidates with
    // s~ >= thr = s~2 - 2 eps.  A split whose second key stays below thr holds at most one
    // of them, its best (listed from the key); a split whose second key reaches thr may
    // hold more and is scanned again with the integer scores.
    int const nO = nOpen;
    int const ns = c.nSplits[phase * b.nPairs + p];
    int const span = c.tilesPerSplit[phase] * kChkTile;  // candidates per split
    for (int o = 0; o < nO; ++o) {
        int const kq = openSlot[o];
        int const ql = chk_query(c, phase, p, kq, nQ, sA.n);
        size_t const qRow = (size_t)b.imgBase[qImg] + ql;
        float const ra = b.denom[qRow];
        float const thr = openThr[o];
        if (k == 0) nCand = nScan = 0;
        __syncthreads();
        int4 const* pp = c.part + ((size_t)phase * b.nPairs + p) * c.maxSplits * kSampleRows + kq;
        for (int sp = k; sp < ns; sp += blockDim.x) {
            int4 const e = pp[(size_t)sp * kSampleRows];
            // keys truncate t by less than 2^-18 relative: the slack keeps every candidate
            // whose untruncated s~ may reach thr
            bool const two = e.y >= 0 && ra * key_value(e.y) * kKeySlack >= thr;
            bool const one = e.x >= 0 && ra * key_value(e.x) * kKeySlack >= thr;
            if (two) {
                int const q = atomicAdd(&nScan, 1);
                if (q < kChkMaxSplits) scanSplit[q] = sp;
            } else if (one) {
                int const q = atomicAdd(&nCand, 1);
                if (q < kChkCand) cand[q] = e.z * kChkTile + (e.x & 31);
            }
        }
        __syncthreads();
        bool const overflow = nScan > kChkMaxSplits;  // every candidate exactly below
        int const nsc = overflow ? 0 : nScan;
        uint4 qv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) qv[u] = reinterpret_cast<uint4 const*>(b.u8 + qRow * kDim)[u];
        for (int e = 0; e < nsc; ++e) {
            int const j0 = scanSplit[e] * span;
            int const j1 = min(j0 + span, nC);
            for (int j = j0 + k; j < j1; j += blockDim.x) {
                size_t const row = cBase + j;
                int2 const csr = c.rowCS[row];
                if (csr.x == kDegenerateC) continue;
                uint4 const* cr = reinterpret_cast<uint4 const*>(b.u8 + row * kDim);
                unsigned int acc = 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    uint4 const y = cr[u];
                    acc = __builtin_amdgcn_udot4(qv[u].x, y.x, acc, false);
                    acc = __builtin_amdgcn_udot4(qv[u].y, y.y, acc, false);
                    acc = __builtin_amdgcn_udot4(qv[u].z, y.z, acc, false);
                    acc = __builtin_amdgcn_udot4(qv[u].w, y.w, acc, false);
                }
                float const sv = ra * (__int_as_float(csr.y) * (float)acc);
                if (sv * kKeySlack >= thr) {
                    int const q = atomicAdd(&nCand, 1);
                    if (q < kChkCand) cand[q] = j;
                }
            }
        }
        __syncthreads();
        int const nc = nCand;
        float const* qf = b.img + qRow * kDim;
        Top2 s = top2_init();
        if (nc <= kChkCand && !overflow) {
            for (int e = k; e < nc; e += blockDim.x) {
                int const j = cand[e];
                float const sc = exact_score(qf, b.img + (cBase + j) * kDim);
                if (sc == sc) s = top2_merge(s, Top2{sc, kInit, j});
            }
        } else {  // a pile of near ties: every candidate exactly
            for (int j = k; j < nC; j += blockDim.x) {
                float const sc = exact_score(qf, b.img + (cBase + j) * kDim);
                if (sc == sc) s = top2_merge(s, Top2{sc, kInit, j});
            }
        }
        Top2 const t = wg_top2(s, red);
        if (k == 0) openDec[o] = decide_exact(t, minScore, minRatio);
        __syncthreads();
    }
    if (d == kFlag) {
        for (int o = 0; o < nO; ++o)
            if (openSlot[o] == k) d = openDec[o];
    }
    int const lane = k & 63, wv = k >> 6;
    if (phase == 0) {
        // decisions, and the matched B features in slot order for phase 1
        if (k < kSampleRows) c.decA[p * kSampleRows + k] = (k < nQ) ? d : -1;
        bool const m = k < nQ && d >= 0;
        unsigned long long const bal = __ballot(m);
        int const below = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wv][0] = __popcll(bal);
        __syncthreads();
        int const pos = (wv == 1 ? wcnt[0][0] : 0) + below;
        if (m) {
            c.qB[p * kSampleRows + pos] = d;
            c.slotB[p * kSampleRows + pos] = k;
        }
        if (k == 0) {
            c.cntB[p] = wcnt[0][0] + wcnt[1][0];
            counts[4 * p + 2] = nO;
        }
    } else {
        int acc = 0, miss = 0;
        if (k < nQ) {
            int const slot = c.slotB[p * kSampleRows + k];
            int const S = min(c.samples, sA.n);
            int const i = (int)(((long long)slot * sA.n) / S);
            int const j = qLocal;
            if (d == i) {
                acc = 1;
                bool const held = sel[sA.qOff + i] == j && sel[sB.qOff + j] == i;
                miss = held ? 0 : 1;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            acc += __shfl_xor(acc, o);
            miss += __shfl_xor(miss, o);
        }
        if (lane == 0) {
            wcnt[wv][0] = acc;
            wcnt[wv][1] = miss;
        }
        __syncthreads();
        if (k == 0) {
            int const a = wcnt[0][0] + wcnt[1][0], mm = wcnt[0][1] + wcnt[1][1];
            int const S = (sA.n > 0 && sB.n > 0) ? min(c.samples, sA.n) : 0;
            counts[4 * p + 0] = a;
            counts[4 * p + 1] = mm;
            counts[4 * p + 2] += nO;
            counts[4 * p + 3] = S;
            pairFlag[p] = (mm > (int)lim.allowed[a]) ? 1 : 0;
        }
    }
}

void launch_batch_check_probe(BatchDev const& b, CheckDev const& c, int phase, int4 const* items, int nItems,
                              int cus, hipStream_t st)
{
    if (nItems <= 0) return;
    // kProbeGroupsPerCU work-groups per CU (40 KiB of LDS and 128 VGPRs each: four would
    // fit; three measured as fast or faster); items are taken in a fixed stride, largest first
    hipLaunchKernelGGL(k_batch_check_probe, dim3(kProbeGroupsPerCU * cus), dim3(256), 0, st, b, c, phase, items,
                       nItems);
}

void launch_batch_check_decide(BatchDev const& b, CheckDev const& c, int phase, float minScore, float minRatio,
                               MissLimits const& lim, int const* sel, int* pairFlag, int* counts, hipStream_t st)
{
    if (b.nPairs <= 0) return;
    hipLaunchKernelGGL(k_batch_check_decide, dim3(b.nPairs), dim3(kSampleRows), 0, st, b, c, phase, minScore,
                       minRatio, lim, sel, pairFlag, counts);
}

}  // namespace v3dm
