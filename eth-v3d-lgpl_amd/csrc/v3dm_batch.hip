// Selection and bookkeeping kernels of batched matching: per-row norms, merge of the FP32
// partial top-2 with certified decisions, exact FP64 re-check of open queries, and the
// mutual check with stable compaction of every pair's matches.
//
// Per pair these compute what V3D_CUDA::SIFT_Matching::findPutativeMatches does
// (V3D/CUDA/v3d_cudamatching.cu:522-552) for that pair alone, as the reference's callers
// do one pair at a time (V3D/Apps/SfM/match_sequence.cpp:165-382).
#include "v3dm_batch.h"
#include "v3dm_kernels.h"

namespace v3dm {

__global__ void k_batch_reset(int* __restrict__ counters, int* __restrict__ pairCount, int nPairs)
{
    int const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kBatchCounters) {
        // totals of the previous run stay readable until the gather overwrites them
        if (t != kBCntPairs && t != kBCntFlagDone && t != kBCntRescoredDone) counters[t] = 0;
    }
    if (t < nPairs) pairCount[t] = 0;
}

__global__ __launch_bounds__(256) void k_batch_fb_filter(int4 const* __restrict__ items, int nItems,
                                                         int const* __restrict__ pairFlag, int nPairs,
                                                         int4* __restrict__ fbItems, int* __restrict__ fbCount,
                                                         int* __restrict__ segMask)
{
    int const t = blockIdx.x * 256 + threadIdx.x;
    if (t < 2 * nPairs) segMask[t] = pairFlag[t < nPairs ? t : t - nPairs] != 0;
    if (t >= nItems) return;
    int4 const w = items[t];
    if (pairFlag[w.x]) fbItems[atomicAdd(fbCount, 1)] = w;
}

void launch_batch_fb_filter(int4 const* items, int nItems, int const* pairFlag, int nPairs, int4* fbItems,
                            int* fbCount, int* segMask, hipStream_t st)
{
    int const n = nItems > 2 * nPairs ? nItems : 2 * nPairs;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_batch_fb_filter, dim3((n + 255) / 256), dim3(256), 0, st, items, nItems, pairFlag, nPairs,
                       fbItems, fbCount, segMask);
}

void launch_batch_reset(int* counters, int* pairCount, int nPairs, hipStream_t st)
{
    int const n = nPairs > kBatchCounters ? nPairs : kBatchCounters;
    hipLaunchKernelGGL(k_batch_reset, dim3((n + 255) / 256), dim3(256), 0, st, counters, pairCount, nPairs);
}

// Row norms of the image buffer and per-image maximum.  blkImg[row / kBatchRowAlign] =
// image of the row; a work-group's 64 rows lie in one image (images are aligned to 256
// rows), so one atomicMax per work-group suffices.
__global__ __launch_bounds__(256) void k_batch_norms(float const* __restrict__ X, int const* __restrict__ blkImg,
                                                     int rows, float* __restrict__ norm,
                                                     unsigned int* __restrict__ maxNorm)
{
    __shared__ unsigned int wmax[4];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    unsigned int mx = 0;
    for (int k = 0; k < 16; ++k) {
        int const row = blockIdx.x * 64 + wave * 16 + k;
        if (row >= rows) break;
        float2 const v = reinterpret_cast<float2 const*>(X + (size_t)row * kDim)[lane];
        float s = v.x * v.x + v.y * v.y;
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        float const nr = sqrtf(s);
        if (lane == 0) norm[row] = nr;
        mx = max(mx, __float_as_uint(nr));
    }
    if (lane == 0) wmax[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0 && blockIdx.x * 64 < rows) {
        unsigned int const m = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        atomicMax(&maxNorm[blkImg[(blockIdx.x * 64) / kBatchRowAlign]], m);
    }
}

void launch_batch_norms(float const* img, int const* blkImg, int rows, float* norm, unsigned int* maxNorm,
                        hipStream_t st)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_batch_norms, dim3((rows + 63) / 64), dim3(256), 0, st, img, blkImg, rows, norm, maxNorm);
}

// 8 lanes (sub = lane & 7) per query row g of segment si: merge the segment's partials,
// then the certified decision (as the single-pair merge).  Open queries are appended to
// flagList (query-space rows).
__device__ __forceinline__ void merge_decide_row(BatchDev const& b, Top2 const* __restrict__ part, int si, int g,
                                                 int sub, float epsRel, float minScore, float minRatio,
                                                 int* __restrict__ sel, int* __restrict__ flagList,
                                                 int* __restrict__ counters)
{
    BatchSeg const sg = b.seg[si];
    int const q = g - sg.qOff;
    int const np = (q < sg.n) ? sg.nParts : 0;
    Top2 const* pp = part + sg.partOff + q;
    Top2 s = top2_init();
#pragma unroll 4
    for (int p = sub; p < np; p += 8) s = top2_merge(s, pp[(size_t)p * sg.ld]);
    s = top2_merge(s, top2_shfl_xor(s, 1));
    s = top2_merge(s, top2_shfl_xor(s, 2));
    s = top2_merge(s, top2_shfl_xor(s, 4));
    if (sub != 0) return;  // the group's result is in every lane; its first lane writes it
    if (q >= sg.n || sg.nO <= 0) {
        sel[g] = -1;
        return;
    }
    float const nq = b.norm[b.imgBase[sg.img] + q];
    float const mo = __uint_as_float(b.maxNorm[sg.oImg]);
    float const eps = epsRel * nq * mo + kEpsAbs;
    int const d = (epsRel > 0.0f) ? decide_certified(s, eps, minScore, minRatio) : kFlag;
    if (d == kFlag) flagList[atomicAdd(&counters[kBCntFlag], 1)] = g;
    sel[g] = d;
}

__global__ __launch_bounds__(256) void k_batch_merge_decide(BatchDev b, Top2 const* __restrict__ part, float epsRel,
                                                            float minScore, float minRatio, int* __restrict__ sel,
                                                            int* __restrict__ flagList, int* __restrict__ counters)
{
    int const g = (int)((blockIdx.x * 256u + threadIdx.x) >> 3);
    if (g >= b.qRows) return;  // uniform over the 8 lanes of a query
    merge_decide_row(b, part, b.segOfBlock[g / kBatchRowAlign], g, threadIdx.x & 7, epsRel, minScore, minRatio, sel,
                     flagList, counters);
}

// The same for the segments with segMask set only: a persistent grid reads the segment
// masks 1024 at a time into LDS (a list of the flagged ones) and shares each flagged
// segment's rows (32 per work-group step) out, so that a run with few or no flagged pairs
// costs a few microseconds.
__global__ __launch_bounds__(256) void k_batch_merge_masked(BatchDev b, Top2 const* __restrict__ part, float epsRel,
                                                            float minScore, float minRatio, int* __restrict__ sel,
                                                            int* __restrict__ flagList, int* __restrict__ counters,
                                                            int const* __restrict__ segMask)
{
    __shared__ int list[1024];
    __shared__ int nList;
    int const sub = threadIdx.x & 7;
    int const r = threadIdx.x >> 3;
    int const nSeg = 2 * b.nPairs;
    for (int base = 0; base < nSeg; base += 1024) {
        __syncthreads();  // the previous chunk's list is consumed
        if (threadIdx.x == 0) nList = 0;
        __syncthreads();
        for (int u = threadIdx.x; u < 1024 && base + u < nSeg; u += 256)
            if (segMask[base + u]) list[atomicAdd(&nList, 1)] = base + u;
        __syncthreads();
        int const n = nList;
        for (int e = 0; e < n; ++e) {
            int const si = list[e];
            BatchSeg const sg = b.seg[si];
            int const chunks = (sg.n + 31) / 32;
            for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
                int const g = sg.qOff + c * 32 + r;
                if (g - sg.qOff < sg.n)
                    merge_decide_row(b, part, si, g, sub, epsRel, minScore, minRatio, sel, flagList, counters);
            }
        }
    }
}

void launch_batch_merge_decide(BatchDev const& b, Top2 const* part, float epsRel, float minScore, float minRatio,
                               int* sel, int* flagList, int* counters, hipStream_t st)
{
    if (b.qRows <= 0) return;
    hipLaunchKernelGGL(k_batch_merge_decide, dim3((b.qRows + 31) / 32), dim3(256), 0, st, b, part, epsRel, minScore,
                       minRatio, sel, flagList, counters);
}

void launch_batch_merge_masked(BatchDev const& b, Top2 const* part, float epsRel, float minScore, float minRatio,
                               int* sel, int* flagList, int* counters, int const* segMask, int grid, hipStream_t st)
{
    if (b.nPairs <= 0) return;
    hipLaunchKernelGGL(k_batch_merge_masked, dim3(grid), dim3(256), 0, st, b, part, epsRel, minScore, minRatio, sel,
                       flagList, counters, segMask);
}

// Exact decision of one open query per work-group (grid-stride over the flag list).
// Thread t scores candidates t, t + 256, ... as fl32(sum_k double(q_k) * double(c_k)) in
// ascending k: the candidates pass through LDS 256 rows at a time, in two halves of 64
// components, and each thread carries its FP64 sum across the halves, so the order of the
// sum is the oracle's.  The per-thread top-2 states (ascending candidates) are merged by
// an order-independent merge.
constexpr int kBRecheckRows = 256;
constexpr int kBRecheckStride = 68;  // floats per staged half row (64 + 4 pad)

__global__ __launch_bounds__(256) void k_batch_recheck(BatchDev b, int const* __restrict__ flagList,
                                                       int const* __restrict__ counters, float minScore,
                                                       float minRatio, int* __restrict__ sel)
{
    __shared__ __attribute__((aligned(16))) float cand[kBRecheckRows * kBRecheckStride];
    __shared__ double qd[kDim];
    __shared__ Top2 wtop[4];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const nFlag = counters[kBCntFlag];
    for (int w = blockIdx.x; w < nFlag; w += gridDim.x) {
        int const g = flagList[w];
        BatchSeg const sg = b.seg[b.segOfBlock[g / kBatchRowAlign]];
        int const q = g - sg.qOff;
        float const* Qd = b.img + (size_t)(b.imgBase[sg.img] + q) * kDim;
        float const* Cd = b.img + (size_t)b.imgBase[sg.oImg] * kDim;
        __syncthreads();  // the previous query's LDS contents are consumed
        if (tid < kDim) qd[tid] = (double)Qd[tid];
        Top2 s = top2_init();
        for (int c0 = 0; c0 < sg.nO; c0 += kBRecheckRows) {
            double acc = 0.0;
            for (int half = 0; half < 2; ++half) {
                __syncthreads();  // every thread is done with the staged half
                // 256 rows x 64 floats = 4096 float4: 16 per thread, row-contiguous
#pragma unroll 4
                for (int u = 0; u < 16; ++u) {
                    int const f = u * 256 + tid;
                    int const r = f >> 4, c4 = f & 15;
                    // rows past nO are the image's zero padding (images are padded to
                    // multiples of 256 rows): in bounds, and masked below
                    float4 const v =
                        reinterpret_cast<float4 const*>(Cd + (size_t)(c0 + r) * kDim + half * 64)[c4];
                    *reinterpret_cast<float4*>(&cand[r * kBRecheckStride + 4 * c4]) = v;
                }
                __syncthreads();
                float const* cr = &cand[tid * kBRecheckStride];
                double const* qh = qd + half * 64;
#pragma unroll 8
                for (int k4 = 0; k4 < 16; ++k4) {
                    float4 const y = *reinterpret_cast<float4 const*>(cr + 4 * k4);
                    acc += qh[4 * k4 + 0] * (double)y.x;
                    acc += qh[4 * k4 + 1] * (double)y.y;
                    acc += qh[4 * k4 + 2] * (double)y.z;
                    acc += qh[4 * k4 + 3] * (double)y.w;
                }
            }
            int const j = c0 + tid;
            float const sc = (float)acc;
            if (j < sg.nO && sc == sc) top2_push(s, sc, j);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s = top2_merge(s, top2_shfl_xor(s, o));
        if (lane == 0) wtop[wave] = s;
        __syncthreads();
        if (tid == 0) {
            Top2 t = top2_merge(top2_merge(wtop[0], wtop[1]), top2_merge(wtop[2], wtop[3]));
            sel[g] = decide_exact(t, minScore, minRatio);
        }
    }
}

void launch_batch_recheck(BatchDev const& b, int const* flagList, int* counters, float minScore, float minRatio,
                          int* sel, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_batch_recheck, dim3(grid), dim3(256), 0, st, b, flagList, counters, minScore, minRatio, sel);
}

// Mutual check over the left-query rows [0, leftRows) of the query space (segments 0 ..
// nPairs-1 in pair order): left i of pair p with sel = l >= 0 is kept iff the right query
// l of pair p chose i (cu:546-551).  Stable compaction in query-space order keeps every
// pair's matches contiguous and sorted by left index; pairCount[p] counts them.
__global__ __launch_bounds__(256) void k_batch_mutual_local(BatchDev b, int const* __restrict__ sel, int leftRows,
                                                            int2* __restrict__ tmp, int* __restrict__ blockCount,
                                                            int* __restrict__ pairCount)
{
    __shared__ int wsum[4];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const g0 = blockIdx.x * 1024 + 4 * tid;  // 4 rows per thread, one 256-row block per wave
    int li[4], lr[4];
    bool m[4];
    int cnt = 0;
    int pairId = -1;
    if (g0 < leftRows) {
        int const si = b.segOfBlock[g0 / kBatchRowAlign];
        BatchSeg const sl = b.seg[si];
        BatchSeg const sr = b.seg[b.nPairs + si];
        pairId = si;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const i = g0 + u - sl.qOff;
            int const l = (i < sl.n) ? sel[g0 + u] : -1;
            li[u] = i;
            lr[u] = l;
            m[u] = (l >= 0) && (l < sr.n) && (sel[sr.qOff + l] == i);
            cnt += m[u] ? 1 : 0;
        }
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) m[u] = false;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int const t = __shfl_up(incl, o);
        incl += (lane >= o) ? t : 0;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    int pos = blockIdx.x * 1024 + off + incl - cnt;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (m[u]) tmp[pos++] = make_int2(li[u], lr[u]);
    // one atomic per wave: its 256 rows are one segment
    int wcnt = cnt;
    for (int o = 32; o > 0; o >>= 1) wcnt += __shfl_xor(wcnt, o);
    if (lane == 0 && wcnt > 0) atomicAdd(&pairCount[pairId], wcnt);
    if (tid == 0) blockCount[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Single work-group: exclusive prefix of the block counts -> blockOff (each thread sums a
// run of blocks, then a work-group scan), and the run's totals.  The copy into place is a
// separate kernel with one work-group per block.
__global__ __launch_bounds__(1024) void k_batch_mutual_prefix(int const* __restrict__ blockCount, int nBlocks,
                                                              int* __restrict__ blockOff, int* __restrict__ stripes,
                                                              int* __restrict__ counters)
{
    __shared__ int wsum[16];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int rescored = 0;
    if (tid < kStripes) {
        rescored = stripes[tid * kStripeStride];
        stripes[tid * kStripeStride] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) rescored += __shfl_xor(rescored, o);  // wave 0 holds the stripes
    int const per = (nBlocks + 1023) / 1024;
    int const b0 = tid * per;
    int mine = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nBlocks) mine += blockCount[b0 + k];
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int const t = __shfl_up(incl, o);
        incl += (lane >= o) ? t : 0;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int off = incl - mine;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    for (int k = 0; k < per; ++k) {
        int const blk = b0 + k;
        if (blk >= nBlocks) break;
        blockOff[blk] = off;
        off += blockCount[blk];
    }
    if (tid == 0) {
        int total = 0;
        for (int w = 0; w < 16; ++w) total += wsum[w];
        counters[kBCntPairs] = total;
        counters[kBCntFlagDone] = counters[kBCntFlag];
        counters[kBCntRescoredDone] = rescored;
    }
}

__global__ __launch_bounds__(256) void k_batch_mutual_copy(int2 const* __restrict__ tmp,
                                                           int const* __restrict__ blockCount,
                                                           int const* __restrict__ blockOff, int2* __restrict__ pairs)
{
    int const blk = blockIdx.x;
    int const n = blockCount[blk];
    int const o = blockOff[blk];
    for (int e = threadIdx.x; e < n; e += 256) pairs[o + e] = tmp[(size_t)blk * 1024 + e];
}

void launch_batch_mutual(BatchDev const& b, int const* sel, int leftRows, int2* tmp, int* blockCount, int2* pairs,
                         int* pairCount, int* stripes, int* counters, hipStream_t st)
{
    int const nb = (leftRows + 1023) / 1024;
    int* blockOff = blockCount + nb;  // the count buffer holds 2 * nb ints
    if (nb > 0)
        hipLaunchKernelGGL(k_batch_mutual_local, dim3(nb), dim3(256), 0, st, b, sel, leftRows, tmp, blockCount,
                           pairCount);
    hipLaunchKernelGGL(k_batch_mutual_prefix, dim3(1), dim3(1024), 0, st, blockCount, nb, blockOff, stripes,
                       counters);
    if (nb > 0) hipLaunchKernelGGL(k_batch_mutual_copy, dim3(nb), dim3(256), 0, st, tmp, blockCount, blockOff, pairs);
}

// Input rows -> the aligned image buffer (one wave per destination row).  Row r of the
// buffer belongs to image m = blkImg[r / kBatchRowAlign]; its local index r - imgBase[m] is
// a feature when below nImg[m] (the count rounded down to 32) and zero padding otherwise.
// The uint8 rows are also kept in the aligned layout with their scale (imgU8, denom): the
// exact rescoring of the shortlist forms reads 128 bytes per candidate instead of 512 and
// forms the same float values.  The sampled check's row constants: rowCS.x (128 * the
// component sum, kDegenerateC for all-zero and padding rows) and rowCS.y (the bits of the
// scale, or of 1).
__global__ __launch_bounds__(256) void k_batch_load_u8(uint8_t const* __restrict__ src, int const* __restrict__ srcRow,
                                                       int const* __restrict__ nImg, int const* __restrict__ blkImg,
                                                       int const* __restrict__ imgBase, int rows,
                                                       float* __restrict__ img, uint8_t* __restrict__ imgU8,
                                                       float* __restrict__ denom, int2* __restrict__ rowCS)
{
    int const lane = threadIdx.x & 63;
    int const row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    int const m = blkImg[row / kBatchRowAlign];
    int const local = row - imgBase[m];
    float2 v = make_float2(0.0f, 0.0f);
    uchar2 q = make_uchar2(0, 0);
    float d = 0.0f;
    if (local < nImg[m]) {
        q = reinterpret_cast<uchar2 const*>(src + (size_t)(srcRow[m] + local) * kDim)[lane];
        d = dequant_u8_denom(q);
        v = make_float2(dequant_u8_value(q.x, d), dequant_u8_value(q.y, d));
    }
    int sum = (int)q.x + (int)q.y;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    reinterpret_cast<float2*>(img + (size_t)row * kDim)[lane] = v;
    reinterpret_cast<uchar2*>(imgU8 + (size_t)row * kDim)[lane] = q;
    if (lane == 0) {
        bool const valid = d > 0.0f && d < INFINITY;  // a feature with a nonzero component
        denom[row] = d;
        rowCS[row] = make_int2(valid ? 128 * sum : kDegenerateC, __float_as_int(valid ? d : 1.0f));
    }
}

__global__ __launch_bounds__(256) void k_batch_load_f32(float const* __restrict__ src, int const* __restrict__ srcRow,
                                                        int const* __restrict__ nImg, int const* __restrict__ blkImg,
                                                        int const* __restrict__ imgBase, int rows,
                                                        float* __restrict__ img)
{
    int const lane = threadIdx.x & 63;
    int const row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= rows) return;
    int const m = blkImg[row / kBatchRowAlign];
    int const local = row - imgBase[m];
    float2 v = make_float2(0.0f, 0.0f);
    if (local < nImg[m]) v = reinterpret_cast<float2 const*>(src + (size_t)(srcRow[m] + local) * kDim)[lane];
    reinterpret_cast<float2*>(img + (size_t)row * kDim)[lane] = v;
}

void launch_batch_load_u8(uint8_t const* src, int const* srcRow, int const* nImg, int const* blkImg,
                          int const* imgBase, int rows, float* img, uint8_t* imgU8, float* denom, int2* rowCS,
                          hipStream_t st)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_batch_load_u8, dim3((rows + 3) / 4), dim3(256), 0, st, src, srcRow, nImg, blkImg, imgBase,
                       rows, img, imgU8, denom, rowCS);
}

// The BF16 copy of a uint8 batch from the uint8 rows and their scales (128 + 4 bytes read per
// row instead of the 512 of the FP32 rows): each element is formed as k_batch_load_u8 forms
// it (dequant_u8_value, the same IEEE product) and rounded to bf16 (nearest even), so the copy
// equals the conversion of the FP32 image bit for bit.  Eight elements per thread.
__global__ __launch_bounds__(256) void k_batch_u8_to_bf16(uint8_t const* __restrict__ u8,
                                                          float const* __restrict__ denom, int rows,
                                                          __bf16* __restrict__ Y)
{
    int const t = blockIdx.x * 256 + threadIdx.x;
    int const row = t >> 4;
    if (row >= rows) return;
    uint2 const q = reinterpret_cast<uint2 const*>(u8)[t];  // elements 8 (t & 15) .. +7 of the row
    float const d = denom[row];
    __bf16 out[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        out[k] = (__bf16)dequant_u8_value((q.x >> (8 * k)) & 0xffu, d);
        out[4 + k] = (__bf16)dequant_u8_value((q.y >> (8 * k)) & 0xffu, d);
    }
    reinterpret_cast<uint4*>(Y)[t] = *reinterpret_cast<uint4 const*>(out);
}

void launch_batch_u8_to_bf16(uint8_t const* u8, float const* denom, int rows, __bf16* Y, hipStream_t st)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_batch_u8_to_bf16, dim3((rows * 16 + 255) / 256), dim3(256), 0, st, u8, denom, rows, Y);
}

void launch_batch_load_f32(float const* src, int const* srcRow, int const* nImg, int const* blkImg,
                           int const* imgBase, int rows, float* img, hipStream_t st)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_batch_load_f32, dim3((rows + 3) / 4), dim3(256), 0, st, src, srcRow, nImg, blkImg, imgBase,
                       rows, img);
}

// ------------------------------------------------------------------------------ sampled check
// FP32 images (uint8 batches use v3dm_check.hip).  Per pair p, S_p = min(samples, n) evenly
// spaced features of the side sample_side picks (its larger image, left on a tie): rows,
// norms and indices gathered into kSampleRows-row blocks (zero rows past S_p).  One wave per
// sample row.
__global__ __launch_bounds__(256) void k_batch_sample_gather(BatchDev b, int samples, float* __restrict__ sRows,
                                                             float* __restrict__ sNorm, int* __restrict__ sIdx,
                                                             int* __restrict__ sCount)
{
    int const lane = threadIdx.x & 63;
    int const r = blockIdx.x * 4 + (threadIdx.x >> 6);
    int const p = r / kSampleRows, k = r - p * kSampleRows;
    if (p >= b.nPairs) return;
    BatchSeg const sl = b.seg[sample_side(b, p) * b.nPairs + p];
    int const S = min(samples, sl.n);
    int const i = (k < S) ? (int)(((long long)k * sl.n) / S) : -1;
    int const row = b.imgBase[sl.img] + (i >= 0 ? i : 0);
    float2 const v = (i >= 0) ? reinterpret_cast<float2 const*>(b.img + (size_t)row * kDim)[lane]
                              : make_float2(0.f, 0.f);
    reinterpret_cast<float2*>(sRows + (size_t)r * kDim)[lane] = v;
    if (lane == 0) {
        sNorm[r] = (i >= 0) ? b.norm[row] : 0.0f;
        sIdx[r] = i;
        if (k == 0) sCount[p] = S;
    }
}

// One work-group (kSampleRows threads) per pair: merge the sample's split partials, take
// the certified FP32 decision (samples whose decision is not certified -- near ties, well
// under 1% -- are left out of the estimate), compare with the shortlist decision sel of the
// same feature of the sampled side.  pairFlag[p] = 1 when the misses exceed
// lim.allowed[accepted]; counts[4p..] = accepted, missed, left out.
__global__ __launch_bounds__(kSampleRows) void k_batch_sample_decide(BatchDev b, Top2 const* __restrict__ sPart,
                                                                     int maxSplits, int const* __restrict__ nSplits,
                                                                     float const* __restrict__ sNorm,
                                                                     int const* __restrict__ sIdx,
                                                                     int const* __restrict__ sCount, float epsRel,
                                                                     float minScore, float minRatio,
                                                                     MissLimits lim, int const* __restrict__ sel,
                                                                     int* __restrict__ pairFlag,
                                                                     int* __restrict__ counts)
{
    __shared__ int red[2][3];
    int const p = blockIdx.x;
    int const k = threadIdx.x;
    BatchSeg const sl = b.seg[sample_side(b, p) * b.nPairs + p];
    int const S = sCount[p];
    int acc = 0, miss = 0, out = 0;
    if (k < S && sl.nO > 0) {
        Top2 s = top2_init();
        Top2 const* pp = sPart + (size_t)p * maxSplits * kSampleRows + k;
        for (int sp = 0; sp < nSplits[p]; ++sp) s = top2_merge(s, pp[(size_t)sp * kSampleRows]);
        float const eps = epsRel * sNorm[p * kSampleRows + k] * __uint_as_float(b.maxNorm[sl.oImg]) + kEpsAbs;
        int const e = (epsRel > 0.0f) ? decide_certified(s, eps, minScore, minRatio) : kFlag;
        if (e == kFlag) {
            out = 1;
        } else {
            int const d = sel[sl.qOff + sIdx[p * kSampleRows + k]];
            acc = e >= 0;
            miss = (e >= 0 && d != e);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_xor(acc, o);
        miss += __shfl_xor(miss, o);
        out += __shfl_xor(out, o);
    }
    if ((k & 63) == 0) {
        red[k >> 6][0] = acc;
        red[k >> 6][1] = miss;
        red[k >> 6][2] = out;
    }
    __syncthreads();
    if (k == 0) {
        int const a = red[0][0] + red[1][0], m = red[0][1] + red[1][1], x = red[0][2] + red[1][2];
        counts[4 * p + 0] = a;
        counts[4 * p + 1] = m;
        counts[4 * p + 2] = x;
        counts[4 * p + 3] = S;
        pairFlag[p] = (m > (int)lim.allowed[a]) ? 1 : 0;
    }
}

void launch_batch_sample_gather(BatchDev const& b, int samples, float* sRows, float* sNorm, int* sIdx, int* sCount,
                                hipStream_t st)
{
    int const rows = b.nPairs * kSampleRows;
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_batch_sample_gather, dim3((rows + 3) / 4), dim3(256), 0, st, b, samples, sRows, sNorm, sIdx,
                       sCount);
}

void launch_batch_sample_decide(BatchDev const& b, Top2 const* sPart, int maxSplits, int const* nSplits,
                                float const* sNorm, int const* sIdx, int const* sCount, float epsRel, float minScore,
                                float minRatio, MissLimits const& lim, int const* sel, int* pairFlag, int* counts,
                                hipStream_t st)
{
    if (b.nPairs <= 0) return;
    hipLaunchKernelGGL(k_batch_sample_decide, dim3(b.nPairs), dim3(kSampleRows), 0, st, b, sPart, maxSplits, nSplits,
                       sNorm, sIdx, sCount, epsRel, minScore, minRatio, lim, sel, pairFlag, counts);
}

}  // namespace v3dm
