// Exact rescoring of the shortlist forms (BF16, FP8, learned codes), in two kernels.
//
//  select: one lane per query scans the query's partition lists (two candidates per
//          (split, half) list, approximate scores), fixes the threshold -- either the
//          window v >= a2 - winRel * |q| * max|c| below the approximate second best a2, or
//          the topK-th largest approximate score -- and writes the candidates at or above
//          it to the query's slot row (at most kRescoreSlots, which every list fits).
//  score:  eight lanes per query split the query's slots; each lane forms the exact score
//          fl32(sum_k double(q_k) * double(c_k)) of its candidate in ascending k (one FP64
//          chain per lane, rows read straight from memory); the eight partial top-2 states
//          merge by xor shuffles; the decision is the reference's rule on the exact top-2.
// The second best candidate (by exact score) is kept as well: it labels training data.
#include "v3dm_batch.h"
#include "v3dm_kernels.h"
#include "v3dm_short.h"

namespace v3dm {

namespace {

struct Top2x {
    float b1, b2;
    int a1, a2;
};

// larger best wins, equal best keeps the smaller index; second = second largest of the
// union (index of the smaller candidate among equal values)
__device__ __forceinline__ Top2x top2x_merge(Top2x a, Top2x b)
{
    bool const takeB = (b.b1 > a.b1) || (b.b1 == a.b1 && b.a1 >= 0 && (a.a1 < 0 || b.a1 < a.a1));
    Top2x const w = takeB ? b : a;
    Top2x const l = takeB ? a : b;
    bool const lo = (l.b1 > w.b2) || (l.b1 == w.b2 && l.a1 >= 0 && (w.a2 < 0 || l.a1 < w.a2));
    return Top2x{w.b1, lo ? l.b1 : w.b2, w.a1, lo ? l.a1 : w.a2};
}

__device__ __forceinline__ Top2x top2x_shfl_xor(Top2x s, int mask)
{
    return Top2x{__shfl_xor(s.b1, mask), __shfl_xor(s.b2, mask), __shfl_xor(s.a1, mask), __shfl_xor(s.a2, mask)};
}

// Threshold and slot list of one query (lane-per-query).  part points at the query's
// (split 0, half 0) entry; entries of split s are at part[s * ldq * 2 + h].  kMax bounds
// topK (the sorted register array of the topK largest values).
template <int kMax>
__device__ __forceinline__ int select_query(Cand2 const* __restrict__ part, size_t ldq2, int nParts, float nq,
                                            float mo, float winRel, int topK, int* __restrict__ slots)
{
    float thr;
    if (topK > 0) {
        float top[kMax];
#pragma unroll
        for (int i = 0; i < kMax; ++i) top[i] = -INFINITY;
        for (int s = 0; s < nParts; ++s)
            for (int h = 0; h < 2; ++h) {
                Cand2 const c = part[s * ldq2 + h];
                float const vs[2] = {c.j1 >= 0 ? c.v1 : -INFINITY, c.j2 >= 0 ? c.v2 : -INFINITY};
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    float v = vs[u];
#pragma unroll
                    for (int i = 0; i < kMax; ++i) {  // insertion: keeps top[] descending
                        float const hi = fmaxf(top[i], v);
                        v = fminf(top[i], v);
                        top[i] = hi;
                    }
                }
            }
        float kth = top[0];
#pragma unroll
        for (int i = 1; i < kMax; ++i) kth = (i < topK && top[i] > -INFINITY) ? top[i] : kth;
        thr = kth;
    } else {
        float a1 = -INFINITY, a2 = -INFINITY;
        for (int s = 0; s < nParts; ++s)
            for (int h = 0; h < 2; ++h) {
                Cand2 const c = part[s * ldq2 + h];
                float const x = c.j1 >= 0 ? c.v1 : -INFINITY;
                float const y = c.j2 >= 0 ? c.v2 : -INFINITY;
                a2 = fmaxf(a2, fminf(a1, x));
                a1 = fmaxf(a1, x);
                a2 = fmaxf(a2, fminf(a1, y));
                a1 = fmaxf(a1, y);
            }
        thr = a2 - winRel * nq * mo;
    }
    int n = 0;
    for (int s = 0; s < nParts; ++s)
        for (int h = 0; h < 2; ++h) {
            Cand2 const c = part[s * ldq2 + h];
            if (c.j1 >= 0 && c.v1 >= thr && n < kRescoreSlots) slots[n++] = c.j1;
            if (c.j2 >= 0 && c.v2 >= thr && n < kRescoreSlots) slots[n++] = c.j2;
        }
    return n;
}

__device__ __forceinline__ int select_query_any(Cand2 const* __restrict__ part, size_t ldq2, int nParts, float nq,
                                                float mo, float winRel, int topK, int* __restrict__ slots)
{
    if (topK <= 8) return select_query<8>(part, ldq2, nParts, nq, mo, winRel, topK, slots);
    return select_query<kRescoreTopKMax>(part, ldq2, nParts, nq, mo, winRel, topK, slots);
}

// Exact top-2 of one query over its slots; 8 lanes (sub = lane & 7) per query.  Returns
// the merged state in every lane of the group.
__device__ __forceinline__ Top2x score_query(float const* __restrict__ Qd, float const* __restrict__ Cd,
                                             int const* __restrict__ slots, int n, int sub)
{
    Top2x s = {kInit, kInit, -1, -1};
    for (int k = sub; k < n; k += 8) {
        int const j = slots[k];
        float4 const* qp = reinterpret_cast<float4 const*>(Qd);
        float4 const* cp = reinterpret_cast<float4 const*>(Cd + (size_t)j * kDim);
        double acc = 0.0;
#pragma unroll 8
        for (int k4 = 0; k4 < 32; ++k4) {
            float4 const x = qp[k4];
            float4 const y = cp[k4];
            acc += (double)x.x * (double)y.x;
            acc += (double)x.y * (double)y.y;
            acc += (double)x.z * (double)y.z;
            acc += (double)x.w * (double)y.w;
        }
        float const sc = (float)acc;
        if (sc == sc) s = top2x_merge(s, Top2x{sc, kInit, j, -1});
    }
    s = top2x_merge(s, top2x_shfl_xor(s, 1));
    s = top2x_merge(s, top2x_shfl_xor(s, 2));
    s = top2x_merge(s, top2x_shfl_xor(s, 4));
    return s;
}

}  // namespace

// ------------------------------------------------------------------------------ one pair
// Queries [0, nL) are left features (lists partL), [nL, nL + nR) right features.
__global__ __launch_bounds__(256) void k_rescore_select(Cand2 const* __restrict__ partL, int nPartL, int ldqL,
                                                        Cand2 const* __restrict__ partR, int nPartR, int ldqR,
                                                        int nL, int nR, float const* __restrict__ normL,
                                                        float const* __restrict__ normR,
                                                        unsigned int const* __restrict__ maxNormBits, float winRel,
                                                        int topK, int* __restrict__ slots, int* __restrict__ count)
{
    int const g = blockIdx.x * 256 + threadIdx.x;
    if (g >= nL + nR) return;
    bool const left = g < nL;
    int const q = left ? g : g - nL;
    Cand2 const* part = (left ? partL : partR) + (size_t)q * 2;
    size_t const ldq2 = 2 * (size_t)(left ? ldqL : ldqR);
    float const nq = left ? normL[q] : normR[q];
    float const mo = __uint_as_float(maxNormBits[left ? 1 : 0]);
    count[g] = select_query_any(part, ldq2, left ? nPartL : nPartR, nq, mo, winRel, topK,
                                slots + (size_t)g * kRescoreSlots);
}

__global__ __launch_bounds__(256) void k_rescore_score(float const* __restrict__ L, float const* __restrict__ R,
                                                       int nL, int nR, int const* __restrict__ slots,
                                                       int const* __restrict__ count, float minScore,
                                                       float minRatio, int* __restrict__ colMax,
                                                       int* __restrict__ rowMax, int* __restrict__ colSecond,
                                                       int* __restrict__ rowSecond, int* __restrict__ stripes)
{
    __shared__ int wsum[4];
    int const tid = threadIdx.x;
    int const sub = tid & 7;
    int const g = blockIdx.x * 32 + (tid >> 3);
    int n = 0;
    if (g < nL + nR) {  // uniform over the 8 lanes of a query
        bool const left = g < nL;
        int const q = left ? g : g - nL;
        n = count[g];
        Top2x const s = score_query((left ? L : R) + (size_t)q * kDim, left ? R : L,
                                    slots + (size_t)g * kRescoreSlots, n, sub);
        if (sub == 0) {
            (left ? colMax : rowMax)[q] = decide_exact(Top2{s.b1, s.b2, s.a1}, minScore, minRatio);
            if (colSecond) (left ? colSecond : rowSecond)[q] = s.a2;
        }
    }
    int c = (sub == 0) ? n : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((tid & 63) == 0) wsum[tid >> 6] = c;
    __syncthreads();
    if (tid == 0) atomicAdd(&stripes[(blockIdx.x % kStripes) * kStripeStride], wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

void launch_rescore(float const* L, float const* R, int nL, int nR, Cand2 const* partL, int nPartL, int ldqL,
                    Cand2 const* partR, int nPartR, int ldqR, float const* normL, float const* normR,
                    unsigned int const* maxNormBits, float winRel, int topK, float minScore, float minRatio,
                    int* slots, int* count, int* colMax, int* rowMax, int* colSecond, int* rowSecond, int* stripes,
                    hipStream_t st)
{
    int const n = nL + nR;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rescore_select, dim3((n + 255) / 256), dim3(256), 0, st, partL, nPartL, ldqL, partR,
                       nPartR, ldqR, nL, nR, normL, normR, maxNormBits, winRel, topK, slots, count);
    hipLaunchKernelGGL(k_rescore_score, dim3((n + 31) / 32), dim3(256), 0, st, L, R, nL, nR, slots, count, minScore,
                       minRatio, colMax, rowMax, colSecond, rowSecond, stripes);
}

// ------------------------------------------------------------------------------ batch
// Query-space rows (padding rows get count 0 and decision -1).
__global__ __launch_bounds__(256) void k_batch_rescore_select(BatchDev b, Cand2 const* __restrict__ part,
                                                              float winRel, int topK, int* __restrict__ slots,
                                                              int* __restrict__ count)
{
    int const g = blockIdx.x * 256 + threadIdx.x;
    if (g >= b.qRows) return;
    BatchSeg const sg = b.seg[b.segOfBlock[g / kBatchRowAlign]];
    int const q = g - sg.qOff;
    int n = 0;
    if (q < sg.n && sg.nO > 0 && sg.nParts > 0) {
        int const row = b.imgBase[sg.img] + q;
        n = select_query_any(part + ((size_t)sg.partOff + q) * 2, 2 * (size_t)sg.ld, sg.nParts, b.norm[row],
                         __uint_as_float(b.maxNorm[sg.oImg]), winRel, topK, slots + (size_t)g * kRescoreSlots);
    }
    count[g] = n;
}

__global__ __launch_bounds__(256) void k_batch_rescore_score(BatchDev b, int const* __restrict__ slots,
                                                             int const* __restrict__ count, float minScore,
                                                             float minRatio, int* __restrict__ sel,
                                                             int* __restrict__ stripes)
{
    __shared__ int wsum[4];
    int const tid = threadIdx.x;
    int const sub = tid & 7;
    int const g = blockIdx.x * 32 + (tid >> 3);
    int n = 0;
    if (g < b.qRows) {
        BatchSeg const sg = b.seg[b.segOfBlock[g / kBatchRowAlign]];
        int const q = g - sg.qOff;
        n = count[g];
        int d = -1;
        if (q < sg.n && n > 0) {
            float const* Qd = b.img + (size_t)(b.imgBase[sg.img] + q) * kDim;
            float const* Cd = b.img + (size_t)b.imgBase[sg.oImg] * kDim;
            Top2x const s = score_query(Qd, Cd, slots + (size_t)g * kRescoreSlots, n, sub);
            d = decide_exact(Top2{s.b1, s.b2, s.a1}, minScore, minRatio);
        }
        if (sub == 0) sel[g] = d;
    }
    int c = (sub == 0) ? n : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((tid & 63) == 0) wsum[tid >> 6] = c;
    __syncthreads();
    if (tid == 0) atomicAdd(&stripes[(blockIdx.x % kStripes) * kStripeStride], wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

void launch_batch_rescore(BatchDev const& b, Cand2 const* part, float winRel, int topK, float minScore,
                          float minRatio, int* slots, int* count, int* sel, int* stripes, hipStream_t st)
{
    if (b.qRows <= 0) return;
    hipLaunchKernelGGL(k_batch_rescore_select, dim3((b.qRows + 255) / 256), dim3(256), 0, st, b, part, winRel, topK,
                       slots, count);
    hipLaunchKernelGGL(k_batch_rescore_score, dim3((b.qRows + 31) / 32), dim3(256), 0, st, b, slots, count, minScore,
                       minRatio, sel, stripes);
}

}  // namespace v3dm
