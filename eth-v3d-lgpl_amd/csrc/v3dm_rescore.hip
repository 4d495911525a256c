// Exact rescoring of the shortlist forms (BF16, FP8, learned codes): one kernel, eight
// lanes per query.
//
// Lane `sub` of a query's group holds the query's partition lists e = sub, sub + 8, ...
// (two candidates with approximate scores per (split, half) list).  The group fixes the
// threshold -- either the window v >= a2 - winRel * |q| * max|c| below the approximate
// second best a2, or the topK-th largest approximate score (a tournament over the group:
// each round takes the largest remaining value) -- compacts the candidates at or above it
// into an LDS row, and the lanes share them out: each forms the exact score
// fl32(sum_k double(q_k) * double(c_k)) (ascending k, one FP64 chain, rows read straight
// from memory) of every eighth one.  The lanes' top-2 states merge by
// xor shuffles and the reference's rule decides on the exact top-2.  The second best
// candidate by exact score is kept as well: it labels training data.
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

constexpr int kPerLane = kRescoreSlots / 8;  // candidates one lane holds (kShortMaxSplits lists)

// Exact score (the oracle's: FP64 products, summed in ascending k).  q: the query row in
// LDS; c: the candidate row, loaded sixteen float4 at a time so that the loads overlap.
__device__ __forceinline__ double exact_dot(float const* q, float const* __restrict__ c)
{
    float4 const* qp = reinterpret_cast<float4 const*>(q);
    float4 const* cp = reinterpret_cast<float4 const*>(c);
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float4 y[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) y[u] = cp[16 * h + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            float4 const x = qp[16 * h + u];
            acc += (double)x.x * (double)y[u].x;
            acc += (double)x.y * (double)y[u].y;
            acc += (double)x.z * (double)y[u].z;
            acc += (double)x.w * (double)y[u].w;
        }
    }
    return acc;
}

// Exact score against a uint8 candidate row: element k = (float)c_k * denom, the value the
// FP32 image holds (same IEEE product).  The whole row (128 bytes) is loaded at once.
__device__ __forceinline__ double exact_dot_u8(float const* q, uint8_t const* __restrict__ c, float d)
{
    float4 const* qp = reinterpret_cast<float4 const*>(q);
    uint4 const* cp = reinterpret_cast<uint4 const*>(c);
    uint4 w[8];
#pragma unroll
    for (int k16 = 0; k16 < 8; ++k16) w[k16] = cp[k16];
    double acc = 0.0;
#pragma unroll
    for (int k16 = 0; k16 < 8; ++k16) {
        unsigned int const ws[4] = {w[k16].x, w[k16].y, w[k16].z, w[k16].w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float4 const x = qp[4 * k16 + u];
            acc += (double)x.x * (double)dequant_u8_value(ws[u] & 0xffu, d);
            acc += (double)x.y * (double)dequant_u8_value((ws[u] >> 8) & 0xffu, d);
            acc += (double)x.z * (double)dequant_u8_value((ws[u] >> 16) & 0xffu, d);
            acc += (double)x.w * (double)dequant_u8_value(ws[u] >> 24, d);
        }
    }
    return acc;
}

// Candidate selection of one query, eight lanes (sub = lane & 7).  part: the query's
// (split 0, half 0) list; the lists of split s are at part[s * ldq2 + h]; slots: the
// group's kRescoreSlots ints of LDS.  The group fixes the threshold (window below the
// approximate second best, or the topK tournament with the margin), compacts the
// candidates at or above it into its slot row (in lane order) and returns their count
// (uniform over the group).
template <typename List>
__device__ __forceinline__ int select_candidates(List const* __restrict__ part, size_t ldq2, int nParts, int span,
                                                 float nq, float mo, float winRel, int topK, float margin, int sub,
                                                 int* slots)
{
    float v[kPerLane];
    int j[kPerLane];
#pragma unroll
    for (int u = 0; u < kPerLane / 2; ++u) {
        int const e = sub + 8 * u;  // list (split e / 2, half e % 2)
        if constexpr (sizeof(List) == sizeof(CandP)) {
            unsigned int const empty = (unsigned int)(kPackedEmpty & 0xffff) << 16;
            CandP c = {empty, empty};
            if (e < 2 * nParts) c = part[(size_t)(e >> 1) * ldq2 + (e & 1)];
            int const base = (e >> 1) * span;
            int const s1 = (int)c.a >> 16, s2 = (int)c.b >> 16;
            v[2 * u] = s1 != kPackedEmpty ? (float)s1 : -INFINITY;
            j[2 * u] = s1 != kPackedEmpty ? base + (int)(c.a & 0xffffu) : -1;
            v[2 * u + 1] = s2 != kPackedEmpty ? (float)s2 : -INFINITY;
            j[2 * u + 1] = s2 != kPackedEmpty ? base + (int)(c.b & 0xffffu) : -1;
        } else {
            Cand2 c = {-INFINITY, -1, -INFINITY, -1};
            if (e < 2 * nParts) c = part[(size_t)(e >> 1) * ldq2 + (e & 1)];
            v[2 * u] = c.j1 >= 0 ? c.v1 : -INFINITY;
            j[2 * u] = c.j1;
            v[2 * u + 1] = c.j2 >= 0 ? c.v2 : -INFINITY;
            j[2 * u + 1] = c.j2;
        }
    }
    float thr;
    int within = kRescoreSlots + 1;
    if (topK > 0) {
        // the group's best value; when at most topK candidates lie within `margin` of it,
        // they are the tournament's result below (it takes them all, in rounds)
        float m = v[0];
#pragma unroll
        for (int u = 1; u < kPerLane; ++u) m = fmaxf(m, v[u]);
        m = fmaxf(m, __shfl_xor(m, 1));
        m = fmaxf(m, __shfl_xor(m, 2));
        m = fmaxf(m, __shfl_xor(m, 4));
        thr = m - margin;
        int c = 0;
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) c += (j[u] >= 0 && !(v[u] < thr)) ? 1 : 0;
        c += __shfl_xor(c, 1);
        c += __shfl_xor(c, 2);
        c += __shfl_xor(c, 4);
        within = (m == -INFINITY) ? 0 : c;
    }
    if (topK > 0 && within <= topK) {
        // thr as set above (no candidate: nothing is at or above it)
    } else if (topK > 0) {
        // tournament: round k takes the largest remaining value of the group; thr = the
        // value taken in round topK, or the last one within `margin` of the first (or the
        // last finite one)
        float w[kPerLane];
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) w[u] = v[u];
        float kth = -INFINITY, top = INFINITY;
        for (int k = 0; k < topK; ++k) {
            float m = w[0];
#pragma unroll
            for (int u = 1; u < kPerLane; ++u) m = fmaxf(m, w[u]);
            float g = m;
            g = fmaxf(g, __shfl_xor(g, 1));
            g = fmaxf(g, __shfl_xor(g, 2));
            g = fmaxf(g, __shfl_xor(g, 4));
            if (g == -INFINITY) break;  // uniform over the group
            if (k == 0) top = g;
            else if (g < top - margin) break;  // uniform over the group
            kth = g;
            // the lowest lane holding g removes one copy
            int own = (m == g) ? sub : 8;
            own = min(own, __shfl_xor(own, 1));
            own = min(own, __shfl_xor(own, 2));
            own = min(own, __shfl_xor(own, 4));
            if (own == sub) {
                bool done = false;
#pragma unroll
                for (int u = 0; u < kPerLane; ++u) {
                    bool const hit = !done && w[u] == g;
                    w[u] = hit ? -INFINITY : w[u];
                    done = done || hit;
                }
            }
        }
        thr = kth;
    } else {
        float a1 = -INFINITY, a2 = -INFINITY;
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) {
            a2 = fmaxf(a2, fminf(a1, v[u]));
            a1 = fmaxf(a1, v[u]);
        }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            float const b1 = __shfl_xor(a1, o), b2 = __shfl_xor(a2, o);
            a2 = fmaxf(fminf(a1, b1), fmaxf(a2, b2));
            a1 = fmaxf(a1, b1);
        }
        thr = a2 - winRel * nq * mo;
    }
    int mine = 0;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) mine += (j[u] >= 0 && v[u] >= thr) ? 1 : 0;
    int incl = mine;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        int const t = __shfl_up(incl, o, 8);
        incl += (sub >= o) ? t : 0;
    }
    int const total = __shfl(incl, 7, 8);
    int pos = incl - mine;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u)
        if (j[u] >= 0 && v[u] >= thr) slots[pos++] = j[u];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // the group's slot writes are visible
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return total;
}

// Exact scores of the group's `total` slot candidates: lane sub scores slots sub, sub + 8,
// ... (the oracle's arithmetic); qs: the query row (128 floats of LDS).  Returns the exact
// top-2 (all lanes) and the number of candidates scored by this lane.
template <bool kU8>
__device__ __forceinline__ Top2x score_exact(float const* qs, float const* __restrict__ Cd,
                                             uint8_t const* __restrict__ Cu8, float const* __restrict__ Cdenom,
                                             int const* slots, int total, int sub, int& nScored)
{
    Top2x s = {kInit, kInit, -1, -1};
    int n = 0;
    for (int k = sub; k < total; k += 8) {
        int const jj = slots[k];
        // an offset the compiler cannot see through: the query row is read from LDS on every
        // candidate rather than held in 128 registers across the loop
        int qo = 0;
        asm volatile("" : "+v"(qo));
        float const sc = kU8 ? (float)exact_dot_u8(qs + qo, Cu8 + (size_t)jj * kDim, Cdenom[jj])
                             : (float)exact_dot(qs + qo, Cd + (size_t)jj * kDim);
        if (sc == sc) s = top2x_merge(s, Top2x{sc, kInit, jj, -1});
        ++n;
    }
    s = top2x_merge(s, top2x_shfl_xor(s, 1));
    s = top2x_merge(s, top2x_shfl_xor(s, 2));
    s = top2x_merge(s, top2x_shfl_xor(s, 4));
    nScored = n;
    return s;
}

// One query, eight lanes: candidate selection, then exact scores of the candidates.  The
// query row is staged in the group's LDS row qs (its loads issued with the list loads).
template <bool kU8, typename List>
__device__ __forceinline__ Top2x finish_query(List const* __restrict__ part, size_t ldq2, int nParts, int span, float nq,
                                              float mo, float winRel, int topK, float margin,
                                              float const* __restrict__ Qd,
                                              float const* __restrict__ Cd, uint8_t const* __restrict__ Cu8,
                                              float const* __restrict__ Cdenom, int sub, int* slots, float* qs,
                                              int& nScored)
{
    {
        float4 qv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) qv[u] = reinterpret_cast<float4 const*>(Qd)[sub + 8 * u];
#pragma unroll
        for (int u = 0; u < 4; ++u) reinterpret_cast<float4*>(qs)[sub + 8 * u] = qv[u];
    }
    int const total = select_candidates(part, ldq2, nParts, span, nq, mo, winRel, topK, margin, sub, slots);
    return score_exact<kU8>(qs, Cd, Cu8, Cdenom, slots, total, sub, nScored);
}

// Certified scores of uint8 rows.  With a_k = fl32(qa_k * ra) and b_k = fl32(qb_k * rb)
// (qa, qb the uint8 components, ra, rb the rows' 1/sqrt scales), the exact score is
// s = fl32(sum_k a_k * b_k).  The integer I = sum_k qa_k * qb_k is exact, and since every
// term is nonnegative, s~ = fl32(fl32(ra * rb) * I) satisfies |s~ - s| <= 5.01 * 2^-24 * s~
// (two rounded dequantisations, the FP64 sum, the FP32 roundings of s and s~), which
// kCertU8 = 2^-21 covers.  The eight lanes of a query hold 16 bytes each of the query row
// (qv) and of every candidate row, so a candidate is one 128-byte segment; four
// v_dot4_u32_u8 per lane and three xor shuffles form I.  All lanes end with the same top-2.
constexpr float kCertU8 = 4.76837158203125e-07f;  // 2^-21

__device__ __forceinline__ Top2x score_u8_certified(uint4 qv, float ra, uint8_t const* __restrict__ Cu8,
                                                    float const* __restrict__ Cdenom, int const* slots, int total,
                                                    int sub)
{
    Top2x s = {kInit, kInit, -1, -1};
    for (int k = 0; k < total; k += 4) {
        uint4 cv[4];
        float rb[4];
        int jj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            jj[u] = (k + u < total) ? slots[k + u] : -1;
            int const jr = max(jj[u], 0);  // row 0 of the candidate image stands in (not used)
            cv[u] = *reinterpret_cast<uint4 const*>(Cu8 + (size_t)jr * kDim + 16 * sub);
            rb[u] = Cdenom[jr];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            unsigned int acc = __builtin_amdgcn_udot4(qv.x, cv[u].x, 0u, false);
            acc = __builtin_amdgcn_udot4(qv.y, cv[u].y, acc, false);
            acc = __builtin_amdgcn_udot4(qv.z, cv[u].z, acc, false);
            acc = __builtin_amdgcn_udot4(qv.w, cv[u].w, acc, false);
            acc += __shfl_xor(acc, 1);
            acc += __shfl_xor(acc, 2);
            acc += __shfl_xor(acc, 4);
            float const sc = mul_nc(mul_nc(ra, rb[u]), (float)acc);
            if (jj[u] >= 0 && sc == sc) s = top2x_merge(s, Top2x{sc, kInit, jj[u], -1});
        }
    }
    return s;
}

__device__ __forceinline__ void add_stripe(int n, int* __restrict__ stripes)
{
    __shared__ int wsum[4];
    int const tid = threadIdx.x;
    int c = n;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((tid & 63) == 0) wsum[tid >> 6] = c;
    __syncthreads();
    if (tid == 0) atomicAdd(&stripes[(blockIdx.x % kStripes) * kStripeStride], wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

}  // namespace

// ------------------------------------------------------------------------------ one pair
// Queries [0, nL) are left features (lists partL), [nL, nL + nR) right features; 32
// queries per work-group.  List: Cand2 (BF16 / FP8 passes) or CandP (learned code pass,
// spanL / spanR candidates per split).
template <typename List>
__global__ __launch_bounds__(256) void k_shortlist_finish(float const* __restrict__ L, float const* __restrict__ R,
                                                          int nL, int nR, List const* __restrict__ partL, int nPartL,
                                                          int ldqL, int spanL, List const* __restrict__ partR,
                                                          int nPartR, int ldqR, int spanR,
                                                          float const* __restrict__ normL,
                                                          float const* __restrict__ normR,
                                                          unsigned int const* __restrict__ maxNormBits, float winRel,
                                                          int topK, float margin, float minScore, float minRatio,
                                                          int* __restrict__ colMax, int* __restrict__ rowMax,
                                                          int* __restrict__ colSecond, int* __restrict__ rowSecond,
                                                          int* __restrict__ stripes)
{
    __shared__ int slots[32][kRescoreSlots];
    __shared__ __attribute__((aligned(16))) float qrow[32][kDim];
    int const sub = threadIdx.x & 7;
    int const g = blockIdx.x * 32 + (threadIdx.x >> 3);
    int n = 0;
    if (g < nL + nR) {  // uniform over the 8 lanes of a query
        bool const left = g < nL;
        int const q = left ? g : g - nL;
        Top2x const s = finish_query<false>((left ? partL : partR) + (size_t)q * 2, 2 * (size_t)(left ? ldqL : ldqR),
                                            left ? nPartL : nPartR, left ? spanL : spanR, left ? normL[q] : normR[q],
                                            __uint_as_float(maxNormBits[left ? 1 : 0]), winRel, topK, margin,
                                            (left ? L : R) + (size_t)q * kDim, left ? R : L, nullptr, nullptr, sub,
                                            slots[threadIdx.x >> 3], qrow[threadIdx.x >> 3], n);
        if (sub == 0) {
            (left ? colMax : rowMax)[q] = decide_exact(Top2{s.b1, s.b2, s.a1}, minScore, minRatio);
            if (colSecond) (left ? colSecond : rowSecond)[q] = s.a2;
        }
    }
    add_stripe(n, stripes);
}

void launch_rescore(float const* L, float const* R, int nL, int nR, Cand2 const* partL, int nPartL, int ldqL,
                    Cand2 const* partR, int nPartR, int ldqR, float const* normL, float const* normR,
                    unsigned int const* maxNormBits, float winRel, int topK, float margin, float minScore,
                    float minRatio, int* colMax, int* rowMax, int* colSecond, int* rowSecond, int* stripes,
                    hipStream_t st)
{
    int const n = nL + nR;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shortlist_finish<Cand2>, dim3((n + 31) / 32), dim3(256), 0, st, L, R, nL, nR, partL, nPartL,
                       ldqL, 0, partR, nPartR, ldqR, 0, normL, normR, maxNormBits, winRel, topK, margin, minScore,
                       minRatio, colMax, rowMax, colSecond, rowSecond, stripes);
}

void launch_rescore_packed(float const* L, float const* R, int nL, int nR, CandP const* partL, int nPartL, int ldqL,
                           int spanL, CandP const* partR, int nPartR, int ldqR, int spanR, float const* normL,
                           float const* normR, unsigned int const* maxNormBits, int topK, float margin,
                           float minScore, float minRatio, int* colMax, int* rowMax, int* colSecond, int* rowSecond,
                           int* stripes, hipStream_t st)
{
    int const n = nL + nR;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shortlist_finish<CandP>, dim3((n + 31) / 32), dim3(256), 0, st, L, R, nL, nR, partL, nPartL,
                       ldqL, spanL, partR, nPartR, ldqR, spanR, normL, normR, maxNormBits, 0.0f, topK, margin,
                       minScore, minRatio, colMax, rowMax, colSecond, rowSecond, stripes);
}

// ------------------------------------------------------------------------------ batch
// Query-space rows; padding rows decide -1.  FP32 images: exact scores of every candidate.
__device__ __forceinline__ void batch_finish(BatchDev const& b, Cand2 const* __restrict__ part, float winRel,
                                             int topK, float margin, float minScore, float minRatio,
                                             int* __restrict__ sel, int* __restrict__ stripes)
{
    __shared__ int slots[32][kRescoreSlots];
    __shared__ __attribute__((aligned(16))) float qrow[32][kDim];
    int const sub = threadIdx.x & 7;
    int const g = blockIdx.x * 32 + (threadIdx.x >> 3);
    int n = 0;
    if (g < b.qRows) {
        BatchSeg const sg = b.seg[b.segOfBlock[g / kBatchRowAlign]];
        int const q = g - sg.qOff;
        int d = -1;
        if (q < sg.n && sg.nO > 0 && sg.nParts > 0) {
            int const row = b.imgBase[sg.img] + q;
            size_t const cb = (size_t)b.imgBase[sg.oImg];
            Top2x const s = finish_query<false>(part + ((size_t)sg.partOff + q) * 2, 2 * (size_t)sg.ld, sg.nParts, 0,
                                                b.norm[row], __uint_as_float(b.maxNorm[sg.oImg]), winRel, topK,
                                                margin, b.img + (size_t)row * kDim, b.img + cb * kDim, nullptr,
                                                nullptr, sub, slots[threadIdx.x >> 3], qrow[threadIdx.x >> 3], n);
            d = decide_exact(Top2{s.b1, s.b2, s.a1}, minScore, minRatio);
        }
        if (sub == 0) sel[g] = d;
    }
    add_stripe(n, stripes);
}

// uint8 images: certified integer scores (score_u8_certified) and the certified decision.
// A query whose decision the bound leaves open (near ties and thresholds; certify = false:
// every query) is appended to `open` (count in *nOpen) and decided by
// k_batch_rescore_open.  Blocks run in XCD order (xcd_block), so the queries of one
// segment, which read the same candidate image, share an L2.  No exact code here: few
// registers, eight waves per SIMD to cover the chain of dependent loads (segment, lists,
// candidate rows).
__global__ __launch_bounds__(256, 8) void k_batch_rescore_u8(BatchDev b, Cand2 const* __restrict__ part, float winRel,
                                                             int topK, float margin, float minScore, float minRatio,
                                                             bool certify, int* __restrict__ sel,
                                                             int* __restrict__ open, int* __restrict__ nOpen,
                                                             int* __restrict__ stripes)
{
    __shared__ int slots[32][kRescoreSlots];
    int const sub = threadIdx.x & 7;
    int const grp = threadIdx.x >> 3;
    int const g = xcd_block(blockIdx.x, gridDim.x) * 32 + grp;
    int n = 0;
    if (g < b.qRows) {
        BatchSeg const sg = b.seg[b.segOfBlock[g / kBatchRowAlign]];
        int const q = g - sg.qOff;
        int d = -1;
        if (q < sg.n && sg.nO > 0 && sg.nParts > 0) {
            int const row = b.imgBase[sg.img] + q;
            size_t const cb = (size_t)b.imgBase[sg.oImg];
            uint4 const qv = reinterpret_cast<uint4 const*>(b.u8 + (size_t)row * kDim)[sub];
            float const ra = b.denom[row];
            int const total = select_candidates(part + ((size_t)sg.partOff + q) * 2, 2 * (size_t)sg.ld, sg.nParts, 0,
                                                b.norm[row], __uint_as_float(b.maxNorm[sg.oImg]), winRel, topK,
                                                margin, sub, slots[grp]);
            n = (sub == 0) ? total : 0;
            d = kFlag;
            if (certify) {
                Top2x const s = score_u8_certified(qv, ra, b.u8 + cb * kDim, b.denom + cb, slots[grp], total, sub);
                d = decide_certified(Top2{s.b1, s.b2, s.a1}, kCertU8 * fmaxf(s.b1, 0.0f) + kEpsAbs, minScore,
                                     minRatio);
            }
            if (d == kFlag && sub == 0) open[atomicAdd(nOpen, 1)] = g;
        }
        if (sub == 0) sel[g] = d;
    }
    add_stripe(n, stripes);
}

// The learned lists of uint8 images (topK in [1, kSelK]) in two kernels.
//
// k_batch_select_lane, one lane per query (a work-group's 256 queries lie in one segment,
// kBatchRowAlign = 256), one pass over the query's lists: the kSelK largest list values
// with their places in the lists (sorted, multiplicity counted), then
// thr = max(m - margin, the topK-th largest) with m the largest -- the eight-lane
// selection's threshold in both of its cases (at most topK values within the margin:
// m - margin, which is then >= the topK-th largest; else the topK-th largest, which is
// then >= m - margin).  The candidates at or above thr are a prefix of the kept ones unless
// all kSelK kept values are >= thr (ties at the threshold may continue past them): such a
// query, like every query when certify is off, goes to the open list (exact path).  The
// prefix goes to slots[g][0 .. cnt) with cnt[g].
//
// k_batch_score8, eight lanes per query: lane k loads bytes 16k .. 16k+15 of the query row
// and of every selected candidate row (coalesced 128-byte rows), forms the eight partial
// integer dot products, and a transposed reduction (three exchange steps over the group)
// leaves lane k with candidate k's whole product.  Score fl(fl(ra rb) I) as the eight-lane
// form forms it, the top-2 (score, then smaller index) merged over the group, the
// certified decision.  Same candidates, scores, order and decision as the eight-lane form.
constexpr int kSelK = 8;

// The learned lists' values are integer code scores (|s| < 2^15): key = s * 64 + e with e
// the candidate's place in the query's lists (2 * list + 0/1, < 64) orders by value (then
// place) and fits an int32; INT_MIN marks no candidate.
__device__ __forceinline__ void topk_insert_key(int (&t)[kSelK], int x)
{
#pragma unroll
    for (int i = 0; i < kSelK; ++i) {  // t descending: keep the larger, carry the smaller
        int const hi = max(t[i], x);
        x = min(t[i], x);
        t[i] = hi;
    }
}

__device__ __forceinline__ float key_value(int key) { return key == INT_MIN ? -INFINITY : (float)(key >> 6); }

__global__ __launch_bounds__(256) void k_batch_select_lane(BatchDev b, Cand2 const* __restrict__ part, int topK,
                                                           float margin, bool certify, int* __restrict__ sel,
                                                           int* __restrict__ cnt, int4* __restrict__ slots,
                                                           int* __restrict__ open, int* __restrict__ nOpen,
                                                           int* __restrict__ stripes)
{
    static_assert(kBatchRowAlign == 256, "a work-group's queries lie in one segment");
    static_assert(2 * 2 * kShortMaxSplits <= 64, "a candidate's place fits 6 bits");
    BatchBlock const bi = b.blkInfo[xcd_block(blockIdx.x, gridDim.x)];  // one scalar load
    int const g = bi.blk * 256 + (int)threadIdx.x;
    int n = 0;
    if (g < b.qRows) {
        int const q = bi.q0 + (int)threadIdx.x;
        int c = 0;
        int d = -1;
        int jj[kSelK];
#pragma unroll
        for (int i = 0; i < kSelK; ++i) jj[i] = -1;
        if (q < bi.n && bi.nO > 0 && bi.nParts > 0) {
            Cand2 const* __restrict__ lp = part + ((size_t)bi.partOff + q) * 2;
            size_t const ld2 = 2 * (size_t)bi.ld;
            int const nLists = 2 * bi.nParts;
            int t[kSelK];
#pragma unroll
            for (int i = 0; i < kSelK; ++i) t[i] = INT_MIN;
            int total = 0;  // listed candidates
            for (int e0 = 0; e0 < nLists; e0 += 4) {
                Cand2 cs[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {  // four lists in flight
                    int const e = e0 + u;
                    cs[u] = Cand2{0.0f, -1, 0.0f, -1};
                    if (e < nLists) cs[u] = lp[(size_t)(e >> 1) * ld2 + (e & 1)];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const e = e0 + u;
                    if (cs[u].j1 >= 0) {
                        topk_insert_key(t, (int)cs[u].v1 * 64 + 2 * e);
                        ++total;
                    }
                    if (cs[u].j2 >= 0) {
                        topk_insert_key(t, (int)cs[u].v2 * 64 + 2 * e + 1);
                        ++total;
                    }
                }
            }
            float kth = key_value(t[0]);
#pragma unroll
            for (int i = 1; i < kSelK; ++i) kth = (i == topK - 1) ? key_value(t[i]) : kth;
            float const thr = fmaxf(key_value(t[0]) - margin, kth);
#pragma unroll
            for (int i = 0; i < kSelK; ++i) c += (t[i] != INT_MIN && key_value(t[i]) >= thr) ? 1 : 0;
            n = c;  // candidates scored here (an overflowing query's further ties uncounted)
            bool const overflow = c == kSelK && total > kSelK;
            if (!certify || overflow) {
                d = kFlag;
                c = 0;
                open[atomicAdd(nOpen, 1)] = g;
            } else if (c > 0) {
                d = kFlag;  // decided by k_batch_score8
#pragma unroll
                for (int i = 0; i < kSelK; ++i)  // the kept candidates' indices (cache hits)
                    if (i < c) {
                        int const e = (t[i] & 63) >> 1;
                        Cand2 const* cp = lp + (size_t)(e >> 1) * ld2 + (e & 1);
                        jj[i] = (t[i] & 1) ? cp->j2 : cp->j1;
                    }
            }
        }
        sel[g] = d;
        cnt[g] = c;
        // slots past cnt repeat slot 0 (read, not used, by k_batch_score8)
#pragma unroll
        for (int i = 1; i < kSelK; ++i) jj[i] = i < c ? jj[i] : jj[0];
        slots[2 * (size_t)g] = make_int4(jj[0], jj[1], jj[2], jj[3]);
        slots[2 * (size_t)g + 1] = make_int4(jj[4], jj[5], jj[6], jj[7]);
    }
    add_stripe(n, stripes);
}

// DPP exchanges inside a group of eight lanes (lane k = lane & 7):
//   kDppXor1 / kDppXor2: quad_perm [1,0,3,2] / [2,3,0,1] (lane k ^ 1, k ^ 2);
//   kDppMirror8: row_half_mirror (lane 7 - k: the other half of the group).
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;
constexpr int kDppMirror8 = 0x141;

template <int kCtrl>
__device__ __forceinline__ unsigned dpp_u(unsigned v)
{
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, kCtrl, 0xF, 0xF, false);
}

// Best score, its candidate and the second best score (the decision needs no second
// index): the merge keeps the larger best (equal: the smaller index, -1 as the largest) and
// the second largest of the four values.
struct Top2s {
    float b1, b2;
    int a1;
};

__device__ __forceinline__ Top2s top2s_merge(Top2s a, Top2s b)
{
    bool const takeB = (b.b1 > a.b1) || (b.b1 == a.b1 && (unsigned)b.a1 < (unsigned)a.a1);
    return Top2s{fmaxf(a.b1, b.b1), fmaxf(fminf(a.b1, b.b1), fmaxf(a.b2, b.b2)), takeB ? b.a1 : a.a1};
}

template <int kCtrl>
__device__ __forceinline__ Top2s top2s_dpp(Top2s s)
{
    return Top2s{__uint_as_float(dpp_u<kCtrl>(__float_as_uint(s.b1))),
                 __uint_as_float(dpp_u<kCtrl>(__float_as_uint(s.b2))), (int)dpp_u<kCtrl>((unsigned)s.a1)};
}

__global__ __launch_bounds__(256) void k_batch_score8(BatchDev b, int const* __restrict__ cnt,
                                                      int4 const* __restrict__ slots, float minScore, float minRatio,
                                                      int* __restrict__ sel, int* __restrict__ open,
                                                      int* __restrict__ nOpen)
{
    int const k = threadIdx.x & 7;
    int const lb = xcd_block(blockIdx.x, gridDim.x);  // 32-query blocks, in blkInfo's order
    BatchBlock const bi = b.blkInfo[lb >> 3];         // one scalar load
    int const sub = (lb & 7) * 32 + (int)(threadIdx.x >> 3);  // query inside the 256-row block
    int const g = bi.blk * 256 + sub;
    int const c = cnt[g];                    // uniform over the group
    if (c <= 0) return;                      // whole groups leave: the exchanges stay inside groups
    int const row = bi.rowBase + sub;
    uint8_t const* __restrict__ cRows = b.u8 + (size_t)bi.cBase * kDim;  // uniform
    float const* __restrict__ cDen = b.denom + bi.cBase;
    int4 const s0 = slots[2 * (size_t)g], s1 = slots[2 * (size_t)g + 1];
    int const js[kSelK] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
    uint4 const qv = reinterpret_cast<uint4 const*>(b.u8 + (size_t)row * kDim)[k];
    float const ra = b.denom[row];
    // chunk k of each slot's row (32-bit offsets from the uniform image base); slots past
    // cnt repeat slot 0's row (cache hits), and their sums only reach lanes that are not used
    uint4 cv[kSelK];
#pragma unroll
    for (int i = 0; i < kSelK; ++i)
        cv[i] = *reinterpret_cast<uint4 const*>(cRows + ((unsigned)js[i] * (unsigned)kDim + 16u * k));
    int const jk = k < c ? js[k] : -1;  // the candidate this lane finishes
    float const rb = jk >= 0 ? cDen[jk] : 0.0f;
    unsigned p[kSelK];
#pragma unroll
    for (int i = 0; i < kSelK; ++i) {
        unsigned acc = __builtin_amdgcn_udot4(qv.x, cv[i].x, 0u, false);
        acc = __builtin_amdgcn_udot4(qv.y, cv[i].y, acc, false);
        acc = __builtin_amdgcn_udot4(qv.z, cv[i].z, acc, false);
        p[i] = __builtin_amdgcn_udot4(qv.w, cv[i].w, acc, false);
    }
    // transposed reduction: step 1 with lane 7 - k (other half: k < 4 keeps candidates
    // 0..3), step 2 with k ^ 2 (bit 1 of k picks the pair), step 3 with k ^ 1
    unsigned r4[4], r2[2];
    bool const lo4 = k < 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        unsigned const keep = lo4 ? p[i] : p[4 + i];
        unsigned const send = lo4 ? p[4 + i] : p[i];
        r4[i] = keep + dpp_u<kDppMirror8>(send);
    }
    bool const lo2 = (k & 2) == 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        unsigned const keep = lo2 ? r4[i] : r4[2 + i];
        unsigned const send = lo2 ? r4[2 + i] : r4[i];
        r2[i] = keep + dpp_u<kDppXor2>(send);
    }
    bool const lo1 = (k & 1) == 0;
    unsigned const keep = lo1 ? r2[0] : r2[1];
    unsigned const send = lo1 ? r2[1] : r2[0];
    unsigned const dot = keep + dpp_u<kDppXor1>(send);  // candidate k's product
    Top2s s = {kInit, kInit, -1};
    if (jk >= 0) {
        float const sc = mul_nc(mul_nc(ra, rb), (float)dot);
        if (sc == sc) s = Top2s{sc, kInit, jk};
    }
    s = top2s_merge(s, top2s_dpp<kDppXor1>(s));
    s = top2s_merge(s, top2s_dpp<kDppXor2>(s));
    s = top2s_merge(s, top2s_dpp<kDppMirror8>(s));
    if (k == 0) {
        int const d = decide_certified(Top2{s.b1, s.b2, s.a1}, kCertU8 * fmaxf(s.b1, 0.0f) + kEpsAbs, minScore,
                                       minRatio);
        sel[g] = d;
        if (d == kFlag) open[atomicAdd(nOpen, 1)] = g;
    }
}

// The open queries of k_batch_rescore_u8, eight lanes each (grid-stride over the list): the
// same candidate selection, exact scores (the oracle's arithmetic on the FP32 rows).
__global__ __launch_bounds__(256) void k_batch_rescore_open(BatchDev b, Cand2 const* __restrict__ part, float winRel,
                                                            int topK, float margin, float minScore, float minRatio,
                                                            int* __restrict__ sel, int const* __restrict__ open,
                                                            int const* __restrict__ nOpen)
{
    __shared__ int slots[32][kRescoreSlots];
    int const sub = threadIdx.x & 7;
    int const grp = threadIdx.x >> 3;
    int const cnt = *nOpen;
    for (int w = blockIdx.x * 32 + grp; w < cnt; w += gridDim.x * 32) {  // uniform over the group
        int const g = open[w];
        BatchSeg const sg = b.seg[b.segOfBlock[g / kBatchRowAlign]];
        int const q = g - sg.qOff;
        int const row = b.imgBase[sg.img] + q;
        size_t const cb = (size_t)b.imgBase[sg.oImg];
        int const total = select_candidates(part + ((size_t)sg.partOff + q) * 2, 2 * (size_t)sg.ld, sg.nParts, 0,
                                            b.norm[row], __uint_as_float(b.maxNorm[sg.oImg]), winRel, topK, margin,
                                            sub, slots[grp]);
        int scored = 0;
        Top2x const s = score_exact<true>(b.img + (size_t)row * kDim, nullptr, b.u8 + cb * kDim, b.denom + cb,
                                          slots[grp], total, sub, scored);
        if (sub == 0) sel[g] = decide_exact(Top2{s.b1, s.b2, s.a1}, minScore, minRatio);
    }
}

__global__ __launch_bounds__(256) void k_batch_shortlist_finish(BatchDev b, Cand2 const* __restrict__ part,
                                                                float winRel, int topK, float margin, float minScore,
                                                                float minRatio, int* __restrict__ sel,
                                                                int* __restrict__ stripes)
{
    batch_finish(b, part, winRel, topK, margin, minScore, minRatio, sel, stripes);
}

void launch_batch_rescore(BatchDev const& b, Cand2 const* part, float winRel, int topK, float margin, float minScore,
                          float minRatio, bool certify, int* sel, int* open, int* nOpen, int* selCnt, int* selSlots,
                          int* stripes, hipStream_t st)
{
    if (b.qRows <= 0) return;
    if (b.u8 && topK >= 1 && topK <= kSelK && selCnt && selSlots) {
        hipLaunchKernelGGL(k_batch_select_lane, dim3((b.qRows + 255) / 256), dim3(256), 0, st, b, part, topK, margin,
                           certify, sel, selCnt, reinterpret_cast<int4*>(selSlots), open, nOpen, stripes);
        hipLaunchKernelGGL(k_batch_score8, dim3((b.qRows + 31) / 32), dim3(256), 0, st, b, selCnt,
                           reinterpret_cast<int4 const*>(selSlots), minScore, minRatio, sel, open, nOpen);
        hipLaunchKernelGGL(k_batch_rescore_open, dim3(256), dim3(256), 0, st, b, part, winRel, topK, margin, minScore,
                           minRatio, sel, open, nOpen);
    } else if (b.u8) {
        hipLaunchKernelGGL(k_batch_rescore_u8, dim3((b.qRows + 31) / 32), dim3(256), 0, st, b, part, winRel, topK,
                           margin, minScore, minRatio, certify, sel, open, nOpen, stripes);
        hipLaunchKernelGGL(k_batch_rescore_open, dim3(256), dim3(256), 0, st, b, part, winRel, topK, margin, minScore,
                           minRatio, sel, open, nOpen);
    } else {
        hipLaunchKernelGGL(k_batch_shortlist_finish, dim3((b.qRows + 31) / 32), dim3(256), 0, st, b, part, winRel,
                           topK, margin, minScore, minRatio, sel, stripes);
    }
}

}  // namespace v3dm
