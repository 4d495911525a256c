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

// The learned lists of uint8 images (topK in [1, kSelK]): k_batch_rescore_learned, one
// work-group per 64 queries of one segment (BatchBlock table, pair by pair).
//
// Selection, one lane per query (wave 0), one pass over the query's lists with eight lists in
// flight (the same 90 registers as four; sixteen would cost a work-group per CU): the kSelK largest list values with their candidates (sorted, multiplicity
// counted), then thr = max(m - margin, the topK-th largest) with m the largest -- the
// eight-lane selection's threshold in both of its cases (at most topK values within the
// margin: m - margin, which is then >= the topK-th largest; else the topK-th largest, which
// is then >= m - margin).  The candidates at or above thr are a prefix of the kept ones
// unless all kSelK kept values are >= thr (ties at the threshold may continue past them):
// such a query, like every query when certify is off, goes to the open list (exact path).
// The prefix goes to LDS.
//
// Scoring, four lanes per query (a quad): lane k holds bytes 32k .. 32k+31 of the query row
// and of every selected candidate row, eight partial integer products per lane; a
// transposed reduction in two quad exchanges (DPP) leaves lane k with the whole products of
// candidates 2k and 2k+1; score fl(fl(ra rb) I) as the eight-lane form forms it, the top-2
// (score, then smaller index) merged over the quad, the certified decision (eps = 2^-21
// best: the integer product is exact, the two roundings give |s~ - s| <= 5.01 2^-24 s~).
// Same candidates, scores, order and decision as the eight-lane form, in three dependent
// memory round trips (block table, lists, rows).
constexpr int kSelK = 8;

__device__ __forceinline__ void topk_insert_pair(float (&t)[kSelK], int (&jj)[kSelK], float x, int jx)
{
#pragma unroll
    for (int i = 0; i < kSelK; ++i) {  // t descending; a new value goes behind equal ones
        bool const sw = x > t[i];
        float const tv = sw ? x : t[i];
        int const tj = sw ? jx : jj[i];
        x = sw ? t[i] : x;
        jx = sw ? jj[i] : jx;
        t[i] = tv;
        jj[i] = tj;
    }
}

// DPP exchanges inside a quad: kDppXor1 / kDppXor2 = quad_perm [1,0,3,2] / [2,3,0,1]
constexpr int kDppXor1 = 0xB1;
constexpr int kDppXor2 = 0x4E;

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

__global__ __launch_bounds__(256) void k_batch_rescore_learned(BatchDev b, Cand2 const* __restrict__ part, int topK,
                                                               float margin, float minScore, float minRatio,
                                                               bool certify, int* __restrict__ sel,
                                                               int* __restrict__ open, int* __restrict__ nOpen,
                                                               int* __restrict__ stripes)
{
    static_assert(kBatchRowAlign == 256, "a work-group's queries lie in one segment");
    __shared__ int sCnt[64];
    __shared__ int sJ[64][kSelK];
    int const tid = threadIdx.x;
    int const lb = xcd_block(blockIdx.x, gridDim.x);  // 64-query blocks, in blkInfo's order
    BatchBlock const bi = b.blkInfo[lb >> 2];         // one scalar load
    int const base = (lb & 3) * 64;                   // first query of the work-group in its block
    // the scoring's query rows, loaded while wave 0 selects: lane k of quad qi holds bytes
    // 32k .. 32k+31 of query row base + qi (rows of the padded image, always in range)
    int const k = tid & 3;
    int const qi = tid >> 2;  // query base + qi
    int const row = bi.rowBase + base + qi;
    uint4 const* qp = reinterpret_cast<uint4 const*>(b.u8 + (size_t)row * kDim + 32 * k);
    uint4 const q0 = qp[0], q1 = qp[1];
    float const ra = b.denom[row];
    int n = 0;
    if (tid < 64) {
        // ---- selection: query base + tid
        int const q = bi.q0 + base + tid;
        int c = 0;
        int d = -1;
        int jj[kSelK];
#pragma unroll
        for (int i = 0; i < kSelK; ++i) jj[i] = -1;
        if (q < bi.n && bi.nO > 0 && bi.nParts > 0) {
            Cand2 const* __restrict__ lp = part + ((size_t)bi.partOff + q) * 2;
            size_t const ld2 = 2 * (size_t)bi.ld;
            int const nLists = 2 * bi.nParts;
            float t[kSelK];
#pragma unroll
            for (int i = 0; i < kSelK; ++i) t[i] = -INFINITY;
            int total = 0;  // listed candidates
            for (int e0 = 0; e0 < nLists; e0 += 8) {
                Cand2 cs[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {  // eight lists in flight
                    int const e = e0 + u;
                    cs[u] = Cand2{-INFINITY, -1, -INFINITY, -1};
                    if (e < nLists) cs[u] = lp[(size_t)(e >> 1) * ld2 + (e & 1)];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (cs[u].j1 >= 0) {
                        topk_insert_pair(t, jj, cs[u].v1, cs[u].j1);
                        ++total;
                    }
                    if (cs[u].j2 >= 0) {
                        topk_insert_pair(t, jj, cs[u].v2, cs[u].j2);
                        ++total;
                    }
                }
            }
            float kth = t[0];
#pragma unroll
            for (int i = 1; i < kSelK; ++i) kth = (i == topK - 1) ? t[i] : kth;
            float const thr = fmaxf(t[0] - margin, kth);
#pragma unroll
            for (int i = 0; i < kSelK; ++i) c += (jj[i] >= 0 && t[i] >= thr) ? 1 : 0;
            n = c;  // candidates scored here (an overflowing query's further ties uncounted)
            if (!certify || (c == kSelK && total > kSelK)) {
                d = kFlag;
                c = 0;
                open[atomicAdd(nOpen, 1)] = bi.blk * 256 + base + tid;
            } else if (c > 0) {
                d = kFlag;  // decided by the scoring below
            }
        }
        if (c == 0) sel[bi.blk * 256 + base + tid] = d;
        // slots past c repeat slot 0 (read, not used, by the scoring)
        sCnt[tid] = c;
#pragma unroll
        for (int i = 0; i < kSelK; ++i) sJ[tid][i] = i < c ? jj[i] : jj[0];
    }
    __syncthreads();  // the selections in LDS
    add_stripe(n, stripes);
    // ---- scoring: four lanes per query
    int const c = sCnt[qi];   // uniform over the quad
    if (c <= 0) return;       // whole quads leave: the exchanges stay inside quads
    int const g = bi.blk * 256 + base + qi;
    uint8_t const* __restrict__ cRows = b.u8 + (size_t)bi.cBase * kDim;  // uniform
    float const* __restrict__ cDen = b.denom + bi.cBase;
    int js[kSelK];
#pragma unroll
    for (int i = 0; i < kSelK; ++i) js[i] = sJ[qi][i];
    uint4 cv[kSelK][2];
#pragma unroll
    for (int i = 0; i < kSelK; ++i) {  // only the c selected rows are fetched (the rest: unused)
        cv[i][0] = cv[i][1] = make_uint4(0u, 0u, 0u, 0u);
        if (i < c) {
            uint4 const* cp = reinterpret_cast<uint4 const*>(cRows + ((unsigned)js[i] * (unsigned)kDim + 32u * k));
            cv[i][0] = cp[0];
            cv[i][1] = cp[1];
        }
    }
    int const ja = 2 * k < c ? js[2 * k] : -1;  // the candidates this lane finishes
    int const jb = 2 * k + 1 < c ? js[2 * k + 1] : -1;
    float const rba = ja >= 0 ? cDen[ja] : 0.0f;
    float const rbb = jb >= 0 ? cDen[jb] : 0.0f;
    unsigned p[kSelK];
#pragma unroll
    for (int i = 0; i < kSelK; ++i) {
        unsigned acc = __builtin_amdgcn_udot4(q0.x, cv[i][0].x, 0u, false);
        acc = __builtin_amdgcn_udot4(q0.y, cv[i][0].y, acc, false);
        acc = __builtin_amdgcn_udot4(q0.z, cv[i][0].z, acc, false);
        acc = __builtin_amdgcn_udot4(q0.w, cv[i][0].w, acc, false);
        acc = __builtin_amdgcn_udot4(q1.x, cv[i][1].x, acc, false);
        acc = __builtin_amdgcn_udot4(q1.y, cv[i][1].y, acc, false);
        acc = __builtin_amdgcn_udot4(q1.z, cv[i][1].z, acc, false);
        p[i] = __builtin_amdgcn_udot4(q1.w, cv[i][1].w, acc, false);
    }
    // step 1 with lane k ^ 2 (bit 1 of k: candidates 0..3 or 4..7), step 2 with k ^ 1
    // (bit 0: the first or the second pair of those): lane k ends with candidates 2k, 2k+1
    unsigned r4[4];
    bool const lo2 = (k & 2) == 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        unsigned const keep = lo2 ? p[i] : p[4 + i];
        unsigned const send = lo2 ? p[4 + i] : p[i];
        r4[i] = keep + dpp_u<kDppXor2>(send);
    }
    bool const lo1 = (k & 1) == 0;
    unsigned dot[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        unsigned const keep = lo1 ? r4[i] : r4[2 + i];
        unsigned const send = lo1 ? r4[2 + i] : r4[i];
        dot[i] = keep + dpp_u<kDppXor1>(send);
    }
    Top2s s = {kInit, kInit, -1};
    if (ja >= 0) {
        float const sc = mul_nc(mul_nc(ra, rba), (float)dot[0]);
        if (sc == sc) s = Top2s{sc, kInit, ja};
    }
    if (jb >= 0) {
        float const sc = mul_nc(mul_nc(ra, rbb), (float)dot[1]);
        if (sc == sc) s = top2s_merge(s, Top2s{sc, kInit, jb});
    }
    s = top2s_merge(s, top2s_dpp<kDppXor1>(s));
    s = top2s_merge(s, top2s_dpp<kDppXor2>(s));
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
                          float minRatio, bool certify, int* sel, int* open, int* nOpen, int* stripes, hipStream_t st)
{
    if (b.qRows <= 0) return;
    if (b.u8 && topK >= 1 && topK <= kSelK) {
        hipLaunchKernelGGL(k_batch_rescore_learned, dim3((b.qRows + 63) / 64), dim3(256), 0, st, b, part, topK, margin,
                           minScore, minRatio, certify, sel, open, nOpen, stripes);
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
