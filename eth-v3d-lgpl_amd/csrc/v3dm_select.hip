// Selection stage shared by every precision form: per-feature norms, merging of partial
// top-2 states, certified decisions, exact FP64 re-check of open queries, mutual check
// and ordered compaction of the matches.
//
// Reference: the selection kernels and host loop of V3D/CUDA/v3d_cudamatching.cu
// (cu:133-210 column/row top-2 with the ratio test, cu:33-85 without it, cu:546-551 the
// mutual check).
#include "v3dm_device.h"
#include "v3dm_kernels.h"

namespace v3dm {

// ------------------------------------------------------------------------------ norms
// norm[r] = ||x_r||_2 (FP32) and the maximum over rows.  Norms are >= 0 or NaN, so the
// max can be taken on the uint bits; a NaN wins, which later flags every query.
// One work-group handles 64 rows (16 per wave) and issues one atomic.
__global__ __launch_bounds__(256) void k_norms(float const* __restrict__ X, int n, float* __restrict__ norm,
                                               unsigned int* __restrict__ maxNormBits)
{
    __shared__ unsigned int wmax[4];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    unsigned int mx = 0;
    for (int k = 0; k < 16; ++k) {
        int const row = blockIdx.x * 64 + wave * 16 + k;
        if (row >= n) break;
        float2 const v = reinterpret_cast<float2 const*>(X + (size_t)row * kDim)[lane];
        float s = v.x * v.x + v.y * v.y;
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        float const nr = sqrtf(s);
        if (lane == 0) norm[row] = nr;
        mx = max(mx, __float_as_uint(nr));
    }
    if (lane == 0) wmax[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int const m = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
        atomicMax(maxNormBits, m);
    }
}

void launch_norms(float const* X, int n, float* norm, unsigned int* maxNormBits, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_norms, dim3((n + 63) / 64), dim3(256), 0, st, X, n, norm, maxNormBits);
}

// ------------------------------------------------------------------------------ merge + decide
// 8 lanes per query (lane = 8*query + sub); each lane merges every 8th partial, then the
// 8 lanes combine with xor shuffles.  Queries [0, nL) are left features, [nL, nL+nR)
// right features.
__global__ __launch_bounds__(256) void k_merge_decide(Top2 const* __restrict__ colPart, int nColParts, int ldc,
                                                      Top2 const* __restrict__ rowPart, int nRowParts, int ldr,
                                                      int nL, int nR, float const* __restrict__ normL,
                                                      float const* __restrict__ normR,
                                                      unsigned int const* __restrict__ maxNormBits, float epsRel,
                                                      float minScore, float minRatio, int* __restrict__ colMax,
                                                      int* __restrict__ rowMax, int* __restrict__ flagList,
                                                      int flagOffR, int* __restrict__ counters)
{
    int const sub = threadIdx.x & 7;
    int const g = (int)((blockIdx.x * 256u + threadIdx.x) >> 3);
    bool const valid = g < nL + nR;
    int const gq = valid ? g : 0;
    bool const left = gq < nL;
    int const q = left ? gq : gq - nL;
    Top2 const* part = left ? colPart : rowPart;
    int const np = valid ? (left ? nColParts : nRowParts) : 0;
    size_t const ld = left ? (size_t)ldc : (size_t)ldr;
    Top2 s = top2_init();
#pragma unroll 4
    for (int p = sub; p < np; p += 8) s = top2_merge(s, part[(size_t)p * ld + q]);
    s = top2_merge(s, top2_shfl_xor(s, 1));
    s = top2_merge(s, top2_shfl_xor(s, 2));
    s = top2_merge(s, top2_shfl_xor(s, 4));
    if (!valid || sub != 0) return;
    float const nq = left ? normL[q] : normR[q];
    float const mo = __uint_as_float(maxNormBits[left ? 1 : 0]);
    float const eps = epsRel * nq * mo + kEpsAbs;
    int const d = (epsRel > 0.0f) ? decide_certified(s, eps, minScore, minRatio) : kFlag;
    if (d == kFlag) {
        int const slot = atomicAdd(&counters[left ? kCntFlag : kCntFlagR], 1);
        flagList[left ? slot : flagOffR + slot] = q;
    }
    (left ? colMax : rowMax)[q] = d;
}

void launch_merge_decide(Top2 const* colPart, int nColParts, int ldc, Top2 const* rowPart, int nRowParts,
                         int ldr, int nL, int nR, float const* normL, float const* normR,
                         unsigned int const* maxNormBits, float epsRel, float minScore, float minRatio,
                         int* colMax, int* rowMax, int* flagList, int flagOffR, int* counters, hipStream_t st)
{
    int const n = nL + nR;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_merge_decide, dim3((n + 31) / 32), dim3(256), 0, st, colPart, nColParts, ldc, rowPart,
                       nRowParts, ldr, nL, nR, normL, normR, maxNormBits, epsRel, minScore, minRatio, colMax,
                       rowMax, flagList, flagOffR, counters);
}

// ------------------------------------------------------------------------------ exact re-check
// Open queries are listed per direction (left queries at flagList[0..), right queries at
// flagList[flagOffR..)).  Pass 1: a work item is (group of up to kRecheckGroup queries of
// one direction, chunk of 256*per candidates); grid-stride over all items.  Each thread
// scores its candidate against every query of the group: fl32(sum_k double(q_k) *
// double(c_k)) in ascending k (products of floats are exact in FP64, so fused
// multiply-adds give the same sum), the group's queries being independent FP64 chains.
// The chunk's top-2 per query (first index on ties) goes to part[(row) * kRecheckMaxChunks
// + chunk].  Pass 2: one wave per open query merges its chunk partials and takes the IEEE
// FP32 decision.  The kernel boundary between the passes orders the partial writes.
constexpr int kRecheckGroup = 16;
constexpr int kRecheckSub = 64;     // candidates staged in LDS at a time
constexpr int kRecheckRow = 132;    // floats per staged candidate row (528 bytes: conflict-free b128 reads)

__global__ __launch_bounds__(256) void k_recheck_chunks(float const* __restrict__ L, float const* __restrict__ R,
                                                        int nL, int nR, int const* __restrict__ flagList,
                                                        int flagOffR, int const* __restrict__ counters,
                                                        int nChunksL, int perL, int nChunksR, int perR,
                                                        Top2* __restrict__ part, unsigned const* __restrict__ gbits,
                                                        int ldi)
{
    __shared__ __attribute__((aligned(16))) double qd[kRecheckGroup][kDim];
    __shared__ __attribute__((aligned(16))) float cand[kRecheckSub * kRecheckRow];
    __shared__ int qidx[kRecheckGroup];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const cntL = counters[kCntFlag];
    int const cntR = counters[kCntFlagR];
    int const grpL = (cntL + kRecheckGroup - 1) / kRecheckGroup;
    int const grpR = (cntR + kRecheckGroup - 1) / kRecheckGroup;
    int const itemsL = grpL * nChunksL;
    int const items = itemsL + grpR * nChunksR;
    for (int w = blockIdx.x; w < items; w += gridDim.x) {
        bool const left = w < itemsL;
        int const wl = left ? w : w - itemsL;
        int const nCh = left ? nChunksL : nChunksR;
        int const per = left ? perL : perR;
        int const g = wl / nCh;
        int const c = wl - g * nCh;
        int const cnt = left ? cntL : cntR;
        int const base = left ? 0 : flagOffR;
        int const nq = min(kRecheckGroup, cnt - g * kRecheckGroup);
        float const* Qs = left ? L : R;
        float const* C = left ? R : L;
        int const nC = left ? nR : nL;
        __syncthreads();
        if (tid < kRecheckGroup) qidx[tid] = (tid < nq) ? flagList[base + g * kRecheckGroup + tid] : -1;
        __syncthreads();
        {
            // 16 queries x 128 floats = 512 float4: two independent float4 loads per thread
            float4 v[2];
            int qi[2], k4[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                int const x = u * 256 + tid;
                qi[u] = x >> 5;
                k4[u] = x & 31;
                int const row = qidx[qi[u]];
                v[u] = (row >= 0) ? reinterpret_cast<float4 const*>(Qs + (size_t)row * kDim)[k4[u]]
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                qd[qi[u]][4 * k4[u] + 0] = v[u].x;
                qd[qi[u]][4 * k4[u] + 1] = v[u].y;
                qd[qi[u]][4 * k4[u] + 2] = v[u].z;
                qd[qi[u]][4 * k4[u] + 3] = v[u].w;
            }
        }
        // wave w scores queries 4w .. 4w+3 of the group; lane = candidate of the sub-tile
        Top2 s[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) s[j] = top2_init();
        int const cbeg = c * 256 * per;
        int const cend = min(cbeg + 256 * per, nC);
        for (int c0 = cbeg; c0 < cend; c0 += kRecheckSub) {
            __syncthreads();  // previous sub-tile consumed (and qd written, first time)
            {
                float4 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int const x = u * 256 + tid;
                    int const row = x >> 5, k4 = x & 31;
                    v[u] = (c0 + row < nC) ? reinterpret_cast<float4 const*>(C + (size_t)(c0 + row) * kDim)[k4]
                                           : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int const x = u * 256 + tid;
                    int const row = x >> 5, k4 = x & 31;
                    *reinterpret_cast<float4*>(&cand[row * kRecheckRow + 4 * k4]) = v[u];
                }
            }
            __syncthreads();
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            float const* cr = &cand[lane * kRecheckRow];
#pragma unroll 4
            for (int k4 = 0; k4 < 32; ++k4) {
                float4 const v = *reinterpret_cast<float4 const*>(cr + 4 * k4);
                double const x0 = v.x, x1 = v.y, x2 = v.z, x3 = v.w;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    double const* q = &qd[4 * wave + j][4 * k4];
                    acc[j] += q[0] * x0;
                    acc[j] += q[1] * x1;
                    acc[j] += q[2] * x2;
                    acc[j] += q[3] * x3;
                }
            }
            int const ci = c0 + lane;
            if (ci < nC) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float v = (float)acc[j];
                    int const qrow = qidx[4 * wave + j];
                    if (gbits && qrow >= 0) {
                        // guided matching: a removed (left, right) pair scores -1e20
                        int const li = left ? qrow : ci;
                        int const rj = left ? ci : qrow;
                        v = ((gbits[(size_t)(rj >> 5) * ldi + li] >> (rj & 31)) & 1u) ? -1e20f : v;
                    }
                    top2_push(s[j], v, ci);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            Top2 t = s[j];
            for (int o = 32; o > 0; o >>= 1) t = top2_merge(t, top2_shfl_xor(t, o));
            int const qg = 4 * wave + j;
            if (lane == 0 && qg < nq) part[(size_t)(base + g * kRecheckGroup + qg) * kRecheckMaxChunks + c] = t;
        }
    }
}

__global__ __launch_bounds__(256) void k_recheck_finish(int const* __restrict__ flagList, int flagOffR,
                                                        int const* __restrict__ counters, int nChunksL,
                                                        int nChunksR, Top2 const* __restrict__ part,
                                                        float minScore, float minRatio, int* __restrict__ colMax,
                                                        int* __restrict__ rowMax)
{
    int const lane = threadIdx.x & 63;
    int const cntL = counters[kCntFlag];
    int const cnt = cntL + counters[kCntFlagR];
    for (int e = blockIdx.x * 4 + (threadIdx.x >> 6); e < cnt; e += gridDim.x * 4) {
        bool const left = e < cntL;
        int const row = left ? e : flagOffR + (e - cntL);
        int const nCh = left ? nChunksL : nChunksR;
        Top2 s = top2_init();
        for (int c = lane; c < nCh; c += 64) s = top2_merge(s, part[(size_t)row * kRecheckMaxChunks + c]);
        for (int o = 32; o > 0; o >>= 1) s = top2_merge(s, top2_shfl_xor(s, o));
        if (lane == 0) (left ? colMax : rowMax)[flagList[row]] = decide_exact(s, minScore, minRatio);
    }
}

RecheckPlan plan_recheck(int nC)
{
    RecheckPlan p;
    p.per = (nC + 256 * kRecheckMaxChunks - 1) / (256 * kRecheckMaxChunks);
    if (p.per < 1) p.per = 1;
    p.nChunks = (nC + 256 * p.per - 1) / (256 * p.per);
    if (p.nChunks < 1) p.nChunks = 1;
    return p;
}

void launch_recheck_exact(float const* L, float const* R, int nL, int nR, int const* flagList, int flagOffR,
                          int const* counters, float minScore, float minRatio, int* colMax, int* rowMax,
                          Top2* part, int grid, hipStream_t st, unsigned const* gbits, int ldi)
{
    if (nL <= 0 || nR <= 0) return;
    RecheckPlan const pl = plan_recheck(nR);  // left queries scan the right features
    RecheckPlan const pr = plan_recheck(nL);
    hipLaunchKernelGGL(k_recheck_chunks, dim3(grid), dim3(256), 0, st, L, R, nL, nR, flagList, flagOffR, counters,
                       pl.nChunks, pl.per, pr.nChunks, pr.per, part, gbits, ldi);
    hipLaunchKernelGGL(k_recheck_finish, dim3(256), dim3(256), 0, st, flagList, flagOffR, counters, pl.nChunks,
                       pr.nChunks, part, minScore, minRatio, colMax, rowMax);
}

// ------------------------------------------------------------------------------ mutual + compaction
// (k, colMax[k]) is a match iff rowMax[colMax[k]] == k (cu:546-551).  One launch, ordered
// by a decoupled look-back: work-group b takes the 1024 left features [1024 b, 1024 b +
// 1024) (4 per thread), counts its matches, publishes the count (flag 1) in its status
// word, then adds up its predecessors' words from the nearest backwards (64 per step, one
// per lane of wave 0) until it reaches an inclusive prefix (flag 2), publishes its own
// inclusive prefix and writes its matches in ascending k from its exclusive prefix on, so
// the output is sorted by k.  A status word is valid for the call whose epoch it carries,
// so nothing is cleared between calls.  A work-group waits only for lower indices, which
// are dispatched before it; the wait is bounded (mw.spinLimit sleeps), and a work-group
// that runs out of it stores the epoch in counters[kCntMutualError] (its matches then land
// below their place, inside the buffer).  The last work-group also publishes the match,
// re-check and rescoring counts and clears the per-call counters.

// Left feature sampled as sample k of S (evenly spaced over [0, nL)).
__device__ __forceinline__ int sample_row(int k, int nL, int S) { return (int)(((long long)k * nL) / S); }

__device__ __forceinline__ unsigned long long mutual_word(unsigned epoch, unsigned flag, unsigned value)
{
    return ((unsigned long long)epoch << 32) | ((unsigned long long)flag << 30) | value;  // value < 2^30
}

__global__ __launch_bounds__(256) void k_mutual(int const* __restrict__ colMax, int const* __restrict__ rowMax,
                                                int nL, int nR, unsigned long long* __restrict__ status,
                                                unsigned epoch, int* __restrict__ stripes,
                                                int2* __restrict__ pairs, int* __restrict__ counters, SampleOut so,
                                                MutualWait mw)
{
    __shared__ int wsum[4];
    __shared__ int sBase;
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const b = blockIdx.x;
    int const i0 = b * kMutualSlice + 4 * tid;
    // all four colMax entries in one load, then all four rowMax gathers (indices clamped to
    // [0, max(nR - 1, 0)]: in bounds, rowMax holds at least one row per allocation) before
    // any test: two round trips
    int l[4];
    if (i0 + 3 < nL) {
        int4 const c = *reinterpret_cast<int4 const*>(colMax + i0);
        l[0] = c.x, l[1] = c.y, l[2] = c.z, l[3] = c.w;
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) l[u] = (i0 + u < nL) ? colMax[i0 + u] : -1;
    }
    int r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) r[u] = rowMax[min(max(l[u], 0), max(nR - 1, 0))];
    bool m[4];
    int cnt = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        m[u] = (l[u] >= 0) && (l[u] < nR) && (r[u] == i0 + u);
        cnt += m[u] ? 1 : 0;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int const t = __shfl_up(incl, o);
        incl += (lane >= o) ? t : 0;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        off += (w < wave) ? wsum[w] : 0;
        total += wsum[w];
    }
    if (wave == 0) {
        if (lane == 0)
            __hip_atomic_store(&status[b], mutual_word(epoch, b == 0 ? 2u : 1u, (unsigned)total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        int excl = 0;
        bool failed = false;
        int top = b - 1;  // nearest predecessor not yet added (uniform over the wave)
        unsigned spins = 0;
        while (top >= 0) {
            int const j = top - lane;
            unsigned long long const w =
                (j >= 0) ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : mutual_word(epoch, 2u, 0u);  // before work-group 0: an inclusive prefix of 0
            unsigned const fl = ((unsigned)(w >> 32) == (epoch ^ mw.readXor)) ? (unsigned)(w >> 30) & 3u : 0u;
            unsigned long long const ready = __ballot(fl != 0u);
            unsigned long long const inclusive = __ballot(fl == 2u);
            int const stop = inclusive ? __ffsll((long long)inclusive) - 1 : 63;  // last lane to add
            unsigned long long const need = (2ull << stop) - 1ull;                 // lanes 0 .. stop
            if ((ready & need) == need) {
                int v = (lane <= stop) ? (int)(w & 0x3fffffffull) : 0;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                excl += v;
                if (inclusive) break;
                top -= 64;
            } else {
                if (++spins > mw.spinLimit) {
                    failed = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (lane == 0) {
            if (b > 0)
                __hip_atomic_store(&status[b], mutual_word(epoch, 2u, (unsigned)(excl + total)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            if (failed) counters[kCntMutualError] = (int)epoch;
            sBase = excl;
        }
    }
    __syncthreads();
    int pos = sBase + off + incl - cnt;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (m[u]) pairs[pos++] = make_int2(i0 + u, l[u]);
    // work-group 0 (off the look-back chain by now): the sampled learned decisions
    if (b == 0 && tid < so.S) so.out[tid] = colMax[sample_row(tid, nL, so.S)];  // vector stores
    if (b != (int)gridDim.x - 1) return;
    // the last work-group: counts of the call
    if (wave == 0) {
        int rescored = 0;
        if (tid < kStripes) {
            rescored = stripes[tid * kStripeStride];
            stripes[tid * kStripeStride] = 0;
        }
        for (int o = 32; o > 0; o >>= 1) rescored += __shfl_xor(rescored, o);
        if (tid == 0) {
            counters[kCntPairs] = sBase + total;
            counters[kCntFlagDone] = counters[kCntFlag] + counters[kCntFlagR];
            counters[kCntFlag] = 0;
            counters[kCntFlagR] = 0;
            counters[kCntRescoredDone] = rescored;
        }
    }
}

void launch_mutual_compact(int const* colMax, int const* rowMax, int nL, int nR, unsigned long long* status,
                           unsigned epoch, int* stripes, int2* pairs, int* counters, hipStream_t st,
                           SampleOut const& so, MutualWait const& mw)
{
    int const nb = (nL + kMutualSlice - 1) / kMutualSlice;
    hipLaunchKernelGGL(k_mutual, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, colMax, rowMax, nL, nR, status, epoch,
                       stripes, pairs, counters, so, mw);
}

// ------------------------------------------------------------------------------ uint8 dequantisation
// One wave per descriptor, 2 components per lane (dequant_u8_wave).
__global__ __launch_bounds__(256) void k_dequant_u8(uint8_t const* __restrict__ q, int n, float* __restrict__ out)
{
    int const lane = threadIdx.x & 63;
    int const row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    uchar2 const v = reinterpret_cast<uchar2 const*>(q + (size_t)row * kDim)[lane];
    reinterpret_cast<float2*>(out + (size_t)row * kDim)[lane] = dequant_u8_wave(v);
}

void launch_dequant_u8(uint8_t const* q, int n, float* out, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dequant_u8, dim3((n + 3) / 4), dim3(256), 0, st, q, n, out);
}

// ------------------------------------------------------------------------------ sampled check
// A shortlist result is checked on S evenly spaced left features (S <= 128): their rows and
// norms are gathered into a 128-row block (zero rows past S) and decided by the plain exact
// path; the comparison counts the sampled features the exact rule accepts, those whose
// shortlist decision differs from it (missed), and those the shortlist accepts wrongly.
__global__ __launch_bounds__(256) void k_sample_gather(float const* __restrict__ L, float const* __restrict__ normL,
                                                       int nL, int S, float* __restrict__ Ls,
                                                       float* __restrict__ normS, int* __restrict__ idx,
                                                       int* __restrict__ counters)
{
    if (blockIdx.x == 0 && threadIdx.x < kNumCounters) counters[threadIdx.x] = 0;  // the check's own counters
    int const lane = threadIdx.x & 63;
    int const k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= kSampleRows) return;
    int const i = (k < S) ? sample_row(k, nL, S) : -1;
    float2 const v = (i >= 0) ? reinterpret_cast<float2 const*>(L + (size_t)i * kDim)[lane] : make_float2(0.f, 0.f);
    reinterpret_cast<float2*>(Ls + (size_t)k * kDim)[lane] = v;
    if (lane == 0) {
        normS[k] = (i >= 0) ? normL[i] : 0.0f;
        idx[k] = i;
    }
}

void launch_sample_gather(float const* L, float const* normL, int nL, int S, float* Ls, float* normS, int* idx,
                          int* counters, hipStream_t st)
{
    hipLaunchKernelGGL(k_sample_gather, dim3(kSampleRows / 4), dim3(256), 0, st, L, normL, nL, S, Ls, normS, idx,
                       counters);
}

// The sampled rows' decisions: one wave per sampled left feature merges its split
// partials (lane l takes every 64th), then the certified decision as in k_merge_decide.
__global__ __launch_bounds__(256) void k_sample_decide(Top2 const* __restrict__ colPart, int nParts, int ldc, int S,
                                                       float const* __restrict__ normS,
                                                       unsigned int const* __restrict__ maxNormBits, float epsRel,
                                                       float minScore, float minRatio, int* __restrict__ dec,
                                                       int* __restrict__ flagList, int* __restrict__ counters)
{
    int const lane = threadIdx.x & 63;
    int const q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= S) return;  // uniform per wave
    Top2 s = top2_init();
    for (int p = lane; p < nParts; p += 64) s = top2_merge(s, colPart[(size_t)p * ldc + q]);
    for (int o = 32; o > 0; o >>= 1) s = top2_merge(s, top2_shfl_xor(s, o));
    if (lane != 0) return;
    float const eps = epsRel * normS[q] * __uint_as_float(maxNormBits[1]) + kEpsAbs;
    int const d = (epsRel > 0.0f) ? decide_certified(s, eps, minScore, minRatio) : kFlag;
    if (d == kFlag) flagList[atomicAdd(&counters[kCntFlag], 1)] = q;
    dec[q] = d;
}

void launch_sample_decide(Top2 const* colPart, int nParts, int ldc, int S, float const* normS,
                          unsigned int const* maxNormBits, float epsRel, float minScore, float minRatio, int* dec,
                          int* flagList, int* counters, hipStream_t st)
{
    if (S <= 0) return;
    hipLaunchKernelGGL(k_sample_decide, dim3((S + 3) / 4), dim3(256), 0, st, colPart, nParts, ldc, S, normS,
                       maxNormBits, epsRel, minScore, minRatio, dec, flagList, counters);
}


}  // namespace v3dm
