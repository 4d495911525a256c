// Plain FP32 putative matching on gfx950: fused MFMA score tiles + top-2 in both
// directions, certified decisions, exact FP64 re-check of open queries, mutual check.
//
// Reference path: V3D_CUDA::SIFT_Matching::findPutativeMatches
// (V3D/CUDA/v3d_cudamatching.cu:522-552), which runs cublasSgemm into an N x M score
// buffer (cu:529-530) and two scans over it (cu:133-210).  Here the score matrix only
// ever lives in accumulator registers and LDS.
#include "v3dm_device.h"
#include "v3dm_kernels.h"

namespace v3dm {

// ------------------------------------------------------------------------------ norms
// norm[r] = ||x_r||_2 (FP32), maxNorm = max over rows (as ordered uint bits: norms are
// non-negative or NaN, and a NaN wins, which later flags every query).
__global__ __launch_bounds__(256) void k_norms(float const* __restrict__ X, int n, float* __restrict__ norm,
                                               unsigned int* __restrict__ maxNormBits)
{
    int const lane = threadIdx.x & 63;
    int const row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    float2 const v = reinterpret_cast<float2 const*>(X + (size_t)row * kDim)[lane];
    float s = v.x * v.x + v.y * v.y;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    float const nr = sqrtf(s);
    if (lane == 0) {
        norm[row] = nr;
        atomicMax(maxNormBits, __float_as_uint(nr));
    }
}

void launch_norms(float const* X, int n, float* norm, unsigned int* maxNormBits, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_norms, dim3((n + 3) / 4), dim3(256), 0, st, X, n, norm, maxNormBits);
}

// ------------------------------------------------------------------------------ fused tile kernel
// Work-group: 4 waves, 128 left features (each wave keeps 32 of them as the B operand of
// v_mfma_f32_32x32x2_f32 in 64 VGPRs), streaming 64-row tiles of right features through
// double-buffered LDS.  Lane l of a wave holds column i = l%32 of the 32x32 blocks and
// rows j = 8*(r/4) + 4*(l/32) + r%4, r = 0..15.
//  * per left i: running top-2 over j in registers (rows of the lane), merged with the
//    partner half-wave at the end -> colPart[split][i]
//  * per right j: the tile is written to LDS as S[j][i] and 4 threads per row reduce it
//    over the work-group's 128 left features -> rowPart[leftBlock][j]
// The k index of MFMA step s is h*64 + s (h = l/32), so each lane reads 4 consecutive
// floats per ds_read_b128; A and B use the same map, so the dot products are unchanged.
constexpr int kBI = 128;
constexpr int kBJ = 64;
constexpr int kStride = 132;  // LDS floats per tile row: 128 + 4 (conflict-free b128 reads)

__global__ __launch_bounds__(256, 2) void k_fused_fp32(float const* __restrict__ L, float const* __restrict__ R,
                                                       int nL, int nR, int tilesPerSplit,
                                                       Top2* __restrict__ colPart, int ldc,
                                                       Top2* __restrict__ rowPart, int ldr)
{
    __shared__ __attribute__((aligned(16))) float lds[2][kBJ * kStride];

    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const lb = blockIdx.x;
    int const split = blockIdx.y;
    int const i0 = lb * kBI;
    int const nTiles = (nR + kBJ - 1) / kBJ;
    int const t0 = split * tilesPerSplit;
    int const t1 = min(t0 + tilesPerSplit, nTiles);

    int const myI = i0 + wave * 32 + c32;
    float bfrag[64];
    {
        float4 const* lp = reinterpret_cast<float4 const*>(L + (size_t)myI * kDim + h * 64);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            float4 const v = lp[q];
            bfrag[4 * q + 0] = v.x;
            bfrag[4 * q + 1] = v.y;
            bfrag[4 * q + 2] = v.z;
            bfrag[4 * q + 3] = v.w;
        }
    }

    float4 const* R4 = reinterpret_cast<float4 const*>(R);
    // first tile -> lds[0]
    if (t0 < t1) {
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            int const idx = p * 256 + tid;
            int const row = idx >> 5, c4 = idx & 31;
            float4 const v = R4[(size_t)(t0 * kBJ + row) * 32 + c4];
            *reinterpret_cast<float4*>(&lds[0][row * kStride + c4 * 4]) = v;
        }
    }
    __syncthreads();

    Top2 colS = top2_init();
    int const rq = tid & 3;   // per-right reduction: quarter of the 128 left features
    int const rj = tid >> 2;  // per-right reduction: tile row

    for (int t = t0; t < t1; ++t) {
        int const buf = (t - t0) & 1;
        bool const hasNext = (t + 1) < t1;
        float4 pre[8];
        if (hasNext) {
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                int const idx = p * 256 + tid;
                int const row = idx >> 5, c4 = idx & 31;
                pre[p] = R4[(size_t)((t + 1) * kBJ + row) * 32 + c4];
            }
        }

        v16f acc0 = {};
        v16f acc1 = {};
        float const* A0 = &lds[buf][c32 * kStride + h * 64];
        float const* A1 = A0 + 32 * kStride;
#pragma unroll
        for (int s4 = 0; s4 < 16; ++s4) {
            float4 const a0 = *reinterpret_cast<float4 const*>(A0 + 4 * s4);
            float4 const a1 = *reinterpret_cast<float4 const*>(A1 + 4 * s4);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, bfrag[4 * s4 + 0], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, bfrag[4 * s4 + 0], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, bfrag[4 * s4 + 1], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, bfrag[4 * s4 + 1], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, bfrag[4 * s4 + 2], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, bfrag[4 * s4 + 2], acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, bfrag[4 * s4 + 3], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, bfrag[4 * s4 + 3], acc1, 0, 0, 0);
        }
        __syncthreads();  // every wave is done reading lds[buf]

        if (hasNext) {
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                int const idx = p * 256 + tid;
                int const row = idx >> 5, c4 = idx & 31;
                *reinterpret_cast<float4*>(&lds[buf ^ 1][row * kStride + c4 * 4]) = pre[p];
            }
        }

        int const jt = t * kBJ;
        float* S = lds[buf];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
            float const v0 = acc0[r];
            float const v1 = acc1[r];
            top2_push(colS, (jt + jl < nR) ? v0 : -INFINITY, jt + jl);
            top2_push(colS, (jt + jl + 32 < nR) ? v1 : -INFINITY, jt + jl + 32);
            S[jl * kStride + wave * 32 + c32] = v0;
            S[(jl + 32) * kStride + wave * 32 + c32] = v1;
        }
        __syncthreads();

        Top2 rs = top2_init();
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            int const il = c * 4 + rq;
            float const v = S[rj * kStride + il];
            top2_push(rs, (i0 + il < nL) ? v : -INFINITY, i0 + il);
        }
        rs = top2_merge(rs, top2_shfl_xor(rs, 1));
        rs = top2_merge(rs, top2_shfl_xor(rs, 2));
        if (rq == 0 && jt + rj < nR) rowPart[(size_t)lb * ldr + jt + rj] = rs;
    }

    colS = top2_merge(colS, top2_shfl_xor(colS, 32));
    if (h == 0 && myI < nL) colPart[(size_t)split * ldc + myI] = colS;
}

FusedPlan plan_fused_fp32(int nL, int nR, int targetGroups)
{
    FusedPlan p;
    p.leftBlocks = (nL + kBI - 1) / kBI;
    int const nTiles = (nR + kBJ - 1) / kBJ;
    int splits = (targetGroups + p.leftBlocks - 1) / p.leftBlocks;
    if (splits > kMaxSplits) splits = kMaxSplits;
    if (splits > nTiles) splits = nTiles;
    if (splits < 1) splits = 1;
    p.tilesPerSplit = (nTiles + splits - 1) / splits;
    p.splits = (nTiles + p.tilesPerSplit - 1) / p.tilesPerSplit;
    return p;
}

void launch_fused_fp32(float const* L, float const* R, int nL, int nR, FusedPlan const& p, Top2* colPart,
                       int ldc, Top2* rowPart, int ldr, hipStream_t st)
{
    if (nL <= 0 || nR <= 0) return;
    hipLaunchKernelGGL(k_fused_fp32, dim3(p.leftBlocks, p.splits), dim3(256), 0, st, L, R, nL, nR,
                       p.tilesPerSplit, colPart, ldc, rowPart, ldr);
}

// ------------------------------------------------------------------------------ merge + decide
__global__ __launch_bounds__(256) void k_merge_decide(Top2 const* __restrict__ colPart, int nColParts, int ldc,
                                                      Top2 const* __restrict__ rowPart, int nRowParts, int ldr,
                                                      int nL, int nR, float const* __restrict__ normL,
                                                      float const* __restrict__ normR,
                                                      unsigned int const* __restrict__ maxNormBits, float epsRel,
                                                      float minScore, float minRatio, int* __restrict__ colMax,
                                                      int* __restrict__ rowMax, int* __restrict__ flagList,
                                                      int* __restrict__ counters)
{
    int const g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nL + nR) return;
    bool const left = g < nL;
    int const q = left ? g : g - nL;
    Top2 const* part = left ? colPart : rowPart;
    int const np = left ? nColParts : nRowParts;
    int const ld = left ? ldc : ldr;
    Top2 s = part[q];
    for (int p = 1; p < np; ++p) s = top2_merge(s, part[(size_t)p * ld + q]);
    float const nq = left ? normL[q] : normR[q];
    float const mo = __uint_as_float(maxNormBits[left ? 1 : 0]);
    float const eps = epsRel * nq * mo + kEpsAbs;
    int d = decide_certified(s, eps, minScore, minRatio);
    if (d == kFlag) {
        int const slot = atomicAdd(&counters[kCntFlag], 1);
        flagList[slot] = g;
    }
    (left ? colMax : rowMax)[q] = d;
}

void launch_merge_decide(Top2 const* colPart, int nColParts, int ldc, Top2 const* rowPart, int nRowParts,
                         int ldr, int nL, int nR, float const* normL, float const* normR,
                         unsigned int const* maxNormBits, float epsRel, float minScore, float minRatio,
                         int* colMax, int* rowMax, int* flagList, int* counters, hipStream_t st)
{
    int const n = nL + nR;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_merge_decide, dim3((n + 255) / 256), dim3(256), 0, st, colPart, nColParts, ldc, rowPart,
                       nRowParts, ldr, nL, nR, normL, normR, maxNormBits, epsRel, minScore, minRatio, colMax,
                       rowMax, flagList, counters);
}

// ------------------------------------------------------------------------------ exact re-check
// One work-group per open query (grid-stride over the list): exact scores
// fl32(sum_k double(q_k) * double(c_k)) in ascending k against every candidate, top-2
// with the first-index tie rule, and the decision in IEEE FP32.
__global__ __launch_bounds__(256) void k_recheck_exact(float const* __restrict__ L, float const* __restrict__ R,
                                                       int nL, int nR, int const* __restrict__ flagList,
                                                       int const* __restrict__ counters, float minScore,
                                                       float minRatio, int* __restrict__ colMax,
                                                       int* __restrict__ rowMax)
{
    __shared__ double qd[kDim];
    __shared__ Top2 red[4];
    int const tid = threadIdx.x;
    int const cnt = counters[kCntFlag];
    for (int e = blockIdx.x; e < cnt; e += gridDim.x) {
        int const code = flagList[e];
        bool const left = code < nL;
        int const q = left ? code : code - nL;
        float const* Q = (left ? L : R) + (size_t)q * kDim;
        float const* C = left ? R : L;
        int const nC = left ? nR : nL;
        __syncthreads();
        if (tid < kDim) qd[tid] = (double)Q[tid];
        __syncthreads();
        Top2 s = top2_init();
        for (int c = tid; c < nC; c += 256) {
            float4 const* cp = reinterpret_cast<float4 const*>(C + (size_t)c * kDim);
            double acc = 0.0;
#pragma unroll 8
            for (int k4 = 0; k4 < 32; ++k4) {
                float4 const v = cp[k4];
                acc += qd[4 * k4 + 0] * (double)v.x;
                acc += qd[4 * k4 + 1] * (double)v.y;
                acc += qd[4 * k4 + 2] * (double)v.z;
                acc += qd[4 * k4 + 3] * (double)v.w;
            }
            top2_push(s, (float)acc, c);
        }
        for (int o = 32; o > 0; o >>= 1) s = top2_merge(s, top2_shfl_xor(s, o));
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            Top2 t = red[0];
            for (int w = 1; w < 4; ++w) t = top2_merge(t, red[w]);
            (left ? colMax : rowMax)[q] = decide_exact(t, minScore, minRatio);
        }
    }
}

void launch_recheck_exact(float const* L, float const* R, int nL, int nR, int const* flagList,
                          int const* counters, float minScore, float minRatio, int* colMax, int* rowMax,
                          int grid, hipStream_t st)
{
    if (nL <= 0 || nR <= 0) return;
    hipLaunchKernelGGL(k_recheck_exact, dim3(grid), dim3(256), 0, st, L, R, nL, nR, flagList, counters, minScore,
                       minRatio, colMax, rowMax);
}

// ------------------------------------------------------------------------------ mutual + compaction
// One work-group: (k, colMax[k]) is a match iff rowMax[colMax[k]] == k (cu:546-551);
// matches are written in ascending k with a block-wide prefix count.
__global__ __launch_bounds__(1024) void k_mutual_compact(int const* __restrict__ colMax,
                                                         int const* __restrict__ rowMax, int nL, int nR,
                                                         int2* __restrict__ pairs, int* __restrict__ counters)
{
    __shared__ int wsum[16];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int base = 0;
    for (int c0 = 0; c0 < nL; c0 += 1024) {
        int const i = c0 + tid;
        int l = -1;
        bool m = false;
        if (i < nL) {
            l = colMax[i];
            m = (l >= 0) && (l < nR) && (rowMax[l] == i);
        }
        unsigned long long const bal = __ballot(m);
        int const pre = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            int const x = wsum[w];
            off += (w < wave) ? x : 0;
            tot += x;
        }
        if (m) pairs[base + off + pre] = make_int2(i, l);
        base += tot;
        __syncthreads();
    }
    if (tid == 0) counters[kCntPairs] = base;
}

void launch_mutual_compact(int const* colMax, int const* rowMax, int nL, int nR, int2* pairs, int* counters,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_mutual_compact, dim3(1), dim3(1024), 0, st, colMax, rowMax, nL, nR, pairs, counters);
}

// ------------------------------------------------------------------------------ counters reset
__global__ void k_reset_counters(int* counters, int n)
{
    if ((int)threadIdx.x < n) counters[threadIdx.x] = 0;
}

void launch_reset_counters(int* counters, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, counters, n);
}

}  // namespace v3dm

namespace v3dm {

// ------------------------------------------------------------------------------ uint8 dequantisation
// One wave per descriptor, 2 components per lane.  The sum of squares of integers <= 255
// over 128 terms stays below 2^24, so it is exact in FP32 in any order.
__global__ __launch_bounds__(256) void k_dequant_u8(uint8_t const* __restrict__ q, int n, float* __restrict__ out)
{
    int const lane = threadIdx.x & 63;
    int const row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    uchar2 const v = reinterpret_cast<uchar2 const*>(q + (size_t)row * kDim)[lane];
    float const a = (float)v.x, b = (float)v.y;
    float s = a * a + b * b;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    float const denom = __fdiv_rn(1.0f, __fsqrt_rn(s));
    reinterpret_cast<float2*>(out + (size_t)row * kDim)[lane] = make_float2(a * denom, b * denom);
}

void launch_dequant_u8(uint8_t const* q, int n, float* out, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_dequant_u8, dim3((n + 3) / 4), dim3(256), 0, st, q, n, out);
}

}  // namespace v3dm
