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
                                                       Top2* __restrict__ rowPart, int ldr, float inf)
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
        if (jt + kBJ <= nR) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
                top2_push_fast(colS, acc0[r], jt + jl, inf);
                top2_push_fast(colS, acc1[r], jt + jl + 32, inf);
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
                top2_push_fast(colS, (jt + jl < nR) ? acc0[r] : -INFINITY, jt + jl, inf);
                top2_push_fast(colS, (jt + jl + 32 < nR) ? acc1[r] : -INFINITY, jt + jl + 32, inf);
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
            S[jl * kStride + wave * 32 + c32] = acc0[r];
            S[(jl + 32) * kStride + wave * 32 + c32] = acc1[r];
        }
        __syncthreads();

        Top2 rs = top2_init();
        if (i0 + kBI <= nL) {
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                int const il = c * 4 + rq;
                top2_push_fast(rs, S[rj * kStride + il], i0 + il, inf);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 32; ++c) {
                int const il = c * 4 + rq;
                float const v = S[rj * kStride + il];
                top2_push_fast(rs, (i0 + il < nL) ? v : -INFINITY, i0 + il, inf);
            }
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
                       p.tilesPerSplit, colPart, ldc, rowPart, ldr, INFINITY);
}

}  // namespace v3dm
