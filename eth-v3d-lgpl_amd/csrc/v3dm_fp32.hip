// Plain FP32 putative matching on gfx950: fused MFMA score tiles + top-2 in both
// directions, certified decisions, exact FP64 re-check of open queries, mutual check.
//
// Reference path: V3D_CUDA::SIFT_Matching::findPutativeMatches
// (V3D/CUDA/v3d_cudamatching.cu:522-552), which runs cublasSgemm into an N x M score
// buffer (cu:529-530) and two scans over it (cu:133-210).  Here the score matrix only
// ever lives in accumulator registers and LDS.
#include "v3dm_device.h"
#include "v3dm_kernels.h"
#include "v3dm_batch.h"

namespace v3dm {

// ------------------------------------------------------------------------------ fused tile kernel
// Work-group: 4 waves, 128 left features (each wave keeps 32 of them as the B operand of
// v_mfma_f32_32x32x2_f32 in 64 VGPRs), streaming 64-row tiles of right features through
// two LDS stages filled by direct global->LDS loads (global_load_lds_dwordx4), so tile
// t+1 is in flight while tile t is multiplied.  Lane l of a wave holds column i = l%32 of
// the 32x32 blocks and rows j = 8*(r/4) + 4*(l/32) + r%4, r = 0..15.
//  * per left i: running top-2 over j in registers (rows of the lane), merged with the
//    partner half-wave at the end -> colPart[split][i]
//  * per right j: the tile's scores are written over the stage just consumed as S[j][i]
//    and 4 threads per row reduce them over the work-group's 128 left features
//    -> rowPart[leftBlock][j]
// The k index of MFMA step s is h*64 + s (h = l/32); A and B use the same map, so the dot
// products are unchanged.  Stage layout: row r, half h (k = 64h .. 64h+63, 256 bytes)
// is segment h*64 + r, stored at 272*(h*64 + r): a lane's 16 A chunks are contiguous
// (immediate offsets) and the 16-byte pad rotates consecutive rows across bank groups.
// The direct loads fill the stage 1 KiB at a time (lane L of a load lands at +16L), so
// each lane fetches whatever chunk belongs at its slot (pad slots refetch a neighbour).
// After the multiply the stage is reused for the transposed tile S[j][i] (row stride 132
// floats, conflict-free for the writes and the per-right reads).
//
// The batch form also has an 8-wave work-group (256 left features): each staged tile then
// serves twice as many left features, and the transposed tile (64 x 264 floats) needs a
// larger stage.  Shape<kW> holds the sizes of a kW-wave work-group.
constexpr int kBI = 128;
constexpr int kBJ = 64;
constexpr int kSeg = 272;             // bytes per stage segment (256 data + 16 pad)
constexpr int kStageData = 34;        // KiB of segments in a stage (2 * kBJ segments)

template <int kW>
struct Shape {
    static constexpr int kRows = 32 * kW;                          // left features per work-group
    static constexpr int kLoads = (kStageData + kW - 1) / kW;      // 1 KiB direct loads per wave per tile
    static constexpr int kSRow = kRows + kW;                       // floats per row of the transposed tile
    static constexpr int kLoadBytes = kLoads * kW * 1024;
    static constexpr int kTBytes = (kBJ * kSRow * 4 + 1023) / 1024 * 1024;
    static constexpr int kStage = kLoadBytes > kTBytes ? kLoadBytes : kTBytes;
};
constexpr int kStageLoads = Shape<4>::kLoads;
constexpr int kStageBytes = Shape<4>::kStage;
static_assert(kStageLoads == 9 && kStageBytes == 9 * 4 * 1024, "4-wave stage: 36 KiB");
static_assert(Shape<8>::kStage <= 80 * 1024, "two 8-wave stages fit one CU's LDS");

typedef __attribute__((address_space(3))) void lds_void;

// Byte offset (within a tile of R) fetched by slot `slot` of the stage.
__device__ __forceinline__ int stage_src_fp32(int slot)
{
    int const seg = slot / 17, within = slot - seg * 17;
    if (seg >= 2 * kBJ) return 0;  // tail of the stage: never read
    int const h = seg / kBJ, row = seg - h * kBJ;
    return row * 512 + (h * 16 + (within < 16 ? within : 15)) * 16;
}

template <int kW>
__device__ __forceinline__ void stage_tile_fp32(float const* __restrict__ R, int t, unsigned char* stage,
                                                int const* src, int wv)
{
    unsigned char const* tile = reinterpret_cast<unsigned char const*>(R + (size_t)t * kBJ * kDim);
#pragma unroll
    for (int p = 0; p < Shape<kW>::kLoads; ++p)
        __builtin_amdgcn_global_load_lds(tile + src[p], (lds_void*)(stage + (kW * p + wv) * 1024), 16, 0, 0);
}

// Removed (left, right) pairs of guided matching: bit (j & 31) of bits[(j >> 5) * ldi + i]
// set means the score of left i and right j is replaced by -1e20 (cu:308, cu:340).
struct GuidedBits {
    unsigned const* bits;
    int ldi;
};

template <bool kFullJ, bool kFullI, bool kGuided, int kW>
__device__ __forceinline__ void tile_fp32(unsigned char* stage, float const (&bfrag)[64], Top2& colS, Top2* rowPart,
                                          int ldr, int lb, int t, int nL, int nR, int i0, int tid, float inf,
                                          GuidedBits gb)
{
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    unsigned gw0 = 0, gw1 = 0;
    if (kGuided) {
        // words of rows 64t .. 64t+31 and 64t+32 .. 64t+63 for this lane's left feature
        size_t const gi = (size_t)(i0 + wave * 32 + c32);
        gw0 = gb.bits[(size_t)(2 * t) * gb.ldi + gi];
        gw1 = gb.bits[(size_t)(2 * t + 1) * gb.ldi + gi];
    }
    v16f acc0 = {};
    v16f acc1 = {};
    unsigned char const* A0 = stage + (h * kBJ + c32) * kSeg;
    unsigned char const* A1 = A0 + 32 * kSeg;
#pragma unroll
    for (int s4 = 0; s4 < 16; ++s4) {
        float4 const a0 = *reinterpret_cast<float4 const*>(A0 + 16 * s4);
        float4 const a1 = *reinterpret_cast<float4 const*>(A1 + 16 * s4);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, bfrag[4 * s4 + 0], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, bfrag[4 * s4 + 0], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, bfrag[4 * s4 + 1], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, bfrag[4 * s4 + 1], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, bfrag[4 * s4 + 2], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, bfrag[4 * s4 + 2], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, bfrag[4 * s4 + 3], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, bfrag[4 * s4 + 3], acc1, 0, 0, 0);
    }
    int const jt = t * kBJ;
    if (kGuided) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
            acc0[r] = ((gw0 >> jl) & 1u) ? -1e20f : acc0[r];
            acc1[r] = ((gw1 >> jl) & 1u) ? -1e20f : acc1[r];
        }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
        float v0 = acc0[r], v1 = acc1[r];
        if (!kFullJ) {
            v0 = (jt + jl < nR) ? v0 : -INFINITY;
            v1 = (jt + jl + 32 < nR) ? v1 : -INFINITY;
        }
        top2_push_fast(colS, v0, jt + jl, inf);
        top2_push_fast(colS, v1, jt + jl + 32, inf);
    }
    if (!rowPart) return;  // left decisions only (uniform: a kernel argument)
    __syncthreads();  // every wave is done reading the stage as the A operand
    constexpr int kSRow = Shape<kW>::kSRow;
    float* S = reinterpret_cast<float*>(stage);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        int const jl = 8 * (r >> 2) + 4 * h + (r & 3);
        S[jl * kSRow + wave * 32 + c32] = acc0[r];
        S[(jl + 32) * kSRow + wave * 32 + c32] = acc1[r];
    }
    __syncthreads();
    // kW threads per tile row, each over every kW-th left feature; the row stride (kRows +
    // kW floats) puts the rows that 32 lanes read on distinct banks
    int const rq = tid % kW;
    int const rj = tid / kW;  // tile row
    Top2 rs = top2_init();
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        int const il = c * kW + rq;
        float v = S[rj * kSRow + il];
        if (!kFullI) v = (i0 + il < nL) ? v : -INFINITY;
        top2_push_fast(rs, v, i0 + il, inf);
    }
#pragma unroll
    for (int x = 1; x < kW; x *= 2) rs = top2_merge(rs, top2_shfl_xor(rs, x));
    if (rq == 0 && jt + rj < nR) rowPart[(size_t)lb * ldr + jt + rj] = rs;
}

// The tiles [t0, t1) of one work-group, double-buffered; kMasked handles a partial last
// tile of right features or a partial block of left features.
template <bool kMasked, bool kGuided, int kW>
__device__ __forceinline__ void run_tiles_fp32(float const* __restrict__ R, unsigned char* stA, unsigned char* stB,
                                               int const* src, float const (&bfrag)[64], Top2& colS,
                                               Top2* rowPart, int ldr, int lb, int t0, int t1, int nL, int nR, int i0,
                                               int tid, int wv, float inf, GuidedBits gb)
{
    // two tiles per trip so that each stage is a distinct LDS object
    for (int t = t0; t < t1; t += 2) {
        wait_lds_loads();
        __syncthreads();  // tile t is in stA; every wave is done with stB
        if (t + 1 < t1) stage_tile_fp32<kW>(R, t + 1, stB, src, wv);
        tile_fp32<!kMasked, !kMasked, kGuided, kW>(stA, bfrag, colS, rowPart, ldr, lb, t, nL, nR, i0, tid, inf, gb);
        if (t + 1 >= t1) break;
        wait_lds_loads();
        __syncthreads();  // tile t+1 is in stB; every wave is done with stA
        if (t + 2 < t1) stage_tile_fp32<kW>(R, t + 2, stA, src, wv);
        tile_fp32<!kMasked, !kMasked, kGuided, kW>(stB, bfrag, colS, rowPart, ldr, lb, t + 1, nL, nR, i0, tid, inf,
                                                    gb);
    }
}

// One work unit: left block `lb` (32 * kW features) against right tiles [split *
// tilesPerSplit, ...) -> colPart[split * ldc + i], rowPart[lb * ldr + j].
template <bool kGuided = false, int kW = 4>
__device__ __forceinline__ void fused_block(float const* __restrict__ L, float const* __restrict__ R, int nL, int nR,
                                            int lb, int split, int tilesPerSplit, Top2* __restrict__ colPart,
                                            int ldc, Top2* __restrict__ rowPart, int ldr, unsigned char* stA,
                                            unsigned char* stB, float inf, GuidedBits gb = GuidedBits{nullptr, 0})
{
    // an opaque copy of the thread index: in the persistent kernel it keeps the per-lane
    // offsets derived from it inside the item loop instead of live across it (spills)
    int tid = threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const wv = __builtin_amdgcn_readfirstlane(wave);
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const i0 = lb * Shape<kW>::kRows;
    int const nTiles = (nR + kBJ - 1) / kBJ;
    int const t0 = split * tilesPerSplit;
    int const t1 = min(t0 + tilesPerSplit, nTiles);

    int src[Shape<kW>::kLoads];
#pragma unroll
    for (int p = 0; p < Shape<kW>::kLoads; ++p) src[p] = stage_src_fp32((kW * p + wv) * 64 + lane);
    if (t0 < t1) stage_tile_fp32<kW>(R, t0, stA, src, wv);

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

    Top2 colS = top2_init();
    bool const masked = (i0 + Shape<kW>::kRows > nL) || (t1 * kBJ > nR);
    if (masked)
        run_tiles_fp32<true, kGuided, kW>(R, stA, stB, src, bfrag, colS, rowPart, ldr, lb, t0, t1, nL, nR, i0, tid,
                                          wv, inf, gb);
    else
        run_tiles_fp32<false, kGuided, kW>(R, stA, stB, src, bfrag, colS, rowPart, ldr, lb, t0, t1, nL, nR, i0, tid,
                                           wv, inf, gb);

    colS = top2_merge(colS, top2_shfl_xor(colS, 32));
    if (h == 0 && myI < nL) colPart[(size_t)split * ldc + myI] = colS;
}

__global__ __launch_bounds__(256, 2) void k_fused_fp32(float const* __restrict__ L, float const* __restrict__ R,
                                                       int nL, int nR, int tilesPerSplit,
                                                       Top2* __restrict__ colPart, int ldc,
                                                       Top2* __restrict__ rowPart, int ldr, float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[kStageBytes];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[kStageBytes];
    fused_block(L, R, nL, nR, blockIdx.x, blockIdx.y, tilesPerSplit, colPart, ldc, rowPart, ldr, stA, stB, inf);
}

// Guided matching: the same tiles with the removed pairs of the bitmap scored -1e20.
__global__ __launch_bounds__(256, 2) void k_guided_fp32(float const* __restrict__ L, float const* __restrict__ R,
                                                        int nL, int nR, int tilesPerSplit,
                                                        Top2* __restrict__ colPart, int ldc,
                                                        Top2* __restrict__ rowPart, int ldr,
                                                        unsigned const* __restrict__ bits, int ldi, float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[kStageBytes];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[kStageBytes];
    fused_block<true>(L, R, nL, nR, blockIdx.x, blockIdx.y, tilesPerSplit, colPart, ldc, rowPart, ldr, stA, stB, inf,
                      GuidedBits{bits, ldi});
}

// Persistent form over a batch: each work-group takes items from the queue until it runs
// dry.  Item = (pair, left block of 32 * kW features, split, tiles per split); the left
// queries of pair p are segment p, its right queries segment nPairs + p.  kW = 4: two
// work-groups per CU; kW = 8: one (132 KiB of LDS).
template <int kW>
__global__ __launch_bounds__(64 * kW, 2) void k_batch_fused_fp32(BatchDev b, int4 const* __restrict__ items,
                                                                 int nItems, int const* __restrict__ nItemsDev,
                                                                 int* __restrict__ queue, Top2* __restrict__ part,
                                                                 float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[Shape<kW>::kStage];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[Shape<kW>::kStage];
    __shared__ int sItem;
    if (nItemsDev) nItems = __builtin_amdgcn_readfirstlane(min(nItems, *nItemsDev));  // written by an earlier kernel
    for (;;) {
        if (threadIdx.x == 0) sItem = atomicAdd(queue, 1);
        __syncthreads();  // also: every wave is done with the stages of the previous item
        int const it = __builtin_amdgcn_readfirstlane(sItem);  // uniform: keeps the item in SGPRs
        __syncthreads();
        if (it >= nItems) break;
        int4 const w = items[it];
        BatchSeg const sl = b.seg[w.x];
        BatchSeg const sr = b.seg[b.nPairs + w.x];
        float const* L = b.img + (size_t)b.imgBase[sl.img] * kDim;
        float const* R = b.img + (size_t)b.imgBase[sr.img] * kDim;
        fused_block<false, kW>(L, R, sl.n, sr.n, w.y, w.z, w.w, part + sl.partOff, sl.ld, part + sr.partOff, sr.ld,
                               stA, stB, inf);
    }
}

// Sampled check of a batch: item = (pair, -, split, tiles per split); the left block is the
// pair's gathered sample rows (kSampleRows, from the side sample_side picks) against the
// other image, the samples' decisions only.
__global__ __launch_bounds__(256, 2) void k_batch_sample_fp32(BatchDev b, float const* __restrict__ sRows,
                                                              int const* __restrict__ sCount,
                                                              int4 const* __restrict__ items, int nItems,
                                                              int* __restrict__ queue, Top2* __restrict__ sPart,
                                                              int maxSplits, float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[kStageBytes];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[kStageBytes];
    __shared__ int sItem;
    for (;;) {
        if (threadIdx.x == 0) sItem = atomicAdd(queue, 1);
        __syncthreads();
        int const it = __builtin_amdgcn_readfirstlane(sItem);
        __syncthreads();
        if (it >= nItems) break;
        int4 const w = items[it];
        BatchSeg const sr = b.seg[(1 - sample_side(b, w.x)) * b.nPairs + w.x];  // the image the samples scan
        float const* R = b.img + (size_t)b.imgBase[sr.img] * kDim;
        fused_block(sRows + (size_t)w.x * kSampleRows * kDim, R, sCount[w.x], sr.n, 0, w.z, w.w,
                    sPart + (size_t)w.x * maxSplits * kSampleRows, kSampleRows, nullptr, 0, stA, stB, inf);
    }
}

void launch_batch_sample_fp32(BatchDev const& b, float const* sRows, int const* sCount, int4 const* items,
                              int nItems, int* queue, Top2* sPart, int maxSplits, int grid, hipStream_t st)
{
    if (nItems <= 0) return;
    hipLaunchKernelGGL(k_batch_sample_fp32, dim3(grid), dim3(256), 0, st, b, sRows, sCount, items, nItems, queue,
                       sPart, maxSplits, INFINITY);
}

void launch_batch_fused_fp32(BatchDev const& b, int4 const* items, int nItems, int* queue, Top2* part, int grid,
                             int leftRows, hipStream_t st, int const* nItemsDev)
{
    if (nItems <= 0) return;
    if (leftRows == Shape<8>::kRows)
        hipLaunchKernelGGL(k_batch_fused_fp32<8>, dim3(grid), dim3(512), 0, st, b, items, nItems, nItemsDev, queue,
                           part, INFINITY);
    else
        hipLaunchKernelGGL(k_batch_fused_fp32<4>, dim3(grid), dim3(256), 0, st, b, items, nItems, nItemsDev, queue,
                           part, INFINITY);
}

FusedPlan plan_fused_fp32(int nL, int nR, int targetGroups, int maxSplits)
{
    FusedPlan p;
    p.leftBlocks = (nL + kBI - 1) / kBI;
    int const nTiles = (nR + kBJ - 1) / kBJ;
    int splits = (targetGroups + p.leftBlocks - 1) / p.leftBlocks;
    if (splits > maxSplits) splits = maxSplits;
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

void launch_guided_fp32(float const* L, float const* R, int nL, int nR, FusedPlan const& p, Top2* colPart,
                        int ldc, Top2* rowPart, int ldr, unsigned const* bits, int ldi, hipStream_t st)
{
    if (nL <= 0 || nR <= 0) return;
    hipLaunchKernelGGL(k_guided_fp32, dim3(p.leftBlocks, p.splits), dim3(256), 0, st, L, R, nL, nR,
                       p.tilesPerSplit, colPart, ldc, rowPart, ldr, bits, ldi, INFINITY);
}

}  // namespace v3dm
