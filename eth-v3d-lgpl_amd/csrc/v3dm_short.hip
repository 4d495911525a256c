// Reduced-precision shortlist forms of the putative matcher (BF16 MFMA).
//
// The reference computes every score in FP32 (cublasSgemm, V3D/CUDA/v3d_cudamatching.cu:
// 529-530).  Here each direction is scanned once with v_mfma_f32_32x32x16_bf16 and FP32
// accumulation; every query keeps, per partition of its candidates, the two best
// approximate scores with their indices.  The union of the partition lists is the
// shortlist; the candidates whose approximate score lies within twice the BF16 error bound
// of the approximate second best are rescored exactly (v3dm_rescore.hip), and the
// reference's selection rule (cu:133-210) runs on those exact scores.
#include "v3dm_device.h"
#include "v3dm_kernels.h"
#include "v3dm_short.h"
#include "v3dm_batch.h"

namespace v3dm {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------------------ FP32 -> BF16
// Round to nearest even; 4 components per thread.
__global__ __launch_bounds__(256) void k_to_bf16(float const* __restrict__ X, int nElems, __bf16* __restrict__ Y)
{
    int const i = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (i >= nElems) return;
    float4 const v = *reinterpret_cast<float4 const*>(X + i);
    Y[i + 0] = (__bf16)v.x;
    Y[i + 1] = (__bf16)v.y;
    Y[i + 2] = (__bf16)v.z;
    Y[i + 3] = (__bf16)v.w;
}

void launch_to_bf16(float const* X, int rows, __bf16* Y, hipStream_t st)
{
    int const n = rows * kDim;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_bf16, dim3((n / 4 + 255) / 256), dim3(256), 0, st, X, n, Y);
}

// ------------------------------------------------------------------------------ one-direction pass
constexpr int kSQ = 256;          // queries per work-group
constexpr int kST = 64;           // candidates per tile (16 KB of bf16)

typedef __attribute__((address_space(3))) void lds_void;

struct KeyState {
    float k1, k2;
    int t1, t2;
};

// Tile t of the candidates -> one LDS stage by direct global->LDS loads
// (global_load_lds_dwordx4, lane L of a load landing at +16L).  Stage layout: row r, half
// h (k = 64h .. 64h+63, 128 bytes) is segment h*64 + r at 144*(h*64 + r): a lane's 8 A
// chunks are contiguous (immediate offsets) and the 16-byte pad rotates consecutive rows
// across bank groups.  Each lane fetches the chunk that belongs at its slot; pad slots
// refetch a neighbour.
constexpr int kSegB = 144;
constexpr int kStageDataB = 18;  // KiB of segments in a stage (2 * kST segments)

// Sizes of a work-group of kQW waves (64 queries each): 1 KiB direct loads per wave per
// tile, and the stage they fill.
template <int kQW>
struct ShapeB {
    static constexpr int kLoads = (kStageDataB + kQW - 1) / kQW;
    static constexpr int kStage = kLoads * kQW * 1024;
};
constexpr int kStageLoadsB = ShapeB<4>::kLoads;
constexpr int kStageBytesB = ShapeB<4>::kStage;
static_assert(kStageLoadsB == 5 && kSQ == 4 * 64, "single-pair pass: 4 waves, 20 KiB stages");

__device__ __forceinline__ int stage_src_bf16(int slot)
{
    int const seg = slot / 9, within = slot - seg * 9;
    if (seg >= 2 * kST) return 0;  // tail of the stage: never read
    int const h = seg / kST, row = seg - h * kST;
    return row * 256 + (h * 8 + (within < 8 ? within : 7)) * 16;
}

template <int kQW>
__device__ __forceinline__ void stage_tile_bf16(__bf16 const* __restrict__ C, int t, unsigned char* stage,
                                                int const* src, int wv)
{
    unsigned char const* tile = reinterpret_cast<unsigned char const*>(C + (size_t)t * kST * kDim);
#pragma unroll
    for (int p = 0; p < ShapeB<kQW>::kLoads; ++p)
        __builtin_amdgcn_global_load_lds(tile + src[p], (lds_void*)(stage + (kQW * p + wv) * 1024), 16, 0, 0);
}

template <bool kFull>
__device__ __forceinline__ void tile_bf16(unsigned char const* stage, v8bf const (&bq)[2][8], KeyState (&st)[2], int t,
                                          int nC, int h, int c32, float inf)
{
    v16f acc[2][2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        acc[cb][0] = v16f{};
        acc[cb][1] = v16f{};
        unsigned char const* rowp = stage + (h * kST + cb * 32 + c32) * kSegB;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            v8bf const a = *reinterpret_cast<v8bf const*>(rowp + 16 * s);
            acc[cb][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[0][s], acc[cb][0], 0, 0, 0);
            acc[cb][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bq[1][s], acc[cb][1], 0, 0, 0);
        }
    }
    int const jt = t * kST;
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        float const o1 = st[qb].k1, o2 = st[qb].k2;
        float const k1 = group_best<kFull>(acc[0][qb], acc[1][qb], 0, o1, jt, nC, h);
        float const k2 = group_best<kFull>(acc[0][qb], acc[1][qb], 1, o2, jt, nC, h);
        st[qb].t1 = (k1 != o1) ? t : st[qb].t1;
        st[qb].t2 = (k2 != o2) ? t : st[qb].t2;
        st[qb].k1 = k1;
        st[qb].k2 = k2;
    }
}

// The tiles [t0, t1) of one work-group, double-buffered (two tiles per trip so that each
// stage is a distinct LDS object); kFull = no tile of the range is partial.
template <bool kFull, int kQW>
__device__ __forceinline__ void run_tiles_bf16(__bf16 const* __restrict__ C, unsigned char* stA, unsigned char* stB,
                                               int const* src, v8bf const (&bq)[2][8],
                                               KeyState (&st)[2], int t0, int t1, int nC, int h, int c32, int wv,
                                               float inf)
{
    for (int t = t0; t < t1; t += 2) {
        wait_lds_loads();
        __syncthreads();  // tile t is in stA; every wave is done with stB
        if (t + 1 < t1) stage_tile_bf16<kQW>(C, t + 1, stB, src, wv);
        tile_bf16<kFull>(stA, bq, st, t, nC, h, c32, inf);
        if (t + 1 >= t1) break;
        wait_lds_loads();
        __syncthreads();  // tile t+1 is in stB; every wave is done with stA
        if (t + 2 < t1) stage_tile_bf16<kQW>(C, t + 2, stA, src, wv);
        tile_bf16<kFull>(stB, bq, st, t + 1, nC, h, c32, inf);
    }
}

// Work-group: 4 waves; each wave keeps 64 queries (two 32-query blocks) as B operands
// (lane l: query l%32 of the block, bf16 k = h*64 + 8*s + v for MFMA step s, h = l/32),
// so the work-group covers 256 queries.  Candidates stream through two LDS stages in
// tiles of 64 (two 32-row blocks, A operand with the same k map): while tile t is
// multiplied, tile t+1 is already in flight into the other stage.  Lane l then holds, per
// 32x32 block, the scores of query l%32 against candidate rows jl = 8*(r/4) + 4*h + r%4,
// r = 0..15.  Per query and lane half the best key of each of two disjoint candidate
// groups is kept (the group maxima stand in for a top-2: the rescoring's window is set from
// the largest two listed values): the FP32 score with its 6 low mantissa bits replaced by
// the position (block*16 + r) inside the tile, so that one max3 tree updates value and
// position together; the tile of each group's key is tracked once per tile.  Masked
// candidates score -1e30.
// One work unit: query block `qblk` (64 * kQW queries; the single-pair pass: 4 waves, 256
// queries) against candidate tiles [split * tilesPerSplit, ...) -> part[((split * ldq) + q)
// * 2 + h].
template <int kQW = 4>
__device__ __forceinline__ void pass_bf16_block(__bf16 const* __restrict__ Q, int nQ, __bf16 const* __restrict__ C,
                                                int nC, int qblk, int split, int tilesPerSplit,
                                                Cand2* __restrict__ part, int ldq, unsigned char* stA,
                                                unsigned char* stB, float inf)
{
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const wv = __builtin_amdgcn_readfirstlane(wave);
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const q0 = qblk * (64 * kQW) + wave * 64;
    int const nTiles = (nC + kST - 1) / kST;
    int const t0 = split * tilesPerSplit;
    int const t1 = min(t0 + tilesPerSplit, nTiles);

    int src[ShapeB<kQW>::kLoads];
#pragma unroll
    for (int p = 0; p < ShapeB<kQW>::kLoads; ++p) src[p] = stage_src_bf16((kQW * p + wv) * 64 + lane);
    if (t0 < t1) stage_tile_bf16<kQW>(C, t0, stA, src, wv);

    v8bf bq[2][8];
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        uint4 const* qp = reinterpret_cast<uint4 const*>(Q + (size_t)(q0 + qb * 32 + c32) * kDim + h * 64);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            uint4 const u = qp[s];
            bq[qb][s] = *reinterpret_cast<v8bf const*>(&u);
        }
    }

    KeyState st[2];
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) st[qb] = KeyState{-INFINITY, -INFINITY, -1, -1};

    if (t1 * kST > nC) run_tiles_bf16<false, kQW>(C, stA, stB, src, bq, st, t0, t1, nC, h, c32, wv, inf);
    else run_tiles_bf16<true, kQW>(C, stA, stB, src, bq, st, t0, t1, nC, h, c32, wv, inf);

#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        int const q = q0 + qb * 32 + c32;
        if (q >= nQ) continue;
        Cand2 c;
        int const p1 = (int)(__float_as_uint(st[qb].k1) & 63u);
        int const p2 = (int)(__float_as_uint(st[qb].k2) & 63u);
        c.v1 = st[qb].k1;
        c.v2 = st[qb].k2;
        c.j1 = (st[qb].t1 >= 0 && c.v1 > -1e29f)
                   ? st[qb].t1 * kST + (p1 >> 4) * 32 + 8 * ((p1 & 15) >> 2) + 4 * h + (p1 & 3) : -1;
        c.j2 = (st[qb].t2 >= 0 && c.v2 > -1e29f)
                   ? st[qb].t2 * kST + (p2 >> 4) * 32 + 8 * ((p2 & 15) >> 2) + 4 * h + (p2 & 3) : -1;
        part[((size_t)split * ldq + q) * 2 + h] = c;
    }
}

// Both directions of a pair in one launch: work-groups [0, queryBlocks0 * splits0) run
// direction 0 (query block fastest), the rest direction 1.
struct Bf16Dir {
    __bf16 const* Q;
    __bf16 const* C;
    Cand2* part;
    int nQ, nC, tilesPerSplit, ldq, queryBlocks, splits;
};

__global__ __launch_bounds__(256, 2) void k_pass_bf16(Bf16Dir d0, Bf16Dir d1, float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[kStageBytesB];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[kStageBytesB];
    int b = blockIdx.x;
    int const n0 = d0.queryBlocks * d0.splits;
    bool const second = b >= n0;  // uniform per work-group
    Bf16Dir const d = second ? d1 : d0;
    if (second) b -= n0;
    int const split = b / d.queryBlocks;
    pass_bf16_block(d.Q, d.nQ, d.C, d.nC, b - split * d.queryBlocks, split, d.tilesPerSplit, d.part, d.ldq, stA, stB,
                    inf);
}

// Persistent form over a batch: each work-group takes items from the queue until it runs
// dry.  Item = (segment, query block of 64 * kQW queries, split, tiles per split).
template <int kQW>
__global__ __launch_bounds__(64 * kQW, 2) void k_batch_pass_bf16(BatchDev b, int4 const* __restrict__ items,
                                                                 int nItems, int* __restrict__ queue,
                                                                 Cand2* __restrict__ part, float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[ShapeB<kQW>::kStage];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[ShapeB<kQW>::kStage];
    __shared__ int sItem;
    for (;;) {
        if (threadIdx.x == 0) sItem = atomicAdd(queue, 1);
        __syncthreads();  // also: every wave is done with the stages of the previous item
        int const it = __builtin_amdgcn_readfirstlane(sItem);  // uniform: keeps the item in SGPRs
        __syncthreads();
        if (it >= nItems) break;
        int4 const w = items[it];
        BatchSeg const sg = b.seg[w.x];
        __bf16 const* Q = b.imgB + (size_t)b.imgBase[sg.img] * kDim;
        __bf16 const* C = b.imgB + (size_t)b.imgBase[sg.oImg] * kDim;
        pass_bf16_block<kQW>(Q, sg.n, C, sg.nO, w.y, w.z, w.w, part + (size_t)sg.partOff * 2, sg.ld, stA, stB, inf);
    }
}

void launch_batch_pass_bf16(BatchDev const& b, int4 const* items, int nItems, int* queue, Cand2* part, int grid,
                            int queryRows, hipStream_t st)
{
    if (nItems <= 0) return;
    if (queryRows == 128)
        hipLaunchKernelGGL(k_batch_pass_bf16<2>, dim3(grid), dim3(128), 0, st, b, items, nItems, queue, part, INFINITY);
    else
        hipLaunchKernelGGL(k_batch_pass_bf16<4>, dim3(grid), dim3(256), 0, st, b, items, nItems, queue, part, INFINITY);
}

ShortPlan plan_short(int nQ, int nC, int targetGroups)
{
    ShortPlan p;
    p.queryBlocks = (nQ + kSQ - 1) / kSQ;
    int const nTiles = (nC + kST - 1) / kST;
    int splits = (targetGroups + p.queryBlocks - 1) / p.queryBlocks;
    if (splits > kShortMaxSplits) splits = kShortMaxSplits;
    if (splits > nTiles) splits = nTiles;
    if (splits < 1) splits = 1;
    p.tilesPerSplit = (nTiles + splits - 1) / splits;
    p.splits = (nTiles + p.tilesPerSplit - 1) / p.tilesPerSplit;
    return p;
}

ShortPlan plan_short_packed(int nQ, int nC, int targetGroups)
{
    ShortPlan p = plan_short(nQ, nC, targetGroups);
    if (p.tilesPerSplit > kPackedMaxTiles) {  // more splits, so that a split's index fits 16 bits
        int const nTiles = (nC + kST - 1) / kST;
        int const splits = (nTiles + kPackedMaxTiles - 1) / kPackedMaxTiles;
        p.tilesPerSplit = (nTiles + splits - 1) / splits;
        p.splits = (nTiles + p.tilesPerSplit - 1) / p.tilesPerSplit;
    }
    return p;
}

void launch_pass_bf16_2(__bf16 const* A, int nA, __bf16 const* B, int nB, ShortPlan const& pa, Cand2* partA,
                        int ldA, ShortPlan const& pb, Cand2* partB, int ldB, hipStream_t st)
{
    if (nA <= 0 || nB <= 0) return;
    Bf16Dir const d0{A, B, partA, nA, nB, pa.tilesPerSplit, ldA, pa.queryBlocks, pa.splits};
    Bf16Dir const d1{B, A, partB, nB, nA, pb.tilesPerSplit, ldB, pb.queryBlocks, pb.splits};
    hipLaunchKernelGGL(k_pass_bf16, dim3(pa.queryBlocks * pa.splits + pb.queryBlocks * pb.splits), dim3(256), 0, st,
                       d0, d1, INFINITY);
}



}  // namespace v3dm
