// Code pass of the learned shortlist: integer MFMA over 32-d codes with the candidate
// position folded into the accumulator, top-2 per query and partition on integer keys.
//
// Codes: the unit-norm network output c quantised to q = round(c * scale) in [-31, 31] and
// stored as the int8 value 4q (32 bytes per feature).  v_mfma_i32_32x32x32_i8 with the
// accumulator initialised to r (the element index 0..15 of the D block) gives, for lane l
// and element r, key = 16 * sum_k q_c q_q + r: the score with the candidate's row inside
// the block in its low 4 bits (row 8*(r/4) + 4*(l/32) + r%4, query l%32).  Keys order as
// (score, row), so the best of a set of candidates is a chain of v_max3_i32.  Each lane
// keeps the best key of two disjoint candidate groups: group g takes the elements with
// r / 8 == g of a tile's first candidate block and those with r / 8 == 1 - g of its second,
// so a key's bit 3 tells the block (g ^ bit 3) and the tile of each kept key is tracked
// once per tile.  A partition (split, lane half) lists two candidates, one per group.
#include "v3dm_batch.h"
#include "v3dm_device.h"
#include "v3dm_learned.h"
#include "v3dm_short.h"

namespace v3dm {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int kCQ = 256;   // queries per work-group of the single-pair pass (4 waves x 2 blocks of 32)
constexpr int kCT = 64;    // candidates per tile (2 blocks of 32)
constexpr int kCodeBytes = 32;

struct KeyStateI {
    int k[2];  // best key of the lane's two candidate groups
    int t[2];  // tile of each best key
};

// The two candidate blocks of tile t against one query block (d[cb], 16 keys per lane) into
// the two running maxima: 16 keys and the old maximum per group, eight v_max3_i32.
template <bool kFull>
__device__ __forceinline__ void key_tile_update(KeyStateI& s, v16i const (&d)[2], int t, int jBase, int nC, int h)
{
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        int const o = s.k[g];
        int x[16];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                int const rr = 8 * (g ^ cb) + r;
                int v = d[cb][rr];
                if (!kFull) v = (jBase + 32 * cb + 8 * (rr >> 2) + 4 * h + (rr & 3) < nC) ? v : INT_MIN;
                x[8 * cb + r] = v;
            }
        int const a0 = max(max(x[0], x[1]), x[2]);
        int const a1 = max(max(x[3], x[4]), x[5]);
        int const a2 = max(max(x[6], x[7]), x[8]);
        int const a3 = max(max(x[9], x[10]), x[11]);
        int const a4 = max(max(x[12], x[13]), x[14]);
        int const b0 = max(max(a0, a1), a2);
        int const b1 = max(max(a3, a4), x[15]);
        int k = max(max(b0, b1), o);
        // opaque to the optimiser: the test below stays k != o (one compare) instead of
        // being rewritten as max(b0, b1) > o, which costs a separate max for k
        __asm__ volatile("" : "+v"(k));
        s.t[g] = (k != o) ? t : s.t[g];
        s.k[g] = k;
    }
}

// The A operands of tile t: the tile's address is uniform (scalar registers), the lane's
// offset inside it (row c32 of each 32-row block, bytes 16h..16h+15) is fixed, so the
// loads take a scalar base and a constant vector offset.
__device__ __forceinline__ void load_code_tile(signed char const* __restrict__ C, int t, int lo, v4i (&a)[2])
{
    signed char const* tp = C + (size_t)t * (kCT * kCodeBytes);
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) a[cb] = *reinterpret_cast<v4i const*>(tp + lo + cb * 32 * kCodeBytes);
}

// The four 32x32 blocks of a tile, one query block at a time (two accumulators live).
template <bool kFull>
__device__ __forceinline__ void code_tile(v4i const (&a)[2], int t, v4i const (&bq)[2], v16i const& pos,
                                          KeyStateI (&st)[2], int nC, int h)
{
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        v16i d[2];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) d[cb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[cb], bq[qb], pos, 0, 0, 0);
        key_tile_update<kFull>(st[qb], d, t, t * kCT, nC, h);
    }
}

// One work unit: query block qblk (256 queries) against candidate tiles [split *
// tilesPerSplit, ...) -> part[((split * ldq) + q) * 2 + h]: code scores as floats and
// candidate indices (Cand2), or packed score and index inside the split (CandP).
template <int kQueries = kCQ, typename List>
__device__ __forceinline__ void code_block(signed char const* __restrict__ Q, int nQ,
                                           signed char const* __restrict__ C, int nC, int qblk, int split,
                                           int tilesPerSplit, List* __restrict__ part, int ldq)
{
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const q0 = qblk * kQueries + wave * 64;
    int const nTiles = (nC + kCT - 1) / kCT;
    int const t0 = split * tilesPerSplit;
    int const t1 = min(t0 + tilesPerSplit, nTiles);

    v4i bq[2];
#pragma unroll
    for (int qb = 0; qb < 2; ++qb)
        bq[qb] = *reinterpret_cast<v4i const*>(Q + (size_t)(q0 + qb * 32 + c32) * kCodeBytes + 16 * h);
    v16i pos;
#pragma unroll
    for (int r = 0; r < 16; ++r) pos[r] = r;
    // held in vector registers across the tile loop (as a constant it is rebuilt from
    // scalar registers for every MFMA)
#pragma unroll
    for (int r = 0; r < 16; ++r) __asm__ volatile("" : "+v"(pos[r]));
    int const lo = c32 * kCodeBytes + 16 * h;
    KeyStateI st[2];
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) st[qb] = KeyStateI{{INT_MIN, INT_MIN}, {-1, -1}};

    int const tFull = (t1 * kCT > nC) ? t1 - 1 : t1;  // the last tile may be partial
    // the next tile's A operands are loaded while the current tile is processed (the index
    // is clamped to the split's last tile, so every load stays in range); the scheduling
    // barriers keep each load ahead of the tile before it (left free, the compiler moved it
    // behind that tile's MFMAs into the same registers, so that every tile waited for its
    // own operands)
    v4i a[2], an[2];
    load_code_tile(C, t0, lo, a);
    int t = t0;
    for (; t + 1 < tFull; t += 2) {  // two tiles a round: the buffers alternate, no copies
        load_code_tile(C, t + 1, lo, an);
        __builtin_amdgcn_sched_barrier(0);
        code_tile<true>(a, t, bq, pos, st, nC, h);
        load_code_tile(C, min(t + 2, t1 - 1), lo, a);
        __builtin_amdgcn_sched_barrier(0);
        code_tile<true>(an, t + 1, bq, pos, st, nC, h);
    }
    // `a` holds tile t here
    if (t < tFull) {
        load_code_tile(C, min(t + 1, t1 - 1), lo, an);
        code_tile<true>(a, t, bq, pos, st, nC, h);
        if (tFull < t1) code_tile<false>(an, tFull, bq, pos, st, nC, h);
    } else if (tFull < t1) {
        code_tile<false>(a, tFull, bq, pos, st, nC, h);
    }

#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        int const q = q0 + qb * 32 + c32;
        if (q >= nQ) continue;
        // the list holds the best candidate of each group (two disjoint candidate sets):
        // element r of block 2 t + ((r >> 3) ^ g)
        int const r1 = st[qb].k[0] & 15, r2 = st[qb].k[1] & 15;
        bool const ok1 = st[qb].t[0] >= 0 && st[qb].k[0] != INT_MIN;
        bool const ok2 = st[qb].t[1] >= 0 && st[qb].k[1] != INT_MIN;
        int const j1 = (2 * st[qb].t[0] + (r1 >> 3)) * 32 + 8 * (r1 >> 2) + 4 * h + (r1 & 3);
        int const j2 = (2 * st[qb].t[1] + ((r2 >> 3) ^ 1)) * 32 + 8 * (r2 >> 2) + 4 * h + (r2 & 3);
        if constexpr (sizeof(List) == sizeof(CandP)) {
            int const j0 = t0 * kCT;  // first candidate of the split
            unsigned int const e = (unsigned int)(kPackedEmpty & 0xffff) << 16;
            CandP c;
            c.a = ok1 ? ((unsigned int)(st[qb].k[0] >> 4) << 16) | (unsigned int)(j1 - j0) : e;
            c.b = ok2 ? ((unsigned int)(st[qb].k[1] >> 4) << 16) | (unsigned int)(j2 - j0) : e;
            part[((size_t)split * ldq + q) * 2 + h] = c;
        } else {
            Cand2 c;
            c.v1 = ok1 ? (float)(st[qb].k[0] >> 4) : -INFINITY;
            c.v2 = ok2 ? (float)(st[qb].k[1] >> 4) : -INFINITY;
            c.j1 = ok1 ? j1 : -1;
            c.j2 = ok2 ? j2 : -1;
            part[((size_t)split * ldq + q) * 2 + h] = c;
        }
    }
}

// Both directions of a pair in one launch: work-groups [0, queryBlocks0 * splits0) run
// direction 0 (query block fastest), the rest direction 1.
struct CodeDir {
    signed char const* Q;
    signed char const* C;
    CandP* part;
    int nQ, nC, tilesPerSplit, ldq, queryBlocks, splits;
};

__global__ __launch_bounds__(256, 4) void k_pass_code2(CodeDir d0, CodeDir d1)
{
    int b = blockIdx.x;
    int const n0 = d0.queryBlocks * d0.splits;
    bool const second = b >= n0;  // uniform per work-group
    CodeDir const d = second ? d1 : d0;
    if (second) b -= n0;
    int const split = b / d.queryBlocks;
    code_block(d.Q, d.nQ, d.C, d.nC, b - split * d.queryBlocks, split, d.tilesPerSplit, d.part, d.ldq);
}

// Persistent form over a batch: item = (segment, query block, first split, tiles per split);
// a work-group runs the splits of the item's row from the first one on.  kQW waves of 64
// queries per work-group: the query block is kQW * 64 rows (the plan's tile shape).
template <int kQW>
__global__ __launch_bounds__(64 * kQW, 4) void k_batch_pass_code(BatchDev b, int4 const* __restrict__ items,
                                                                        int nItems, int* __restrict__ queue,
                                                                        Cand2* __restrict__ part)
{
    __shared__ int sItem;
    for (;;) {
        if (threadIdx.x == 0) sItem = atomicAdd(queue, 1);
        __syncthreads();
        int const it = __builtin_amdgcn_readfirstlane(sItem);
        __syncthreads();
        if (it >= nItems) break;
        int4 const w = items[it];
        BatchSeg const sg = b.seg[w.x];
        signed char const* Q = b.codes + (size_t)b.imgBase[sg.img] * kCodeBytes;
        signed char const* C = b.codes + (size_t)b.imgBase[sg.oImg] * kCodeBytes;
        for (int sp = w.z; sp < sg.nParts; ++sp)  // the item: every split of its row, in turn
            code_block<kQW * 64>(Q, sg.n, C, sg.nO, w.y, sp, w.w, part + (size_t)sg.partOff * 2, sg.ld);
    }
}

void launch_batch_pass_code(BatchDev const& b, int4 const* items, int nItems, int* queue, Cand2* part, int grid,
                            int queryRows, hipStream_t st)
{
    if (nItems <= 0) return;
    if (queryRows == 128)
        hipLaunchKernelGGL(k_batch_pass_code<2>, dim3(grid), dim3(128), 0, st, b, items, nItems, queue, part);
    else
        hipLaunchKernelGGL(k_batch_pass_code<4>, dim3(grid), dim3(256), 0, st, b, items, nItems, queue, part);
}

void launch_pass_code2(signed char const* A, int nA, signed char const* B, int nB, ShortPlan const& pa,
                       CandP* partA, int ldA, ShortPlan const& pb, CandP* partB, int ldB, hipStream_t st)
{
    if (nA <= 0 || nB <= 0) return;
    CodeDir const d0{A, B, partA, nA, nB, pa.tilesPerSplit, ldA, pa.queryBlocks, pa.splits};
    CodeDir const d1{B, A, partB, nB, nA, pb.tilesPerSplit, ldB, pb.queryBlocks, pb.splits};
    hipLaunchKernelGGL(k_pass_code2, dim3(pa.queryBlocks * pa.splits + pb.queryBlocks * pb.splits), dim3(256), 0, st,
                       d0, d1);
}


}  // namespace v3dm
