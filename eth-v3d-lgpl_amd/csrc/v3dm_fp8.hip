// FP8 (e4m3) shortlist form of the putative matcher.
//
// Each descriptor x is stored as e4m3(x * 2^e) with its own power-of-two scale 2^e chosen
// so that max|x| * 2^e <= 448 (the largest e4m3 normal).  The scores are computed with the
// block-scaled gfx950 instruction v_mfma_scale_f32_32x32x64_f8f6f4: every lane hands the
// E8M0 byte 127 - e of its row (A, candidates) or column (B, queries), and the matrix core
// applies 2^-e_c * 2^-e_q to the products, so the FP32 accumulators hold the dot products
// of the original descriptors up to e4m3 rounding.  The selection of the partition lists
// and the exact rescoring are those of the BF16 form (v3dm_short.hip).
#include "v3dm_device.h"
#include "v3dm_kernels.h"
#include "v3dm_short.h"

namespace v3dm {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void lds_void;

// ------------------------------------------------------------------------------ FP32 -> e4m3
// One wave per descriptor, 2 components per lane.  e = floor(log2(448 / max|x|)), clamped;
// scale byte = 127 - e.  Rows of zeros get e = 0.
__global__ __launch_bounds__(256) void k_to_fp8(float const* __restrict__ X, int n, unsigned char* __restrict__ Y,
                                                unsigned char* __restrict__ E)
{
    int const lane = threadIdx.x & 63;
    int const row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    float2 const v = reinterpret_cast<float2 const*>(X + (size_t)row * kDim)[lane];
    float m = fmaxf(fabsf(v.x), fabsf(v.y));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    int e = 0;
    if (m > 0.0f && m < 3.0e38f) {
        e = (int)floorf(log2f(448.0f / m));
        while (ldexpf(m, e) > 448.0f) --e;  // guard the rounding of log2f
        e = max(-100, min(100, e));
    }
    unsigned const packed = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(v.x, e), ldexpf(v.y, e), 0, false);
    reinterpret_cast<unsigned short*>(Y + (size_t)row * kDim)[lane] = (unsigned short)(packed & 0xffffu);
    if (lane == 0) E[row] = (unsigned char)(127 - e);
}

void launch_to_fp8(float const* X, int rows, unsigned char* Y, unsigned char* E, hipStream_t st)
{
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_to_fp8, dim3((rows + 3) / 4), dim3(256), 0, st, X, rows, Y, E);
}

// ------------------------------------------------------------------------------ one-direction pass
// Same work split as k_pass_bf16: 4 waves x 64 queries (B operands: lane l holds query l%32
// of its block, k = 64h + 32s + v for MFMA step s, h = l/32, 64 bytes per query), 64
// candidates per tile through two LDS stages filled by direct global->LDS loads.  Stage
// layout: row r, half h (64 bytes) is segment h*64 + r at 80*(h*64 + r) (16-byte pad:
// consecutive rows fall on different bank groups); the tile's 64 candidate scale bytes
// follow at byte 10240.  Per query and lane half the best index-packed key of each of two
// disjoint candidate groups is kept, as in the BF16 pass.
constexpr int kFQ = 256;
constexpr int kFT = 64;
constexpr int kSegF = 80;
constexpr int kStageLoadsF = 3;  // 1 KiB direct loads per wave per tile (4 waves)
constexpr int kStageBytesF = kStageLoadsF * 4 * 1024;
constexpr int kScaleOffF = 640 * 16;

// Running best key of the lane's two candidate groups (k1: group 0, k2: group 1) and the
// tile each was found in.
struct KeyStateF {
    float k1, k2;
    int t1, t2;
};

// Source of stage slot `slot`: descriptor bytes (offset within the tile) or, for the four
// slots after the 640 data slots, the tile's scale bytes (flag bit 30 set).
__device__ __forceinline__ int stage_src_fp8(int slot)
{
    if (slot < 640) {
        int const seg = slot / 5, within = slot - seg * 5;
        int const h = seg / kFT, row = seg - h * kFT;
        return row * 128 + h * 64 + (within < 4 ? within : 3) * 16;
    }
    if (slot < 644) return (1 << 30) | ((slot - 640) * 16);
    return 0;  // tail of the stage: never read
}

__device__ __forceinline__ void stage_tile_fp8(unsigned char const* __restrict__ C, unsigned char const* __restrict__ E,
                                               int t, unsigned char* stage, int const (&src)[kStageLoadsF], int wv)
{
    unsigned char const* tile = C + (size_t)t * kFT * kDim;
    unsigned char const* etile = E + (size_t)t * kFT;
#pragma unroll
    for (int p = 0; p < kStageLoadsF; ++p) {
        int const s = src[p];
        unsigned char const* g = (s & (1 << 30)) ? etile + (s & 0xffff) : tile + s;
        __builtin_amdgcn_global_load_lds(g, (lds_void*)(stage + (4 * p + wv) * 1024), 16, 0, 0);
    }
}

template <bool kFull>
__device__ __forceinline__ void tile_fp8(unsigned char const* stage, v8i const (&bq)[2][2], int const (&sq)[2],
                                         KeyStateF (&st)[2], int t, int nC, int h, int c32, float inf)
{
    v16f acc[2][2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        acc[cb][0] = v16f{};
        acc[cb][1] = v16f{};
        unsigned char const* rowp = stage + (h * kFT + cb * 32 + c32) * kSegF;
        int const sa = stage[kScaleOffF + cb * 32 + c32];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            v8i const a = *reinterpret_cast<v8i const*>(rowp + 32 * s);
            acc[cb][0] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, bq[0][s], acc[cb][0], 0, 0, 0, sa, 0,
                                                                         sq[0]);
            acc[cb][1] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, bq[1][s], acc[cb][1], 0, 0, 0, sa, 0,
                                                                         sq[1]);
        }
    }
    int const jt = t * kFT;
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

template <bool kFull>
__device__ __forceinline__ void run_tiles_fp8(unsigned char const* __restrict__ C, unsigned char const* __restrict__ E,
                                              unsigned char* stA, unsigned char* stB, int const (&src)[kStageLoadsF],
                                              v8i const (&bq)[2][2], int const (&sq)[2], KeyStateF (&st)[2], int t0,
                                              int t1, int nC, int h, int c32, int wv, float inf)
{
    for (int t = t0; t < t1; t += 2) {
        wait_lds_loads();
        __syncthreads();  // tile t is in stA; every wave is done with stB
        if (t + 1 < t1) stage_tile_fp8(C, E, t + 1, stB, src, wv);
        tile_fp8<kFull>(stA, bq, sq, st, t, nC, h, c32, inf);
        if (t + 1 >= t1) break;
        wait_lds_loads();
        __syncthreads();  // tile t+1 is in stB; every wave is done with stA
        if (t + 2 < t1) stage_tile_fp8(C, E, t + 2, stA, src, wv);
        tile_fp8<kFull>(stB, bq, sq, st, t + 1, nC, h, c32, inf);
    }
}

// One work unit: query block qblk against the candidate tiles of split `split`.
__device__ __forceinline__ void pass_fp8_block(unsigned char const* __restrict__ Q,
                                               unsigned char const* __restrict__ EQ, int nQ,
                                               unsigned char const* __restrict__ C,
                                               unsigned char const* __restrict__ EC, int nC, int qblk, int split,
                                               int tilesPerSplit, Cand2* __restrict__ part, int ldq,
                                               unsigned char* stA, unsigned char* stB, float inf)
{
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const wv = __builtin_amdgcn_readfirstlane(wave);
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const q0 = qblk * kFQ + wave * 64;
    int const nTiles = (nC + kFT - 1) / kFT;
    int const t0 = split * tilesPerSplit;
    int const t1 = min(t0 + tilesPerSplit, nTiles);

    int src[kStageLoadsF];
#pragma unroll
    for (int p = 0; p < kStageLoadsF; ++p) src[p] = stage_src_fp8((4 * p + wv) * 64 + lane);
    if (t0 < t1) stage_tile_fp8(C, EC, t0, stA, src, wv);

    v8i bq[2][2];
    int sq[2];
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) {
        int const q = q0 + qb * 32 + c32;
        uint4 const* qp = reinterpret_cast<uint4 const*>(Q + (size_t)q * kDim + h * 64);
        uint4 const u0 = qp[0], u1 = qp[1], u2 = qp[2], u3 = qp[3];
        bq[qb][0] = v8i{(int)u0.x, (int)u0.y, (int)u0.z, (int)u0.w, (int)u1.x, (int)u1.y, (int)u1.z, (int)u1.w};
        bq[qb][1] = v8i{(int)u2.x, (int)u2.y, (int)u2.z, (int)u2.w, (int)u3.x, (int)u3.y, (int)u3.z, (int)u3.w};
        sq[qb] = EQ[q];
    }

    KeyStateF st[2];
#pragma unroll
    for (int qb = 0; qb < 2; ++qb) st[qb] = KeyStateF{-INFINITY, -INFINITY, -1, -1};

    if (t1 * kFT > nC) run_tiles_fp8<false>(C, EC, stA, stB, src, bq, sq, st, t0, t1, nC, h, c32, wv, inf);
    else run_tiles_fp8<true>(C, EC, stA, stB, src, bq, sq, st, t0, t1, nC, h, c32, wv, inf);

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
                   ? st[qb].t1 * kFT + (p1 >> 4) * 32 + 8 * ((p1 & 15) >> 2) + 4 * h + (p1 & 3) : -1;
        c.j2 = (st[qb].t2 >= 0 && c.v2 > -1e29f)
                   ? st[qb].t2 * kFT + (p2 >> 4) * 32 + 8 * ((p2 & 15) >> 2) + 4 * h + (p2 & 3) : -1;
        part[((size_t)split * ldq + q) * 2 + h] = c;
    }
}

// Both directions of a pair in one launch: work-groups [0, queryBlocks0 * splits0) run
// direction 0 (query block fastest), the rest direction 1.
struct Fp8Dir {
    unsigned char const* Q;
    unsigned char const* EQ;
    unsigned char const* C;
    unsigned char const* EC;
    Cand2* part;
    int nQ, nC, tilesPerSplit, ldq, queryBlocks, splits;
};

__global__ __launch_bounds__(256, 2) void k_pass_fp8(Fp8Dir d0, Fp8Dir d1, float inf)
{
    __shared__ __attribute__((aligned(1024))) unsigned char stA[kStageBytesF];
    __shared__ __attribute__((aligned(1024))) unsigned char stB[kStageBytesF];
    int b = blockIdx.x;
    int const n0 = d0.queryBlocks * d0.splits;
    bool const second = b >= n0;  // uniform per work-group
    Fp8Dir const d = second ? d1 : d0;
    if (second) b -= n0;
    int const split = b / d.queryBlocks;
    pass_fp8_block(d.Q, d.EQ, d.nQ, d.C, d.EC, d.nC, b - split * d.queryBlocks, split, d.tilesPerSplit, d.part,
                   d.ldq, stA, stB, inf);
}

void launch_pass_fp8_2(unsigned char const* A, unsigned char const* EA, int nA, unsigned char const* B,
                       unsigned char const* EB, int nB, ShortPlan const& pa, Cand2* partA, int ldA, ShortPlan const& pb,
                       Cand2* partB, int ldB, hipStream_t st)
{
    if (nA <= 0 || nB <= 0) return;
    Fp8Dir const d0{A, EA, B, EB, partA, nA, nB, pa.tilesPerSplit, ldA, pa.queryBlocks, pa.splits};
    Fp8Dir const d1{B, EB, A, EA, partB, nB, nA, pb.tilesPerSplit, ldB, pb.queryBlocks, pb.splits};
    hipLaunchKernelGGL(k_pass_fp8, dim3(pa.queryBlocks * pa.splits + pb.queryBlocks * pb.splits), dim3(256), 0, st,
                       d0, d1, INFINITY);
}

}  // namespace v3dm
