// Vocabulary-tree quantisation on gfx950: the greedy descent of
// VocabularyTree::findBestLeafIndex (V3D/Base/v3d_voctree.h:63-99) for many keys at once,
// level by level.
//
// Per level the keys are bucketed by their current node (a counting sort into buckets
// padded to 32 keys), so that a wave takes 32 keys with the same parent and scores them
// against the parent's children with v_mfma_f32_32x32x2_f32: the children are the A
// operand (32 per block, read straight from L2), the keys the resident B operand (64
// VGPRs per lane, as in the matcher's fused kernel).  The reference's choice is the first
// child of maximal sequential-FP32 score; the MFMA score differs from that score by at most
// eps = 2^-15 * ||key|| * max||node|| (a 128-term FP32 dot product in any order errs by
// <= ~129 * 2^-24 * sum|k_i c_i| <= 2^-17 * ||k|| ||c|| against the exact value; two of
// them differ by less than 2^-16 ||k|| ||c||), so a key whose best approximate score leads
// the second by more than 2 eps has the reference's child.  The other keys (near ties,
// non-finite values, scores near the -1e30 floor) are decided exactly: sequential FP32
// sums with IEEE products, the scan's strict '>' (first index on ties, NaN never wins).
#include <algorithm>

#include "v3dm_device.h"
#include "v3dm_voctree.h"

namespace v3dm {

constexpr float kVtEpsRel = 3.0517578125e-05f;  // 2^-15

__global__ void k_vt_reset(int* counters, int all)
{
    if (threadIdx.x < kVtCounters && (all || threadIdx.x != kVtCntFlagged)) counters[threadIdx.x] = 0;
}

void launch_vt_reset(int* counters, bool all, hipStream_t st)
{
    hipLaunchKernelGGL(k_vt_reset, dim3(1), dim3(64), 0, st, counters, all ? 1 : 0);
}

// ------------------------------------------------------------------------------ bucketing
__global__ __launch_bounds__(256) void k_vt_zero(int* a, int n, int value)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a[i] = value;
}

// Bucket counts live kVtBinStride ints apart (one L2 line per bucket): the global atomics
// of different buckets then never share a line.  For up to kVtLdsBins parents every
// work-group first counts its kVtKeysPerGroup keys in LDS and adds one total per bucket it
// saw; the scatter reserves each bucket's range for the work-group with one atomic and
// places the keys by their LDS rank.  Wider parent levels use one global atomic per key.
constexpr int kVtKeysPerThread = 16;
constexpr int kVtKeysPerGroup = 256 * kVtKeysPerThread;

__device__ __forceinline__ int vt_bucket(int const* __restrict__ node, int k, long long parentStart, int nParents,
                                         int* __restrict__ exactList, int* __restrict__ counters)
{
    int const nd = node[k];
    if (nd == kVtNone) return -1;
    long long const b = (long long)nd - parentStart;
    if (b >= 0 && b < nParents) return (int)b;
    exactList[atomicAdd(&counters[kVtCntExact], 1)] = k;  // not a node of the parent level
    return -1;
}

__global__ __launch_bounds__(256) void k_vt_count_lds(int const* __restrict__ node, int n, long long parentStart,
                                                      int nParents, int* __restrict__ cnt,
                                                      int* __restrict__ exactList, int* __restrict__ counters)
{
    extern __shared__ int hist[];
    int const tid = threadIdx.x;
    for (int b = tid; b < nParents; b += 256) hist[b] = 0;
    __syncthreads();
    int const k0 = blockIdx.x * kVtKeysPerGroup;
    for (int i = 0; i < kVtKeysPerThread; ++i) {
        int const k = k0 + i * 256 + tid;
        if (k >= n) break;
        int const b = vt_bucket(node, k, parentStart, nParents, exactList, counters);
        if (b >= 0) atomicAdd(&hist[b], 1);
    }
    __syncthreads();
    for (int b = tid; b < nParents; b += 256) {
        int const c = hist[b];
        if (c) atomicAdd(&cnt[(size_t)b * kVtBinStride], c);
    }
}

__global__ __launch_bounds__(256) void k_vt_scatter_lds(int const* __restrict__ node, int n, long long parentStart,
                                                        int nParents, int* __restrict__ cursor,
                                                        int* __restrict__ order)
{
    extern __shared__ int hist[];
    int const tid = threadIdx.x;
    for (int b = tid; b < nParents; b += 256) hist[b] = 0;
    __syncthreads();
    int const k0 = blockIdx.x * kVtKeysPerGroup;
    int bin[kVtKeysPerThread], rank[kVtKeysPerThread];
#pragma unroll
    for (int i = 0; i < kVtKeysPerThread; ++i) {
        int const k = k0 + i * 256 + tid;
        bin[i] = -1;
        if (k < n) {
            int const nd = node[k];
            long long const b = (long long)nd - parentStart;
            if (nd != kVtNone && b >= 0 && b < nParents) {
                bin[i] = (int)b;
                rank[i] = atomicAdd(&hist[b], 1);
            }
        }
    }
    __syncthreads();
    for (int b = tid; b < nParents; b += 256) {
        int const c = hist[b];
        if (c) hist[b] = atomicAdd(&cursor[(size_t)b * kVtBinStride], c);  // the group's range in the bucket
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kVtKeysPerThread; ++i)
        if (bin[i] >= 0) order[hist[bin[i]] + rank[i]] = k0 + i * 256 + tid;
}

__global__ __launch_bounds__(256) void k_vt_count(int const* __restrict__ node, int n, long long parentStart,
                                                  int nParents, int* __restrict__ cnt, int* __restrict__ exactList,
                                                  int* __restrict__ counters)
{
    int const k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int const b = vt_bucket(node, k, parentStart, nParents, exactList, counters);
    if (b >= 0) atomicAdd(&cnt[(size_t)b * kVtBinStride], 1);
}

// One work-group: exclusive prefix of the bucket sizes rounded up to 32 (so that no slot of
// 32 keys spans two buckets), slot count for the score kernel.
__global__ __launch_bounds__(1024) void k_vt_scan(int const* __restrict__ cnt, int nParents, int* __restrict__ cursor,
                                                  int* __restrict__ counters)
{
    __shared__ int part[1024];
    int const t = threadIdx.x;
    int const per = (nParents + 1023) / 1024;
    int const b0 = t * per, b1 = min(b0 + per, nParents);
    int s = 0;
    for (int b = b0; b < b1; ++b) s += (cnt[(size_t)b * kVtBinStride] + 31) & ~31;
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
        int const v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int b = b0; b < b1; ++b) {
        cursor[(size_t)b * kVtBinStride] = run;
        run += (cnt[(size_t)b * kVtBinStride] + 31) & ~31;
    }
    if (t == 1023) counters[kVtCntSlots] = part[1023] / 32;
}

__global__ __launch_bounds__(256) void k_vt_scatter(int const* __restrict__ node, int n, long long parentStart,
                                                    int nParents, int* __restrict__ cursor, int* __restrict__ order)
{
    int const k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int const nd = node[k];
    if (nd == kVtNone) return;
    long long const b = (long long)nd - parentStart;
    if (b >= 0 && b < nParents) order[atomicAdd(&cursor[(size_t)b * kVtBinStride], 1)] = k;
}

void launch_vt_bucket(int const* node, int n, VtLevel const& lv, int* cnt, int* cursor, int* order, int orderCap,
                      int* exactList, int* counters, hipStream_t st)
{
    int const binInts = lv.nParents * kVtBinStride;
    hipLaunchKernelGGL(k_vt_zero, dim3(std::min((binInts + 255) / 256, 4096)), dim3(256), 0, st, cnt, binInts, 0);
    hipLaunchKernelGGL(k_vt_zero, dim3(std::min((orderCap + 255) / 256, 4096)), dim3(256), 0, st, order, orderCap,
                       -1);
    if (lv.nParents <= kVtLdsBins) {
        int const g = (n + kVtKeysPerGroup - 1) / kVtKeysPerGroup;
        size_t const lds = (size_t)lv.nParents * sizeof(int);
        hipLaunchKernelGGL(k_vt_count_lds, dim3(g), dim3(256), lds, st, node, n, lv.parentStart, lv.nParents, cnt,
                           exactList, counters);
        hipLaunchKernelGGL(k_vt_scan, dim3(1), dim3(1024), 0, st, cnt, lv.nParents, cursor, counters);
        hipLaunchKernelGGL(k_vt_scatter_lds, dim3(g), dim3(256), lds, st, node, n, lv.parentStart, lv.nParents,
                           cursor, order);
    } else {
        int const g = (n + 255) / 256;
        hipLaunchKernelGGL(k_vt_count, dim3(g), dim3(256), 0, st, node, n, lv.parentStart, lv.nParents, cnt,
                           exactList, counters);
        hipLaunchKernelGGL(k_vt_scan, dim3(1), dim3(1024), 0, st, cnt, lv.nParents, cursor, counters);
        hipLaunchKernelGGL(k_vt_scatter, dim3(g), dim3(256), 0, st, node, n, lv.parentStart, lv.nParents, cursor,
                           order);
    }
}

// ------------------------------------------------------------------------------ certified scores
// Wave = one slot of 32 keys with the same parent.  Lane l holds key l % 32, dims
// [64 h, 64 h + 64) (h = l / 32) as the B operand; the A operand of MFMA step 4 s4 + q is
// dim 64 h + 4 s4 + q of child row l % 32 of the block (same k map on both sides).
// Accumulator element r of lane l: child 8 (r / 4) + 4 h + r % 4 of the block, key l % 32.
__global__ __launch_bounds__(256) void k_vt_score(float const* __restrict__ keys, float const* __restrict__ keyNorm,
                                                  int n, float const* __restrict__ nodes, float maxNodeNorm,
                                                  int branch, int identity, int const* __restrict__ order,
                                                  int* __restrict__ node, int* __restrict__ exactList,
                                                  int* __restrict__ counters, float inf)
{
    int const lane = threadIdx.x & 63;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const slot = blockIdx.x * 4 + (threadIdx.x >> 6);
    int const nSlots = identity ? (n + 31) / 32 : counters[kVtCntSlots];
    if (slot >= nSlots) return;  // uniform per wave
    int kk;
    long long c0;
    if (identity) {
        kk = (slot * 32 + c32 < n) ? slot * 32 + c32 : -1;
        c0 = 0;
    } else {
        kk = order[slot * 32 + c32];
        c0 = ((long long)node[order[slot * 32]] + 1) * branch;  // a slot's first entry is a key
    }
    float bfrag[64];
    if (kk >= 0) {
        float4 const* kp = reinterpret_cast<float4 const*>(keys + (size_t)kk * kDim + h * 64);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            float4 const v = kp[q];
            bfrag[4 * q + 0] = v.x;
            bfrag[4 * q + 1] = v.y;
            bfrag[4 * q + 2] = v.z;
            bfrag[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 64; ++q) bfrag[q] = 0.0f;
    }
    Top2 s = top2_init();
    int const nb = (branch + 31) / 32;
    for (int cb = 0; cb < nb; ++cb) {
        // rows past the last child exist (the node array has 32 zero rows of padding)
        float4 const* ap = reinterpret_cast<float4 const*>(nodes + (size_t)(c0 + cb * 32 + c32) * kDim + h * 64);
        float4 a[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] = ap[q];
        v16f acc = {};
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].x, bfrag[4 * q + 0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].y, bfrag[4 * q + 1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].z, bfrag[4 * q + 2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].w, bfrag[4 * q + 3], acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            int const j = cb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            top2_push_fast(s, (j < branch) ? acc[r] : -INFINITY, j, inf);
        }
    }
    s = top2_merge(s, top2_shfl_xor(s, 32));
    if (h != 0 || kk < 0) return;
    float const eps = kVtEpsRel * keyNorm[kk] * maxNodeNorm + 1e-30f;
    bool const settled = (eps < 3.0e38f) && (s.b1 == s.b1) && (s.b2 == s.b2) && s.a1 >= 0 &&
                         (s.b1 - s.b2 > 2.0f * eps) && (s.b1 - eps > -1e30f);
    if (settled) node[kk] = (int)(c0 + s.a1);
    else exactList[atomicAdd(&counters[kVtCntExact], 1)] = kk;
}

void launch_vt_score(float const* keys, float const* keyNorm, int n, float const* nodes, float maxNodeNorm,
                     VtLevel const& lv, int const* order, bool identity, int maxSlots, int* node, int* exactList,
                     int* counters, hipStream_t st)
{
    if (n <= 0 || maxSlots <= 0) return;
    hipLaunchKernelGGL(k_vt_score, dim3((maxSlots + 3) / 4), dim3(256), 0, st, keys, keyNorm, n, nodes, maxNodeNorm,
                       lv.branch, identity ? 1 : 0, order, node, exactList, counters, INFINITY);
}

// ------------------------------------------------------------------------------ exact decisions
// One wave per listed key: lane j scans children j, j + 64, ... with the reference's
// sequential FP32 score (IEEE product, then IEEE sum, ascending dims) and strict '>'
// against its running best (start: the -1e30 floor); the lanes' bests combine to the
// first child of maximal score.  No valid child gives -1, so the key moves to node
// c0 - 1, as the reference's leafId = nextLeafId + bestNode does.
__global__ __launch_bounds__(256) void k_vt_exact(float const* __restrict__ keys, float const* __restrict__ nodes,
                                                  long long nNodes, int branch, int level,
                                                  int const* __restrict__ exactList, int* __restrict__ node,
                                                  int* __restrict__ counters)
{
    __shared__ float keyRow[4][kDim];
    int const lane = threadIdx.x & 63;
    int const cnt = counters[kVtCntExact];
    for (int e = blockIdx.x * 4 + (threadIdx.x >> 6); e < cnt; e += gridDim.x * 4) {
        int const kk = exactList[e];
        long long const c0 = (level == 0) ? 0 : ((long long)node[kk] + 1) * branch;
        if (c0 < 0 || c0 + branch > nNodes) {
            if (lane == 0) node[kk] = kVtNone;
            continue;
        }
        // the key in LDS (one row per wave), each child row in registers 32 floats at a
        // time: the loads are issued together instead of one per step of the sum
        float* kw = keyRow[threadIdx.x >> 6];
        __builtin_amdgcn_wave_barrier();  // the previous key's reads are done
        reinterpret_cast<float2*>(kw)[lane] = reinterpret_cast<float2 const*>(keys + (size_t)kk * kDim)[lane];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        float bestS = -INFINITY;
        int bestJ = -1;
        for (int j = lane; j < branch; j += 64) {
            float4 const* c = reinterpret_cast<float4 const*>(nodes + (size_t)(c0 + j) * kDim);
            float score = 0.0f;
            for (int q0 = 0; q0 < 32; q0 += 8) {
                float4 v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = c[q0 + u];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    float const* kq = kw + 4 * (q0 + u);
                    score = add_nc(score, mul_nc(kq[0], v[u].x));
                    score = add_nc(score, mul_nc(kq[1], v[u].y));
                    score = add_nc(score, mul_nc(kq[2], v[u].z));
                    score = add_nc(score, mul_nc(kq[3], v[u].w));
                }
            }
            if (score > -1e30f && score > bestS) {
                bestS = score;
                bestJ = j;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            float const os = __shfl_xor(bestS, o);
            int const oj = __shfl_xor(bestJ, o);
            bool const take = oj >= 0 && (bestJ < 0 || os > bestS || (os == bestS && oj < bestJ));
            bestS = take ? os : bestS;
            bestJ = take ? oj : bestJ;
        }
        if (lane == 0) {
            node[kk] = (int)(c0 + bestJ);
            atomicAdd(&counters[kVtCntFlagged], 1);
        }
    }
}

void launch_vt_exact(float const* keys, float const* nodes, VtLevel const& lv, int const* exactList, int* node,
                     int* counters, hipStream_t st)
{
    hipLaunchKernelGGL(k_vt_exact, dim3(512), dim3(256), 0, st, keys, nodes, lv.nNodes, lv.branch, lv.level,
                       exactList, node, counters);
}

__global__ __launch_bounds__(256) void k_vt_finish(int const* __restrict__ node, int n, long long leafStart,
                                                   int globalIds, int32_t* __restrict__ out)
{
    int const k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    int const nd = node[k];
    out[k] = (nd == kVtNone || globalIds) ? nd : (int32_t)((long long)nd - leafStart);
}

void launch_vt_finish(int const* node, int n, long long leafStart, bool globalIds, int32_t* out, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_vt_finish, dim3((n + 255) / 256), dim3(256), 0, st, node, n, leafStart, globalIds ? 1 : 0,
                       out);
}

}  // namespace v3dm
