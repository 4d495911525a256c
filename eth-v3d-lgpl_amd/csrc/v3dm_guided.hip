// Guided matching masks on gfx950: the geometric predicate of runGuidedMatching /
// runGuidedHomographyMatching (V3D/CUDA/v3d_cudamatching.cu:554-694) evaluated once per
// (left, right) pair into a bitmap of removed pairs.  The reference overwrites the removed
// scores of its stored score matrix with -1e20 (cu:308, cu:340); here the matrix is never
// stored, so the guided score pass rescores the descriptors and substitutes -1e20 where a
// bit is set.  Masks of repeated guided calls are OR-ed into the bitmap, which is what
// overwriting the same stored matrix again amounts to.
//
// Arithmetic: plain IEEE FP32 in the reference's operand order, no contraction (mul_nc /
// add_nc / sub_nc), correctly rounded division.
#include "v3dm_device.h"
#include "v3dm_kernels.h"

namespace v3dm {

constexpr int kGuidedRows = 256;  // right features per work-group (8 bitmap words)

struct Mat3f {
    float m[9];
};

__device__ __forceinline__ float fm(float a, float b) { return mul_nc(a, b); }
__device__ __forceinline__ float fa(float a, float b) { return add_nc(a, b); }
__device__ __forceinline__ float fs(float a, float b) { return sub_nc(a, b); }

// Work-group: 256 left features (one per thread) x 256 right features staged in LDS.
// Epipolar (cu:277-311): Fp = F p per left feature; per right feature q the squares of
// the first two components of F^T q; Sampson distance (q.Fp)^2 / (Fp.x^2 + Fp.y^2 +
// Ftq.x^2 + Ftq.y^2) summed left to right.  Homography (cu:313-343): Hp = H p divided by
// its third component; squared distance (Hp.x - q.x)^2 + (Hp.y - q.y)^2.  A pair is
// removed when the distance is > sqrThreshold (a NaN distance keeps the pair).
template <int kKind>
__global__ __launch_bounds__(256) void k_guided_mask(Mat3f M, float sqrThreshold, float const* __restrict__ leftXY,
                                                     float const* __restrict__ rightXY, int nL, int nR,
                                                     unsigned* __restrict__ bits, int ldi, int first)
{
    __shared__ float4 rq[kGuidedRows];
    float const* m = M.m;
    int const tid = threadIdx.x;
    int const j0 = blockIdx.y * kGuidedRows;
    {
        int const j = j0 + tid;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < nR) {
            float2 const q = reinterpret_cast<float2 const*>(rightXY)[j];
            if (kKind == 0) {
                // Ftq.x = F1.x q.x + F2.x q.y + F3.x, Ftq.y = F1.y q.x + F2.y q.y + F3.y
                float const tx = fa(fa(fm(m[0], q.x), fm(m[3], q.y)), m[6]);
                float const ty = fa(fa(fm(m[1], q.x), fm(m[4], q.y)), m[7]);
                v = make_float4(q.x, q.y, fm(tx, tx), fm(ty, ty));
            } else {
                v = make_float4(q.x, q.y, 0.f, 0.f);
            }
        }
        rq[tid] = v;
    }
    __syncthreads();
    int const i = blockIdx.x * 256 + tid;
    if (i >= nL) return;
    float2 const p = reinterpret_cast<float2 const*>(leftXY)[i];
    float ax = fa(fa(fm(m[0], p.x), fm(m[1], p.y)), m[2]);
    float ay = fa(fa(fm(m[3], p.x), fm(m[4], p.y)), m[5]);
    float const az = fa(fa(fm(m[6], p.x), fm(m[7], p.y)), m[8]);
    float a2 = 0.f;
    if (kKind == 0) {
        a2 = fa(fm(ax, ax), fm(ay, ay));
    } else {
        ax = __fdiv_rn(ax, az);
        ay = __fdiv_rn(ay, az);
    }
    int const nw = min(kGuidedRows / 32, (nR - j0 + 31) / 32);
    for (int w = 0; w < nw; ++w) {
        unsigned word = 0;
#pragma unroll 8
        for (int b = 0; b < 32; ++b) {
            float4 const q = rq[w * 32 + b];  // same address in every lane: a broadcast
            bool removed;
            if (kKind == 0) {
                float num = fa(fa(fm(q.x, ax), fm(q.y, ay)), az);
                num = fm(num, num);
                float const denom = fa(fa(a2, q.z), q.w);
                removed = __fdiv_rn(num, denom) > sqrThreshold;
            } else {
                float const dx = fs(ax, q.x), dy = fs(ay, q.y);
                removed = fa(fm(dx, dx), fm(dy, dy)) > sqrThreshold;
            }
            removed = removed && (j0 + w * 32 + b < nR);
            word |= removed ? (1u << b) : 0u;
        }
        size_t const o = (size_t)(blockIdx.y * (kGuidedRows / 32) + w) * ldi + i;
        bits[o] = first ? word : (bits[o] | word);
    }
}

void launch_guided_mask(int kind, float const (&m)[9], float sqrThreshold, float const* leftXY,
                        float const* rightXY, int nL, int nR, unsigned* bits, int ldi, bool first, hipStream_t st)
{
    if (nL <= 0 || nR <= 0) return;
    Mat3f M;
    for (int t = 0; t < 9; ++t) M.m[t] = m[t];
    dim3 const grid((nL + 255) / 256, (nR + kGuidedRows - 1) / kGuidedRows);
    if (kind == 0)
        hipLaunchKernelGGL(k_guided_mask<0>, grid, dim3(256), 0, st, M, sqrThreshold, leftXY, rightXY, nL, nR, bits,
                           ldi, first ? 1 : 0);
    else
        hipLaunchKernelGGL(k_guided_mask<1>, grid, dim3(256), 0, st, M, sqrThreshold, leftXY, rightXY, nL, nR, bits,
                           ldi, first ? 1 : 0);
}

}  // namespace v3dm
