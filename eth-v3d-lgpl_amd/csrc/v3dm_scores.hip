// Score blocks for measurement: the approximate scores each form's pass computes, written
// out for the first nRows left x nCols right features of a handle (v3dm_score_block).
//
// Every form's pass keeps its scores in accumulators and never writes them; the kernels
// here use the same instruction, operand rounding and k order, one 32x32 block per wave,
// operands read straight from the FP32 rows:
//  * FP32: v_mfma_f32_32x32x2_f32, lane (c32, h) supplies k = 64h + s at step s = 0..63
//    (k_fused_fp32);
//  * BF16: v_mfma_f32_32x32x16_bf16 on round-to-nearest-even bf16 operands, k = 64h + 8s + v
//    at step s = 0..7 (k_pass_bf16, operands as k_to_bf16 rounds them);
//  * FP8: v_mfma_scale_f32_32x32x64_f8f6f4 on e4m3 operands with per-row power-of-two
//    scales chosen as k_to_fp8 chooses them, k = 64h + 32s + v at step s = 0..1.
// The score of left i and right j lands in D element r of lane l with i = 32 bi + l % 32
// and j = 32 bj + 8 (r / 4) + 4 (l / 32) + r % 4.
#include "v3dm_device.h"
#include "v3dm_kernels.h"

namespace v3dm {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));

// e4m3 operands of one lane: 64 components of row x (k = 64h .. 64h+63) scaled by 2^e, e as
// k_to_fp8 chooses it from the row's max |x| (the two lane halves together)
__device__ __forceinline__ void fp8_operand(float const* x, int h, v8i (&op)[2], int& scaleByte)
{
    float const* xp = x + 64 * h;
    float m = 0.0f;
    for (int k = 0; k < 64; ++k) m = fmaxf(m, fabsf(xp[k]));
    m = fmaxf(m, __shfl_xor(m, 32));
    int e = 0;
    if (m > 0.0f && m < 3.0e38f) {
        e = (int)floorf(log2f(448.0f / m));
        while (ldexpf(m, e) > 448.0f) --e;
        e = max(-100, min(100, e));
    }
    scaleByte = 127 - e;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int w = 0; w < 8; ++w) {
            float const* p = xp + 32 * s + 4 * w;
            unsigned lo = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(p[0], e), ldexpf(p[1], e), 0, false);
            unsigned hi = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(p[2], e), ldexpf(p[3], e), 0, false);
            op[s][w] = (int)((lo & 0xffffu) | (hi << 16));
        }
}

template <int kMode>
__global__ __launch_bounds__(64) void k_score_block(float const* __restrict__ L, float const* __restrict__ R,
                                                    int nRows, int nCols, float* __restrict__ out)
{
    int const lane = threadIdx.x;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    float const* q = L + (size_t)(i0 + c32) * kDim;  // B operand: left feature (column)
    float const* c = R + (size_t)(j0 + c32) * kDim;  // A operand: right feature (row)
    v16f d = v16f{};
    if (kMode == 0) {
        for (int s = 0; s < 64; ++s)
            d = __builtin_amdgcn_mfma_f32_32x32x2f32(c[64 * h + s], q[64 * h + s], d, 0, 0, 0);
    } else if (kMode == 1) {
        for (int s = 0; s < 8; ++s) {
            v8bf a, b;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                a[v] = (__bf16)c[64 * h + 8 * s + v];
                b[v] = (__bf16)q[64 * h + 8 * s + v];
            }
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d, 0, 0, 0);
        }
    } else {
        v8i a[2], b[2];
        int sa, sq;
        fp8_operand(c, h, a, sa);
        fp8_operand(q, h, b, sq);
#pragma unroll
        for (int s = 0; s < 2; ++s)
            d = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[s], b[s], d, 0, 0, 0, sa, 0, sq);
    }
    int const i = i0 + c32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        int const j = j0 + 8 * (r >> 2) + 4 * h + (r & 3);
        if (i < nRows && j < nCols) out[(size_t)i * nCols + j] = d[r];
    }
}

void launch_score_block(int mode, float const* L, float const* R, int nRows, int nCols, float* out, hipStream_t st)
{
    dim3 const grid((nCols + 31) / 32, (nRows + 31) / 32);
    if (mode == 1)
        hipLaunchKernelGGL(k_score_block<1>, grid, dim3(64), 0, st, L, R, nRows, nCols, out);
    else if (mode == 2)
        hipLaunchKernelGGL(k_score_block<2>, grid, dim3(64), 0, st, L, R, nRows, nCols, out);
    else
        hipLaunchKernelGGL(k_score_block<0>, grid, dim3(64), 0, st, L, R, nRows, nCols, out);
}

}  // namespace v3dm
