// Device-side helpers shared by the matcher kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace v3dm {

constexpr int kDim = 128;          // descriptor length
constexpr int kFlag = -2;          // decision still open: re-checked exactly
constexpr float kInit = -1e30f;    // initial best / second best (cu:138-140)

// Best, second best (multiset order statistics) and index of the best of one query.
struct Top2 {
    float b1, b2;
    int a1;
};

__device__ __forceinline__ Top2 top2_init() { return Top2{kInit, kInit, -1}; }


// One candidate, the update rule of cu:145-147.  NaN scores are skipped.
__device__ __forceinline__ void top2_push(Top2& s, float v, int idx)
{
    bool const gt = v > s.b1;
    s.a1 = gt ? idx : s.a1;
    s.b2 = gt ? s.b1 : fmaxf(v, s.b2);
    s.b1 = fmaxf(s.b1, v);
}

// max / second-largest without the operand canonicalisation fmaxf needs for NaN:
// v_med3_f32(a, b, +inf) = max(a, b) and, for b1 >= b2, med3(b1, b2, v) is the new
// second best.  `inf` must be +infinity supplied at run time (a kernel argument): with a
// literal the compiler folds med3 back into a canonicalising max.  Used on approximate
// scores only; a NaN score there comes from a non-finite descriptor, whose error bound is
// non-finite, so its queries are re-checked.
__device__ __forceinline__ float fast_max(float a, float b, float inf) { return __builtin_amdgcn_fmed3f(a, b, inf); }

__device__ __forceinline__ void top2_push_fast(Top2& s, float v, int idx, float inf)
{
    bool const gt = v > s.b1;
    s.a1 = gt ? idx : s.a1;
    s.b2 = __builtin_amdgcn_fmed3f(s.b1, s.b2, v);
    s.b1 = fast_max(s.b1, v, inf);
}

// Order-independent merge of two partial states: larger best wins, an equal best keeps
// the smaller index (the first index of an ascending scan), the second best is the
// second largest of the union.  Commutative and associative.
__device__ __forceinline__ Top2 top2_merge(Top2 a, Top2 b)
{
    bool const takeB = (b.b1 > a.b1) || (b.b1 == a.b1 && b.a1 >= 0 && (a.a1 < 0 || b.a1 < a.a1));
    Top2 r;
    r.a1 = takeB ? b.a1 : a.a1;
    r.b2 = fmaxf(fminf(a.b1, b.b1), fmaxf(a.b2, b.b2));
    r.b1 = fmaxf(a.b1, b.b1);
    return r;
}

__device__ __forceinline__ Top2 top2_shfl_xor(Top2 s, int mask)
{
    Top2 o;
    o.b1 = __shfl_xor(s.b1, mask);
    o.b2 = __shfl_xor(s.b2, mask);
    o.a1 = __shfl_xor(s.a1, mask);
    return o;
}

// Decision of cu:150-161 (ratio) / cu:48 (minRatio < 0), IEEE arithmetic.
__device__ __forceinline__ bool accept_rule(float b1, float b2, float minScore, float minRatio)
{
    if (minRatio < 0.0f) return b1 > minScore;
    float const nextDist = (float)fmax(0.001, (double)(2.0f - 2.0f * b2));
    float const bestDist = 2.0f - 2.0f * b1;
    float const ratio = bestDist / nextDist;
    return b1 > minScore && ratio < minRatio;
}

__device__ __forceinline__ int decide_exact(Top2 s, float minScore, float minRatio)
{
    return accept_rule(s.b1, s.b2, minScore, minRatio) ? s.a1 : -1;
}

// Decision from approximate scores whose error against the exact scores is at most eps.
// Returns the exact decision when every exact score set within the error box gives the
// same answer, kFlag otherwise.  The rule is monotone (accept grows with b1 and falls
// with b2) when 2 - 2*b1 > 0 over the box, so the two corners settle it; outside that
// regime, with a non-finite bound, or with a best/second margin that could reorder the
// argmax, the query is flagged.
__device__ __forceinline__ int decide_certified(Top2 s, float eps, float minScore, float minRatio)
{
    if (!(eps < 3.0e38f) || !(s.b1 == s.b1) || !(s.b2 == s.b2)) return kFlag;
    float const lo1 = s.b1 - eps, hi1 = s.b1 + eps;
    float const lo2 = s.b2 - eps, hi2 = s.b2 + eps;
    if (minRatio >= 0.0f && !(hi1 < 1.0f)) return kFlag;
    if (!accept_rule(hi1, lo2, minScore, minRatio)) return -1;
    if (accept_rule(lo1, hi2, minScore, minRatio) && (s.b1 - s.b2 > 2.0f * eps) && s.a1 >= 0) return s.a1;
    return kFlag;
}

// IEEE FP32 product / sum / difference that the compiler may not fuse into an FMA, for
// code that must reproduce the reference's plain float arithmetic.  (__fmul_rn / __fadd_rn
// are plain operators in this HIP version, so after inlining they can be contracted under
// the default -ffp-contract=fast-honor-pragmas; the pragma here takes the contract flag off
// the operation itself.)
__device__ __forceinline__ float mul_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}

__device__ __forceinline__ float add_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}

__device__ __forceinline__ float sub_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a - b;
}

// One uint8 SIFT descriptor per wave, lane l holding components 2l and 2l+1, dequantised
// as SIFT_Feature::normalizeDescriptor (V3D/Base/v3d_feature.h:30-38): the sum of squares
// of integers <= 255 over 128 terms stays below 2^24, so it is exact in FP32 in any order,
// and the scale is 1/sqrt with IEEE rounding.
__device__ __forceinline__ float dequant_u8_denom(uchar2 v)
{
    float const a = (float)v.x, b = (float)v.y;
    float s = a * a + b * b;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    return __fdiv_rn(1.0f, __fsqrt_rn(s));
}

// element value: the IEEE product (float)q * denom, the same in every kernel that forms it
__device__ __forceinline__ float dequant_u8_value(unsigned int q, float denom) { return mul_nc((float)q, denom); }

__device__ __forceinline__ float2 dequant_u8_wave(uchar2 v)
{
    float const denom = dequant_u8_denom(v);
    return make_float2(dequant_u8_value(v.x, denom), dequant_u8_value(v.y, denom));
}

// Wait until this wave's direct global->LDS loads (global_load_lds_dwordx4) have written
// LDS: s_waitcnt vmcnt(0).  A double-buffered stage is published to the other waves by a
// barrier; the compiler does not insert this wait on every path into that barrier (the
// loop back edge of the stage pipelines lacked it), so the stage loops call it before each
// such barrier.  The loads were issued a whole tile earlier, so the wait rarely stalls.
__device__ __forceinline__ void wait_lds_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Work-groups of a launch are dispatched to the eight XCDs round-robin (work-group b runs on
// XCD b % 8).  xcd_block(b, G) renumbers the G work-groups so that each XCD takes one
// contiguous range of logical blocks: neighbouring blocks, which read the same rows,
// share that XCD's L2.  A bijection of [0, G).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int G)
{
    int const per = G / kXcds, rem = G % kXcds;
    int const x = b % kXcds, i = b / kXcds;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
}

typedef float v16f __attribute__((ext_vector_type(16)));

// Shortlist keys (BF16 and FP8 passes): the FP32 score with its 6 low mantissa bits
// replaced by the position (block * 16 + r) of the score inside a 64-candidate tile.
__device__ __forceinline__ float pack_key(float v, int pos)
{
    return __uint_as_float((__float_as_uint(v) & ~63u) | (unsigned int)pos);
}

// max of three keys in one v_max3_f32 (fmaxf would first canonicalise each key, which the
// compiler cannot prove canonical after the bit packing).  The keys come from VALU
// instructions, never straight from an MFMA result.
__device__ __forceinline__ float key_max3(float a, float b, float c)
{
    float r;
    __asm__("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Group g of a tile against one query block: elements r with r / 8 == g of the first
// candidate block and r / 8 == 1 - g of the second (16 keys), folded into the running best
// o by a tree of eight max3 (the caller tracks the tile the best comes from).
template <bool kFull>
__device__ __forceinline__ float group_best(v16f const& a0, v16f const& a1, int g, float o, int jt, int nC, int h)
{
    float x[16];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int rr = 0; rr < 8; ++rr) {
            int const r = 8 * (g ^ cb) + rr;
            float v = cb == 0 ? a0[r] : a1[r];
            if (!kFull) {
                int const jl = cb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                v = (jt + jl < nC) ? v : -1e30f;
            }
            x[8 * cb + rr] = pack_key(v, cb * 16 + r);
        }
    float const b0 = key_max3(key_max3(x[0], x[1], x[2]), key_max3(x[3], x[4], x[5]), key_max3(x[6], x[7], x[8]));
    float const b1 = key_max3(key_max3(x[9], x[10], x[11]), key_max3(x[12], x[13], x[14]), x[15]);
    return key_max3(b0, b1, o);
}

typedef float v4f __attribute__((ext_vector_type(4)));

}  // namespace v3dm
