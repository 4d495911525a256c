// Learned shortlist network: 128 -> 256 -> 32 ReLU MLP, BF16 MFMA forward and backward
// kernels with FP32 accumulation, contrastive loss, Adam on FP32 master weights.
//
// The reference matcher has no learned stage; this network only proposes candidates, whose
// scores are then computed exactly.  Row-major parameter vector (FP32 masters):
//   W1 [256][128] at 0, b1 [256] at 32768, W2 [32][256] at 33024, b2 [32] at 41216.
//
// Activation layouts of a minibatch of R rows (padded to Rp, a multiple of 64):
//   X  bf16 [Rp][128]    XT  bf16 [128][Rp]     HT  bf16 [256][Rp] (after ReLU)
//   Z  f32  [Rp][32]     dC  f32  [Rp][32]      dZ  bf16 [Rp][32], dZT bf16 [32][Rp]
//   dZf f32 [Rp][32]     dHT bf16 [256][Rp]
// The weight-gradient GEMMs reduce over the rows, so they read the transposed activations
// (contiguous along the reduction); their per-chunk partials are summed by the Adam kernel
// in a fixed order (no atomics: the gradients are deterministic).
//
// MFMA v_mfma_f32_32x32x16_bf16: lane l supplies row (A) / column (B) l%32 and k = 8*(l/32)
// + 0..7 of the 16-deep step; D element r of lane l is (row 8*(r/4) + 4*(l/32) + r%4,
// column l%32).
#include "v3dm_device.h"
#include "v3dm_learned.h"

namespace v3dm {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v16f mfma_bf16(v8bf a, v8bf b, v16f c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v8bf load8(__bf16 const* p) { return *reinterpret_cast<v8bf const*>(p); }

// 4 consecutive rows (r%4 = 0..3) of one column of a D block, as 4 bf16 (8 bytes)
__device__ __forceinline__ void store4_bf16(__bf16* p, float a, float b, float c, float d)
{
    typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
    v4bf v = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    *reinterpret_cast<v4bf*>(p) = v;
}

// ------------------------------------------------------------------------------ gather
// One wave per row: row r of the minibatch is pool row tri[3*(r % B) + r / B] (anchors,
// then positives, then negatives); rows >= R are zero.
__global__ __launch_bounds__(256) void k_mlp_gather(float const* __restrict__ pool, int const* __restrict__ tri,
                                                    int const* __restrict__ order, int orderOff, int B, int Rp,
                                                    __bf16* __restrict__ X, __bf16* __restrict__ XT)
{
    int const lane = threadIdx.x & 63;
    int const r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= Rp) return;
    float2 v = make_float2(0.0f, 0.0f);
    if (r < 3 * B) {
        int const t = order[orderOff + r % B];
        int const row = tri[3 * t + r / B];
        v = reinterpret_cast<float2 const*>(pool + (size_t)row * kDim)[lane];
    }
    __bf16 const a = (__bf16)v.x, b = (__bf16)v.y;
    X[(size_t)r * kDim + 2 * lane] = a;
    X[(size_t)r * kDim + 2 * lane + 1] = b;
    XT[(size_t)(2 * lane) * Rp + r] = a;
    XT[(size_t)(2 * lane + 1) * Rp + r] = b;
}

// FP32 rows -> bf16 rows (inference input), rows >= n zero.
__global__ __launch_bounds__(256) void k_mlp_rows_bf16(float const* __restrict__ src, int n, int Rp,
                                                       __bf16* __restrict__ X)
{
    int const lane = threadIdx.x & 63;
    int const r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= Rp) return;
    float2 v = make_float2(0.0f, 0.0f);
    if (r < n) v = reinterpret_cast<float2 const*>(src + (size_t)r * kDim)[lane];
    X[(size_t)r * kDim + 2 * lane] = (__bf16)v.x;
    X[(size_t)r * kDim + 2 * lane + 1] = (__bf16)v.y;
}

// ------------------------------------------------------------------------------ forward
// 64 rows per work-group.  Layer 1: wave w computes hidden units [64w, 64w + 64) of both
// 32-row blocks (4 D blocks, 8 k-steps); bias + ReLU; h -> LDS (bf16, [row][unit]) and,
// when HT is given, -> HT.  Layer 2: waves 0 and 1 compute the 32 codes of row block w
// (16 k-steps over the units, A from LDS).
constexpr int kHsStride = 264;  // bf16 per LDS row of h (256 + 8: rows start 16 bytes apart mod 128)

// A-operand rows: bf16 (training minibatch) or FP32 (descriptors, converted in registers)
__device__ __forceinline__ v8bf load_row8(__bf16 const* p) { return load8(p); }
__device__ __forceinline__ v8bf load_row8(float const* p)
{
    float4 const u = reinterpret_cast<float4 const*>(p)[0];
    float4 const v = reinterpret_cast<float4 const*>(p)[1];
    return v8bf{(__bf16)u.x, (__bf16)u.y, (__bf16)u.z, (__bf16)u.w,
                (__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
}

template <typename TIn>
__global__ __launch_bounds__(256) void k_mlp_forward(TIn const* __restrict__ X, int Rp,
                                                     __bf16 const* __restrict__ W1b, float const* __restrict__ b1,
                                                     __bf16 const* __restrict__ W2b, float const* __restrict__ b2,
                                                     __bf16* __restrict__ HT, float* __restrict__ Z)
{
    __shared__ __attribute__((aligned(16))) __bf16 Hs[64 * kHsStride];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const row0 = blockIdx.x * 64;

    v16f acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[rb][u] = v16f{};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        int const k = ks * 16 + 8 * h;
        v8bf a[2], b[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) a[rb] = load_row8(X + (size_t)(row0 + rb * 32 + c32) * kDim + k);
#pragma unroll
        for (int u = 0; u < 2; ++u) b[u] = load8(W1b + (size_t)((2 * wave + u) * 32 + c32) * kDim + k);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[rb][u] = mfma_bf16(a[rb], b[u], acc[rb][u]);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        int const unit = (2 * wave + u) * 32 + c32;
        float const bias = b1[unit];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = fmaxf(acc[rb][u][4 * g + e] + bias, 0.0f);
                int const rl = rb * 32 + 8 * g + 4 * h;  // first of 4 consecutive rows
#pragma unroll
                for (int e = 0; e < 4; ++e) Hs[(rl + e) * kHsStride + unit] = (__bf16)v[e];
                if (HT) store4_bf16(HT + (size_t)unit * Rp + row0 + rl, v[0], v[1], v[2], v[3]);
            }
        }
    }
    __syncthreads();
    if (wave >= 2) return;
    int const rb = wave;
    v16f z = v16f{};
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) {
        int const k = ks * 16 + 8 * h;
        v8bf const a = *reinterpret_cast<v8bf const*>(&Hs[(rb * 32 + c32) * kHsStride + k]);
        v8bf const b = load8(W2b + (size_t)c32 * kHid + k);
        z = mfma_bf16(a, b, z);
    }
    float const bias = b2[c32];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        int const row = row0 + rb * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
        Z[(size_t)row * kCode + c32] = z[r] + bias;
    }
}

// ------------------------------------------------------------------------------ loss
// 64 anchors per work-group, 4 threads per anchor (quarter q of each 64-row positive tile).
// Codes are normalised on the fly.  The in-batch hardest positive of anchor t is the
// argmax over t' != t of c_t . c_{p_t'} (first index on ties).  Gradients with respect to
// the normalised codes go to dC: anchor and negative rows are owned by their triplet,
// positive rows also receive the hardest-negative terms of other anchors (atomic adds on a
// zeroed dC).
__device__ __forceinline__ float softplus_f(float x) { return x > 20.0f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoid_f(float x) { return 1.0f / (1.0f + expf(-x)); }

__global__ __launch_bounds__(256) void k_mlp_loss(float const* __restrict__ Z, int B, float invTau,
                                                  float* __restrict__ dC, int* __restrict__ hardOut,
                                                  float* __restrict__ lossPart)
{
    __shared__ float tile[64][kCode + 1];
    __shared__ float wl[4];
    int const tid = threadIdx.x;
    int const q = tid & 3;
    int const al = tid >> 2;
    int const t = blockIdx.x * 64 + al;
    bool const valid = t < B;

    // normalised anchor code (8 components per thread, gathered to all 32 by shuffles)
    float ca[kCode];
    {
        int const tt = valid ? t : 0;
        float s = 0.0f;
        float part[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            part[e] = Z[(size_t)tt * kCode + 8 * q + e];
            s += part[e] * part[e];
        }
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        float const inv = 1.0f / fmaxf(sqrtf(s), 1e-12f);
#pragma unroll
        for (int qq = 0; qq < 4; ++qq)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float const v = __shfl(part[e], ((tid & 63) & ~3) | qq) * inv;
                ca[8 * qq + e] = v;
            }
    }
    float best = -INFINITY;
    int bestIdx = -1;
    for (int p0 = 0; p0 < B; p0 += 64) {
        __syncthreads();
        {
            // stage positives p0 .. p0+63, normalised (4 threads per row, 8 components each)
            int const pr = tid >> 2;
            int const pi = p0 + pr;
            float part[8];
            float s = 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                part[e] = (pi < B) ? Z[(size_t)(B + pi) * kCode + 8 * q + e] : 0.0f;
                s += part[e] * part[e];
            }
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            float const inv = 1.0f / fmaxf(sqrtf(s), 1e-12f);
#pragma unroll
            for (int e = 0; e < 8; ++e) tile[pr][8 * q + e] = part[e] * inv;
        }
        __syncthreads();
        for (int j = q; j < 64; j += 4) {
            int const pj = p0 + j;
            if (pj >= B || pj == t) continue;
            float d = 0.0f;
#pragma unroll
            for (int e = 0; e < kCode; ++e) d += ca[e] * tile[j][e];
            if (d > best) {
                best = d;
                bestIdx = pj;
            }
        }
    }
    // combine the 4 quarters (first index on ties)
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
        float const ob = __shfl_xor(best, o);
        int const oi = __shfl_xor(bestIdx, o);
        bool const take = ob > best || (ob == best && oi >= 0 && (bestIdx < 0 || oi < bestIdx));
        best = take ? ob : best;
        bestIdx = take ? oi : bestIdx;
    }
    float lossT = 0.0f;
    if (valid) {
        // positive and negative codes of triplet t
        float cp[kCode], cn[kCode];
        float sp2 = 0.0f, sn2 = 0.0f;
#pragma unroll
        for (int e = 0; e < kCode; ++e) {
            cp[e] = Z[(size_t)(B + t) * kCode + e];
            cn[e] = Z[(size_t)(2 * B + t) * kCode + e];
            sp2 += cp[e] * cp[e];
            sn2 += cn[e] * cn[e];
        }
        float const ip = 1.0f / fmaxf(sqrtf(sp2), 1e-12f), in = 1.0f / fmaxf(sqrtf(sn2), 1e-12f);
        float sp = 0.0f, sn = 0.0f;
#pragma unroll
        for (int e = 0; e < kCode; ++e) {
            cp[e] *= ip;
            cn[e] *= in;
            sp += ca[e] * cp[e];
            sn += ca[e] * cn[e];
        }
        float const xh = (best - sp) * invTau, xn = (sn - sp) * invTau;
        lossT = softplus_f(xn) + (bestIdx >= 0 ? softplus_f(xh) : 0.0f);
        float const scale = invTau / (float)B;
        float const gn = sigmoid_f(xn) * scale;
        float const gh = bestIdx >= 0 ? sigmoid_f(xh) * scale : 0.0f;
        // quarter q writes components 8q .. 8q+7
        float chq[8];
        if (bestIdx >= 0) {
            float s2 = 0.0f;
            float full[kCode];
#pragma unroll
            for (int e = 0; e < kCode; ++e) {
                full[e] = Z[(size_t)(B + bestIdx) * kCode + e];
                s2 += full[e] * full[e];
            }
            float const ih = 1.0f / fmaxf(sqrtf(s2), 1e-12f);
#pragma unroll
            for (int e = 0; e < 8; ++e) chq[e] = full[8 * q + e] * ih;
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) chq[e] = 0.0f;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int const k = 8 * q + e;
            float const a = ca[k];
            dC[(size_t)t * kCode + k] = gn * (cn[k] - cp[k]) + gh * (chq[e] - cp[k]);
            atomicAdd(&dC[(size_t)(B + t) * kCode + k], -(gn + gh) * a);
            dC[(size_t)(2 * B + t) * kCode + k] = gn * a;
            if (bestIdx >= 0) atomicAdd(&dC[(size_t)(B + bestIdx) * kCode + k], gh * a);
        }
        if (q == 0 && hardOut) hardOut[t] = bestIdx;
    }
    // work-group loss sum (quarter 0 of each anchor contributes)
    float v = (q == 0) ? lossT : 0.0f;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((tid & 63) == 0) wl[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) lossPart[blockIdx.x] = wl[0] + wl[1] + wl[2] + wl[3];
}

// Code gradient through the normalisation: dz = (dc - c (c . dc)) / |z|.  One thread per
// row; rows >= R get zeros.
__global__ __launch_bounds__(256) void k_mlp_norm_backward(float const* __restrict__ Z, float const* __restrict__ dC,
                                                           int R, int Rp, float* __restrict__ dZf,
                                                           __bf16* __restrict__ dZ, __bf16* __restrict__ dZT)
{
    int const r = blockIdx.x * 256 + threadIdx.x;
    if (r >= Rp) return;
    float z[kCode], d[kCode];
    float s2 = 0.0f, cd = 0.0f;
#pragma unroll
    for (int e = 0; e < kCode; ++e) {
        z[e] = r < R ? Z[(size_t)r * kCode + e] : 0.0f;
        d[e] = r < R ? dC[(size_t)r * kCode + e] : 0.0f;
        s2 += z[e] * z[e];
    }
    float const n = fmaxf(sqrtf(s2), 1e-12f);
    float const inv = 1.0f / n;
#pragma unroll
    for (int e = 0; e < kCode; ++e) cd += z[e] * inv * d[e];
#pragma unroll
    for (int e = 0; e < kCode; ++e) {
        float const g = (d[e] - z[e] * inv * cd) * inv;
        dZf[(size_t)r * kCode + e] = g;
        dZ[(size_t)r * kCode + e] = (__bf16)g;
        dZT[(size_t)e * Rp + r] = (__bf16)g;
    }
}

// ------------------------------------------------------------------------------ backward
// dH = dZ W2 (64 rows per work-group; wave w: units [64w, 64w+64) of both row blocks, two
// 16-deep steps over the 32 codes, B operand from W2T [unit][code]), masked by h > 0,
// written transposed (dHT).
__global__ __launch_bounds__(256) void k_mlp_backward_h(__bf16 const* __restrict__ dZ, int Rp,
                                                        __bf16 const* __restrict__ W2Tb,
                                                        __bf16 const* __restrict__ HT, __bf16* __restrict__ dHT)
{
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const row0 = blockIdx.x * 64;
    v16f acc[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[rb][u] = v16f{};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        int const k = ks * 16 + 8 * h;
        v8bf a[2], b[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) a[rb] = load8(dZ + (size_t)(row0 + rb * 32 + c32) * kCode + k);
#pragma unroll
        for (int u = 0; u < 2; ++u) b[u] = load8(W2Tb + (size_t)((2 * wave + u) * 32 + c32) * kCode + k);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[rb][u] = mfma_bf16(a[rb], b[u], acc[rb][u]);
    }
    typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        int const unit = (2 * wave + u) * 32 + c32;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                int const rl = row0 + rb * 32 + 8 * g + 4 * h;
                v4bf const hv = *reinterpret_cast<v4bf const*>(HT + (size_t)unit * Rp + rl);
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = ((float)hv[e] > 0.0f) ? acc[rb][u][4 * g + e] : 0.0f;
                store4_bf16(dHT + (size_t)unit * Rp + rl, v[0], v[1], v[2], v[3]);
            }
    }
}

// Weight gradients over a chunk of kWgChunk rows: grid (chunks, 5); y < 4: dW1 blocks (unit
// block 2y + w/2 ... ) -- each wave computes 2 of the 32 dW1 blocks of its y, y == 4: the 8
// dW2 blocks (2 per wave).  Bias gradients: y == 4 also sums dZf (db2) and dHT (db1) over
// the chunk.  Partials -> part[chunk][param].
constexpr int kWgChunk = 512;

__global__ __launch_bounds__(256) void k_mlp_wgrad(__bf16 const* __restrict__ XT, __bf16 const* __restrict__ HT,
                                                   __bf16 const* __restrict__ dZT, __bf16 const* __restrict__ dHT,
                                                   float const* __restrict__ dZf, int Rp, float* __restrict__ part)
{
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const chunk = blockIdx.x;
    int const y = blockIdx.y;
    int const i0 = chunk * kWgChunk;
    float* out = part + (size_t)chunk * kParams;
    v16f acc[2] = {v16f{}, v16f{}};
    if (y < 4) {
        // dW1 blocks b = 8y + 2*wave + s, s = 0,1: unit block ub = b / 4, input block db = b % 4
        int ub[2], db[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int const b = 8 * y + 2 * wave + s;
            ub[s] = b >> 2;
            db[s] = b & 3;
        }
        for (int ks = 0; ks < kWgChunk / 16; ++ks) {
            int const i = i0 + ks * 16 + 8 * h;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                v8bf const a = load8(dHT + (size_t)(ub[s] * 32 + c32) * Rp + i);
                v8bf const b = load8(XT + (size_t)(db[s] * 32 + c32) * Rp + i);
                acc[s] = mfma_bf16(a, b, acc[s]);
            }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int const u = ub[s] * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                int const d = db[s] * 32 + c32;
                out[kOffW1 + u * kDim + d] = acc[s][r];
            }
    } else {
        // dW2 blocks: unit block ub = 2*wave + s; D[code][unit]
        for (int ks = 0; ks < kWgChunk / 16; ++ks) {
            int const i = i0 + ks * 16 + 8 * h;
            v8bf const a = load8(dZT + (size_t)c32 * Rp + i);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                v8bf const b = load8(HT + (size_t)((2 * wave + s) * 32 + c32) * Rp + i);
                acc[s] = mfma_bf16(a, b, acc[s]);
            }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                int const c = 8 * (r >> 2) + 4 * h + (r & 3);
                int const u = (2 * wave + s) * 32 + c32;
                out[kOffW2 + c * kHid + u] = acc[s][r];
            }
        // bias gradients of the chunk: thread t < 256 -> db1[t], t < 32 also db2[t]
        int const t = threadIdx.x;
        float s1 = 0.0f;
        for (int i = 0; i < kWgChunk; ++i) s1 += (float)dHT[(size_t)t * Rp + i0 + i];
        out[kOffB1 + t] = s1;
        if (t < kCode) {
            float s2 = 0.0f;
            for (int i = 0; i < kWgChunk; ++i) s2 += dZf[(size_t)(i0 + i) * kCode + t];
            out[kOffB2 + t] = s2;
        }
    }
}

// Sum of the chunk partials (fixed order) -> grad; Adam on the FP32 masters; refresh the
// bf16 copies (W1b, W2b, W2Tb).  One thread per parameter.
__global__ __launch_bounds__(256) void k_mlp_adam(float const* __restrict__ part, int nChunks,
                                                  float* __restrict__ grad, float* __restrict__ P,
                                                  float* __restrict__ M, float* __restrict__ V, float lr,
                                                  float beta1, float beta2, float eps, float bc1, float bc2,
                                                  int apply, __bf16* __restrict__ W1b, __bf16* __restrict__ W2b,
                                                  __bf16* __restrict__ W2Tb, __bf16* __restrict__ W2p)
{
    int const p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kParams) return;
    float g = 0.0f;
    for (int c = 0; c < nChunks; ++c) g += part[(size_t)c * kParams + p];
    grad[p] = g;
    float w = P[p];
    if (apply) {
        float const m = beta1 * M[p] + (1.0f - beta1) * g;
        float const v = beta2 * V[p] + (1.0f - beta2) * g * g;
        M[p] = m;
        V[p] = v;
        w = w - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
        P[p] = w;
    }
    if (p < kOffB1) {
        W1b[p] = (__bf16)w;
    } else if (p >= kOffW2 && p < kOffB2) {
        int const q = p - kOffW2;
        int const c = q / kHid, u = q - c * kHid;
        W2b[q] = (__bf16)w;
        W2Tb[u * kCode + c] = (__bf16)w;
        // W2p: A operand (code c, lane half hh, element v) of step s of unit block ub, for
        // the unit that element r = 8s + v of the first layer's D block holds:
        // u % 32 = 8 (r / 4) + 4 hh + r % 4
        int const ub = u >> 5, ul = u & 31;
        int const r = 4 * (ul >> 3) + (ul & 3), hh = (ul >> 2) & 1;
        int const s = r >> 3, v = r & 7;
        W2p[((ub * 2 + s) * 64 + hh * 32 + c) * 8 + v] = (__bf16)w;
    }
}

// ------------------------------------------------------------------------------ inference codes
// Codes of descriptor rows.  A work-group of kMlpWaves waves stages the network once in LDS
// and its waves then take 32-row units independently (unit u of the grid-stride sequence):
//   W1 rows padded to 272 bytes (the 32 rows one A-operand read touches fall on distinct
//     bank groups); the pad holds the first layer's bias split into bf16 (bh, bh, bl);
//   W2 in the lane order of the second layer's A operands (W2p), 16 contiguous bytes a lane.
// Per unit:
//   layer 1, transposed: D1[unit][row] = W1 . x + s b1 (A = W1 rows, B = the wave's rows in
//     bf16, kept in registers; the bias as a ninth product whose B operand is (sh, sl, sh):
//     bh sh + bh sl + bl sh = s b1 to about 2^-16), one 32-unit block at a time;
//   ReLU on the bf16 values (a packed int16 max); the 16 hidden units a lane holds are the B
//     operand of layer 2 as they are (k order permuted), with W2 read in the same permutation;
//   layer 2: D2[code][row] + s b2; lane (row, half) holds 16 of the row's codes: norm over the
//     two halves, quantisation, one 16-byte store (bytes 16*half .. +15 of the row, in element
//     order: a fixed permutation of the code dimensions, the same for every row).
// s is 1 for FP32 rows.  uint8 rows enter as the integers q (exact in bf16) with s = 1/d: the
// network's output on q d is z = z'/s with z' = W2 relu(W1 q + s b1) + s b2, and the code is
// z normalised, so z' gives it.
// The LDS operands of a unit (88 reads of 16 bytes) pass through a ring of registers kRing
// reads ahead of their MFMA, and the next 32-unit block's first-layer products are issued
// before the current block's ReLU, so LDS latency and the VALU work overlap the MFMA chain.
constexpr int kW1Stride = 136;  // bf16 per LDS row of W1 (128 + 8)
constexpr int kMlpWaves = 8;    // waves per work-group (LDS: 85 KB, one group per CU)
constexpr int kMlpRing = 8;     // LDS operand reads in flight per wave

// Rows of one feature set: FP32 rows, or uint8 rows with their dequantisation scale
// (element k = (float)q_k * denom, the FP32 value the batch image holds).
struct MlpRows {
    float const* f;
    uint8_t const* q;
    float const* denom;
};

// The inputs of one 32-row unit for lane (c32, h): elements 16 ks + 8 h + 0..7, ks < 8.
template <bool kU8>
struct MlpRaw;
template <>
struct MlpRaw<false> {
    float4 v[16];
};
template <>
struct MlpRaw<true> {
    uint2 v[8];
    float d;
};

__device__ __forceinline__ void mlp_load(MlpRows const& in, size_t row, int h, MlpRaw<false>& r)
{
    float4 const* xp = reinterpret_cast<float4 const*>(in.f + row * kDim);
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        r.v[2 * ks] = xp[4 * ks + 2 * h];
        r.v[2 * ks + 1] = xp[4 * ks + 2 * h + 1];
    }
}

__device__ __forceinline__ void mlp_load(MlpRows const& in, size_t row, int h, MlpRaw<true>& r)
{
    uint8_t const* qp = in.q + row * kDim + 8 * h;
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) r.v[ks] = *reinterpret_cast<uint2 const*>(qp + 16 * ks);
    r.d = in.denom[row];
}

__device__ __forceinline__ void mlp_convert(MlpRaw<false> const& r, v8bf (&xb)[8])
{
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        float4 const u = r.v[2 * ks], v = r.v[2 * ks + 1];
        xb[ks] = v8bf{(__bf16)u.x, (__bf16)u.y, (__bf16)u.z, (__bf16)u.w,
                      (__bf16)v.x, (__bf16)v.y, (__bf16)v.z, (__bf16)v.w};
    }
}

// uint8 rows: the integers themselves (at most 8 significant bits: exact in bf16)
__device__ __forceinline__ void mlp_convert(MlpRaw<true> const& r, v8bf (&xb)[8])
{
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        unsigned int const a = r.v[ks].x, b = r.v[ks].y;
        xb[ks] = v8bf{(__bf16)(float)(a & 0xffu), (__bf16)(float)((a >> 8) & 0xffu),
                      (__bf16)(float)((a >> 16) & 0xffu), (__bf16)(float)(a >> 24),
                      (__bf16)(float)(b & 0xffu), (__bf16)(float)((b >> 8) & 0xffu),
                      (__bf16)(float)((b >> 16) & 0xffu), (__bf16)(float)(b >> 24)};
    }
}

// s of the header: 1 for FP32 rows, 1/d for uint8 rows (0 for an all-zero row, d = inf)
__device__ __forceinline__ float mlp_bias_scale(MlpRaw<false> const&) { return 1.0f; }
__device__ __forceinline__ float mlp_bias_scale(MlpRaw<true> const& r) { return 1.0f / r.d; }

typedef short v2s16 __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf16 __attribute__((ext_vector_type(2)));
typedef float v2f32 __attribute__((ext_vector_type(2)));

// max(x, 0) of two values rounded to bf16 (one v_cvt_pk_bf16_f32): a bf16 is negative
// exactly when its bits are negative as int16, so one packed int16 max clears both lanes
__device__ __forceinline__ unsigned relu_bf16x2(float x0, float x1)
{
    v2bf16 const p = __builtin_convertvector((v2f32){x0, x1}, v2bf16);
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(v2s16, p), (v2s16){0, 0}));
}

// LDS operand t of a unit, t < 88, in the order the unit uses them: for block ub, its nine
// layer-1 operands (W1 ks = 0..7, the bias column) are read at 9 ub + j (ub = 0..7), its two
// layer-2 operands at 72 + 2 ub + s
__device__ __forceinline__ __bf16 const* mlp_operand(__bf16 const* W1s, __bf16 const* W2s, int t, int c32, int h,
                                                     int lane)
{
    if (t < 72) {
        int const ub = t / 9, j = t % 9;
        return W1s + (ub * 32 + c32) * kW1Stride + (j < 8 ? 16 * j + 8 * h : 128);
    }
    return W2s + ((t - 72) * 64 + lane) * 8;
}

// The network -> LDS (one work-group of kMlpWaves waves): W1 rows padded to kW1Stride with
// the bias split (bh, bh, bl) in the pad, W2 in its operand order.  All loads of a thread are
// issued before its first store.
__device__ __forceinline__ void mlp_stage_network(__bf16* W1s, __bf16* W2s, __bf16 const* __restrict__ W1b,
                                                  float const* __restrict__ b1, __bf16 const* __restrict__ W2p,
                                                  int tid)
{
    constexpr int kW1Chunks = kHid * 16 / (64 * kMlpWaves), kW2Chunks = kHid * kCode / 8 / (64 * kMlpWaves);
    uint4 w[kW1Chunks], w2[kW2Chunks];
#pragma unroll
    for (int i = 0; i < kW1Chunks; ++i) {
        int const f = tid + 64 * kMlpWaves * i;
        w[i] = reinterpret_cast<uint4 const*>(W1b + (size_t)(f >> 4) * 128)[f & 15];
    }
#pragma unroll
    for (int i = 0; i < kW2Chunks; ++i) w2[i] = reinterpret_cast<uint4 const*>(W2p)[tid + 64 * kMlpWaves * i];
    float const bv = tid < kHid ? b1[tid] : 0.0f;
#pragma unroll
    for (int i = 0; i < kW1Chunks; ++i) {
        int const f = tid + 64 * kMlpWaves * i;
        *reinterpret_cast<uint4*>(&W1s[(f >> 4) * kW1Stride + (f & 15) * 8]) = w[i];
    }
#pragma unroll
    for (int i = 0; i < kW2Chunks; ++i) reinterpret_cast<uint4*>(W2s)[tid + 64 * kMlpWaves * i] = w2[i];
    if (tid < kHid) {
        __bf16 const bh = (__bf16)bv;
        __bf16 const bl = (__bf16)(bv - (float)bh);
        __bf16 const zz = (__bf16)0.0f;
        *reinterpret_cast<v8bf*>(&W1s[tid * kW1Stride + 128]) = v8bf{bh, bh, bl, zz, zz, zz, zz, zz};
    }
}

// Normalise and quantise the 16 codes a lane holds (z: the unit's output plus bias) and store
// them: bytes 16 h .. 16 h + 15 of the row's 32.
__device__ __forceinline__ void mlp_store_codes(v16f const& z, float scale, int h, signed char* __restrict__ dst)
{
    float s2 = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) s2 += z[r] * z[r];
    s2 += __shfl_xor(s2, 32);
    float const inv = scale / fmaxf(sqrtf(s2), 1e-30f);
    unsigned int w[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        unsigned int p = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float const c = z[4 * g + e] * inv;
            int const q = (c == c) ? (int)rintf(fminf(fmaxf(c, -31.0f), 31.0f)) : 0;
            p |= ((unsigned int)(4 * q) & 0xffu) << (8 * e);
        }
        w[g] = p;
    }
    *reinterpret_cast<uint4*>(dst + 16 * h) = make_uint4(w[0], w[1], w[2], w[3]);
}

// kPrefetch: the next unit's inputs are loaded while the current unit is computed (the
// uint8 rows take 17 registers, few enough to hold a second set).
template <bool kU8, bool kPrefetch>
__device__ __forceinline__ void mlp_codes_body(MlpRows const& in0, int nUnits0, signed char* __restrict__ codes0,
                                               MlpRows const& in1, int nUnits1, signed char* __restrict__ codes1,
                                               __bf16 const* __restrict__ W1b, float const* __restrict__ b1,
                                               __bf16 const* __restrict__ W2p, float const* __restrict__ b2,
                                               float scale)
{
    __shared__ __attribute__((aligned(16))) __bf16 W1s[kHid * kW1Stride];
    __shared__ __attribute__((aligned(16))) __bf16 W2s[kHid * kCode];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const nUnits = nUnits0 + nUnits1;
    int const stride = (int)gridDim.x * kMlpWaves;
    // unit -> row of this lane in its feature set
    auto rowOf = [&](int u) -> size_t { return (size_t)((u >= nUnits0 ? u - nUnits0 : u) * 32 + c32); };
    int unit = (int)blockIdx.x * kMlpWaves + wave;
    MlpRaw<kU8> raw;
    if (kPrefetch && unit < nUnits) mlp_load(unit >= nUnits0 ? in1 : in0, rowOf(unit), h, raw);
    mlp_stage_network(W1s, W2s, W1b, b1, W2p, tid);
    __syncthreads();
    float bz[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) bz[r] = b2[8 * (r >> 2) + 4 * h + (r & 3)];

    for (; unit < nUnits; unit += stride) {
        bool const second = unit >= nUnits0;
        signed char* __restrict__ codes = second ? codes1 : codes0;
        size_t const row = rowOf(unit);
        if (!kPrefetch) mlp_load(second ? in1 : in0, row, h, raw);
        v8bf xb[8];
        mlp_convert(raw, xb);
        float const sc = mlp_bias_scale(raw);
        __bf16 const sh = (__bf16)sc;
        __bf16 const sl = (__bf16)(sc - (float)sh);
        __bf16 const zz = (__bf16)0.0f;
        v8bf const xbias = h == 0 ? v8bf{sh, sl, sh, zz, zz, zz, zz, zz} : v8bf{zz, zz, zz, zz, zz, zz, zz, zz};
        if (kPrefetch) {
            int const nxt = unit + stride;
            if (nxt < nUnits) mlp_load(nxt >= nUnits0 ? in1 : in0, rowOf(nxt), h, raw);
        }
        // LDS operand ring: operand i (in order of use) sits in ring[i % kMlpRing] from its
        // read, issued kMlpRing operands ahead, until its MFMA.  Order of use: L1(0), then
        // per ub: L1(ub + 1) (ub < 7), L2(ub).  Each step (one MFMA, the read it frees room
        // for, a share of the ReLU work) is closed by a scheduling barrier, so the compiler
        // keeps this interleaving instead of hoisting every read to the top.
        v8bf ring[kMlpRing];
        auto order = [](int i) -> int {  // i-th operand used -> its LDS operand index
            if (i < 9) return i;
            int const k = i - 9, ub = k / 11, j = k % 11;
            if (ub < 7) return j < 9 ? 9 * (ub + 1) + j : 72 + 2 * ub + (j - 9);
            return 72 + 14 + (k - 77);
        };
#pragma unroll
        for (int i = 0; i < kMlpRing; ++i) ring[i] = load8(mlp_operand(W1s, W2s, order(i), c32, h, lane));
        int used = 0;
        auto next = [&]() -> v8bf {
            v8bf const op = ring[used % kMlpRing];
            if (used + kMlpRing < 88)
                ring[used % kMlpRing] = load8(mlp_operand(W1s, W2s, order(used + kMlpRing), c32, h, lane));
            ++used;
            return op;
        };
        __builtin_amdgcn_sched_barrier(0);
        v16f z = v16f{};
        v16f a = v16f{};
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            v8bf const op = next();
            a = mfma_bf16(op, j < 8 ? xb[j] : xbias, a);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int ub = 0; ub < 8; ++ub) {
            unsigned w[8];
            if (ub < 7) {
                v16f an = v16f{};
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    v8bf const op = next();
                    an = mfma_bf16(op, j < 8 ? xb[j] : xbias, an);
                    if (j < 8) w[j] = relu_bf16x2(a[2 * j], a[2 * j + 1]);
                    __builtin_amdgcn_sched_barrier(0);
                }
                v8bf const hb0 = __builtin_bit_cast(v8bf, (uint4){w[0], w[1], w[2], w[3]});
                v8bf const hb1 = __builtin_bit_cast(v8bf, (uint4){w[4], w[5], w[6], w[7]});
                z = mfma_bf16(next(), hb0, z);
                __builtin_amdgcn_sched_barrier(0);
                z = mfma_bf16(next(), hb1, z);
                __builtin_amdgcn_sched_barrier(0);
                a = an;
            } else {
#pragma unroll
                for (int g = 0; g < 8; ++g) w[g] = relu_bf16x2(a[2 * g], a[2 * g + 1]);
                v8bf const hb0 = __builtin_bit_cast(v8bf, (uint4){w[0], w[1], w[2], w[3]});
                v8bf const hb1 = __builtin_bit_cast(v8bf, (uint4){w[4], w[5], w[6], w[7]});
                z = mfma_bf16(next(), hb0, z);
                z = mfma_bf16(next(), hb1, z);
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) z[r] = fmaf(bz[r], sc, z[r]);
        mlp_store_codes(z, scale, h, codes + row * kCode);
    }
}

__global__ __launch_bounds__(64 * kMlpWaves, 1) void k_mlp_codes(
    float const* __restrict__ X0, int nUnits0, signed char* __restrict__ codes0, float const* __restrict__ X1,
    int nUnits1, signed char* __restrict__ codes1, __bf16 const* __restrict__ W1b, float const* __restrict__ b1,
    __bf16 const* __restrict__ W2p, float const* __restrict__ b2, float scale)
{
    mlp_codes_body<false, false>(MlpRows{X0, nullptr, nullptr}, nUnits0, codes0, MlpRows{X1, nullptr, nullptr},
                                 nUnits1, codes1, W1b, b1, W2p, b2, scale);
}

// uint8 rows (the batch image), 64-row units: the wave's two 32-row groups share every LDS
// operand read (88 reads of 16 bytes per 64 rows, half the LDS traffic of 32-row units).  Per
// 32-unit block: the first layer's 18 MFMAs (each operand read once, used for both groups),
// the ReLU of group 0, then the second layer's MFMAs of group 0 interleaved with group 1's
// ReLU, then group 1's.  The next unit's rows are loaded during the current unit.
__device__ __forceinline__ void mlp_codes_u8_body(MlpRows const& in, int nUnits, signed char* __restrict__ codes,
                                                  __bf16 const* __restrict__ W1b, float const* __restrict__ b1,
                                                  __bf16 const* __restrict__ W2p, float const* __restrict__ b2,
                                                  float scale)
{
    __shared__ __attribute__((aligned(16))) __bf16 W1s[kHid * kW1Stride];
    __shared__ __attribute__((aligned(16))) __bf16 W2s[kHid * kCode];
    __shared__ __attribute__((aligned(16))) float b2s[kCode];
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const h = lane >> 5;
    int const c32 = lane & 31;
    int const stride = (int)gridDim.x * kMlpWaves;
    int unit = (int)blockIdx.x * kMlpWaves + wave;
    MlpRaw<true> raw[2];
    if (unit < nUnits) {
#pragma unroll
        for (int g = 0; g < 2; ++g) mlp_load(in, (size_t)unit * 64 + 32 * g + c32, h, raw[g]);
    }
    mlp_stage_network(W1s, W2s, W1b, b1, W2p, tid);
    if (tid < kCode) b2s[tid] = b2[tid];
    __syncthreads();

    for (; unit < nUnits; unit += stride) {
        v8bf xb[2][8];
        v8bf xbias[2];
        float sc[2];
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            mlp_convert(raw[g], xb[g]);
            sc[g] = mlp_bias_scale(raw[g]);
            __bf16 const sh = (__bf16)sc[g];
            __bf16 const sl = (__bf16)(sc[g] - (float)sh);
            __bf16 const zz = (__bf16)0.0f;
            xbias[g] = h == 0 ? v8bf{sh, sl, sh, zz, zz, zz, zz, zz} : v8bf{zz, zz, zz, zz, zz, zz, zz, zz};
        }
        int const nxt = unit + stride;
        if (nxt < nUnits) {
#pragma unroll
            for (int g = 0; g < 2; ++g) mlp_load(in, (size_t)nxt * 64 + 32 * g + c32, h, raw[g]);
        }
        // LDS operand ring in the order of use: per ub, the nine first-layer operands (W1
        // ks = 0..7, the bias column), then the two second-layer ones; each step (the MFMAs of
        // one operand, the read it frees room for, a share of the VALU work) is closed by a
        // scheduling barrier so that the compiler keeps the interleaving
        v8bf ring[kMlpRing];
        auto order = [](int i) -> int {  // i-th operand used -> its LDS operand index
            int const ub = i / 11, j = i % 11;
            return j < 9 ? 9 * ub + j : 72 + 2 * ub + (j - 9);
        };
#pragma unroll
        for (int i = 0; i < kMlpRing; ++i) ring[i] = load8(mlp_operand(W1s, W2s, order(i), c32, h, lane));
        int used = 0;
        auto next = [&]() -> v8bf {
            v8bf const op = ring[used % kMlpRing];
            if (used + kMlpRing < 88)
                ring[used % kMlpRing] = load8(mlp_operand(W1s, W2s, order(used + kMlpRing), c32, h, lane));
            ++used;
            return op;
        };
        __builtin_amdgcn_sched_barrier(0);
        v16f z[2] = {v16f{}, v16f{}};
#pragma unroll
        for (int ub = 0; ub < 8; ++ub) {
            v16f a[2] = {v16f{}, v16f{}};
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                v8bf const op = next();
                a[0] = mfma_bf16(op, j < 8 ? xb[0][j] : xbias[0], a[0]);
                a[1] = mfma_bf16(op, j < 8 ? xb[1][j] : xbias[1], a[1]);
                __builtin_amdgcn_sched_barrier(0);
            }
            unsigned w0[8], w1[8];
#pragma unroll
            for (int g = 0; g < 8; ++g) w0[g] = relu_bf16x2(a[0][2 * g], a[0][2 * g + 1]);
            v8bf const op0 = next();
            v8bf const op1 = next();
            z[0] = mfma_bf16(op0, __builtin_bit_cast(v8bf, (uint4){w0[0], w0[1], w0[2], w0[3]}), z[0]);
#pragma unroll
            for (int g = 0; g < 4; ++g) w1[g] = relu_bf16x2(a[1][2 * g], a[1][2 * g + 1]);
            __builtin_amdgcn_sched_barrier(0);
            z[0] = mfma_bf16(op1, __builtin_bit_cast(v8bf, (uint4){w0[4], w0[5], w0[6], w0[7]}), z[0]);
#pragma unroll
            for (int g = 4; g < 8; ++g) w1[g] = relu_bf16x2(a[1][2 * g], a[1][2 * g + 1]);
            __builtin_amdgcn_sched_barrier(0);
            z[1] = mfma_bf16(op0, __builtin_bit_cast(v8bf, (uint4){w1[0], w1[1], w1[2], w1[3]}), z[1]);
            z[1] = mfma_bf16(op1, __builtin_bit_cast(v8bf, (uint4){w1[4], w1[5], w1[6], w1[7]}), z[1]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int g = 0; g < 2; ++g) {
#pragma unroll
            for (int r = 0; r < 16; ++r) z[g][r] = fmaf(b2s[8 * (r >> 2) + 4 * h + (r & 3)], sc[g], z[g][r]);
            mlp_store_codes(z[g], scale, h, codes + ((size_t)unit * 64 + 32 * g + c32) * kCode);
        }
    }
}

__global__ __launch_bounds__(64 * kMlpWaves, 1) void k_mlp_codes_u8(
    uint8_t const* __restrict__ Q, float const* __restrict__ denom, int nUnits, signed char* __restrict__ codes,
    __bf16 const* __restrict__ W1b, float const* __restrict__ b1, __bf16 const* __restrict__ W2p,
    float const* __restrict__ b2, float scale)
{
    mlp_codes_u8_body(MlpRows{nullptr, Q, denom}, nUnits, codes, W1b, b1, W2p, b2, scale);
}

namespace {
int mlp_grid(int nUnits)
{
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int const groups = (nUnits + kMlpWaves - 1) / kMlpWaves;
    return groups < cus ? groups : cus;
}
}  // namespace

void launch_mlp_codes2_f32(float const* X0, int rows0, signed char* codes0, float const* X1, int rows1,
                           signed char* codes1, MlpWeights const& w, float scale, hipStream_t st)
{
    int const nu0 = rows0 / 32, nu1 = rows1 / 32;
    if (nu0 + nu1 <= 0) return;
    hipLaunchKernelGGL(k_mlp_codes, dim3(mlp_grid(nu0 + nu1)), dim3(64 * kMlpWaves), 0, st, X0, nu0, codes0, X1, nu1,
                       codes1, w.W1b, w.P + kOffB1, w.W2p, w.P + kOffB2, scale);
}

void launch_mlp_codes_f32(float const* X, int rows128, MlpWeights const& w, float scale, signed char* codes,
                          hipStream_t st)
{
    launch_mlp_codes2_f32(X, rows128, codes, nullptr, 0, nullptr, w, scale, st);
}

void launch_mlp_codes_u8(uint8_t const* Q, float const* denom, int rows128, MlpWeights const& w, float scale,
                         signed char* codes, hipStream_t st)
{
    int const nu = rows128 / 64;  // 64-row units
    if (nu <= 0) return;
    hipLaunchKernelGGL(k_mlp_codes_u8, dim3(mlp_grid(nu)), dim3(64 * kMlpWaves), 0, st, Q, denom, nu, codes, w.W1b,
                       w.P + kOffB1, w.W2p, w.P + kOffB2, scale);
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v)
{
    size_t const i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------ launchers
void launch_mlp_gather(float const* pool, int const* tri, int const* order, int orderOff, int B, int Rp,
                       __bf16* X, __bf16* XT, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_gather, dim3((Rp + 3) / 4), dim3(256), 0, st, pool, tri, order, orderOff, B, Rp, X, XT);
}

void launch_mlp_rows_bf16(float const* src, int n, int Rp, __bf16* X, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_rows_bf16, dim3((Rp + 3) / 4), dim3(256), 0, st, src, n, Rp, X);
}

void launch_mlp_forward(__bf16 const* X, int Rp, MlpWeights const& w, __bf16* HT, float* Z, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_forward<__bf16>, dim3(Rp / 64), dim3(256), 0, st, X, Rp, w.W1b, w.P + kOffB1, w.W2b,
                       w.P + kOffB2, HT, Z);
}



void launch_mlp_loss(float const* Z, int B, float tau, float* dC, int* hard, float* lossPart, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_loss, dim3((B + 63) / 64), dim3(256), 0, st, Z, B, 1.0f / tau, dC, hard, lossPart);
}

void launch_mlp_norm_backward(float const* Z, float const* dC, int R, int Rp, float* dZf, __bf16* dZ, __bf16* dZT,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_norm_backward, dim3((Rp + 255) / 256), dim3(256), 0, st, Z, dC, R, Rp, dZf, dZ, dZT);
}

void launch_mlp_backward_h(__bf16 const* dZ, int Rp, MlpWeights const& w, __bf16 const* HT, __bf16* dHT,
                           hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_backward_h, dim3(Rp / 64), dim3(256), 0, st, dZ, Rp, w.W2Tb, HT, dHT);
}

int mlp_wgrad_chunks(int Rp) { return (Rp + kWgChunk - 1) / kWgChunk; }

void launch_mlp_wgrad(__bf16 const* XT, __bf16 const* HT, __bf16 const* dZT, __bf16 const* dHT, float const* dZf,
                      int Rp, float* part, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(Rp / kWgChunk, 5), dim3(256), 0, st, XT, HT, dZT, dHT, dZf, Rp, part);
}

void launch_mlp_adam(float const* part, int nChunks, float* grad, MlpWeights const& w, float* M, float* V, float lr,
                     int step, int apply, hipStream_t st)
{
    float const beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float const bc1 = 1.0f - powf(beta1, (float)step), bc2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, st, part, nChunks, grad, w.P, M, V,
                       lr, beta1, beta2, eps, apply ? bc1 : 1.0f, apply ? bc2 : 1.0f, apply, w.W1b, w.W2b, w.W2Tb,
                       w.W2p);
}

void launch_fill(float* p, size_t n, float v, hipStream_t st)
{
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace v3dm
