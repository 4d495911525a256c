/*
 * CPU BASELINE of the putative matcher -- TEST / BENCHMARK INFRASTRUCTURE ONLY.
 *
 * The reference has no CPU matcher; this is a CPU port of its algorithm as it runs on
 * the GPU (V3D/CUDA/v3d_cudamatching.cu:522-552): an FP32 score GEMM (cublasSgemm,
 * cu:529-530), per-left and per-right best / second best with the ratio test
 * (cu:133-210) and the mutual check (cu:546-551).  The GEMM is cache-blocked and
 * vectorised by the compiler, and left blocks are spread over OpenMP threads.  It is
 * timed next to the GPU library and its match set is compared with the exact oracle
 * (FP32 summation order differs from the oracle's exact scores, so rows within a few
 * ulp of a decision boundary may differ).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define D 128
#define IB 64   /* left rows per block */
#define JB 256  /* right columns per block */

typedef struct { float b1, b2; int a1; } top2_t;

static inline void upd(top2_t* s, float v, int idx)
{
    s->a1 = (v > s->b1) ? idx : s->a1;
    s->b2 = (v > s->b1) ? s->b1 : fmaxf(v, s->b2);
    s->b1 = fmaxf(s->b1, v);
}

/* merge b (later indices) into a (earlier indices) */
static inline void merge(top2_t* a, top2_t const* b)
{
    float const lo1 = fminf(a->b1, b->b1);
    float const hi2 = fmaxf(a->b2, b->b2);
    if (b->b1 > a->b1) a->a1 = b->a1;
    a->b1 = fmaxf(a->b1, b->b1);
    a->b2 = fmaxf(lo1, hi2);
}

static inline int decide(top2_t s, float minScore, float minRatio)
{
    if (minRatio < 0.0f) return (s.b1 > minScore) ? s.a1 : -1;
    float const nextDist = (float)fmax(0.001, (double)(2.0f - 2.0f * s.b2));
    float const bestDist = 2.0f - 2.0f * s.b1;
    return (s.b1 > minScore && bestDist / nextDist < minRatio) ? s.a1 : -1;
}

int v3dcpu_find_putative_matches(float const* L, int nL, float const* R, int nR, float minScore,
                                 float minRatio, int32_t* pairs, int cap, int nThreads)
{
    if (nL <= 0 || nR <= 0) return 0;
    if (nThreads > 0) omp_set_num_threads(nThreads);
    int const nIB = (nL + IB - 1) / IB;
    int const nRp = (nR + JB - 1) / JB * JB;

    /* right descriptors transposed to k-major, padded with zeros */
    float* Rt = (float*)aligned_alloc(64, sizeof(float) * (size_t)D * nRp);
    top2_t* colS = (top2_t*)malloc(sizeof(top2_t) * (size_t)nL);
    top2_t* rowP = (top2_t*)malloc(sizeof(top2_t) * (size_t)nIB * nR);
    int* colMax = (int*)malloc(sizeof(int) * (size_t)nL);
    int* rowMax = (int*)malloc(sizeof(int) * (size_t)nR);

#pragma omp parallel for schedule(static)
    for (int j = 0; j < nRp; ++j)
        for (int k = 0; k < D; ++k) Rt[(size_t)k * nRp + j] = (j < nR) ? R[(size_t)j * D + k] : 0.0f;

#pragma omp parallel
    {
        float* Sb = (float*)aligned_alloc(64, sizeof(float) * IB * JB);
#pragma omp for schedule(dynamic, 1)
        for (int ib = 0; ib < nIB; ++ib) {
            int const i0 = ib * IB;
            int const ni = (nL - i0) < IB ? (nL - i0) : IB;
            top2_t cs[IB];
            for (int t = 0; t < IB; ++t) { cs[t].b1 = -1e30f; cs[t].b2 = -1e30f; cs[t].a1 = -1; }
            for (int j0 = 0; j0 < nR; j0 += JB) {
                int const nj = (nR - j0) < JB ? (nR - j0) : JB;
                memset(Sb, 0, sizeof(float) * IB * JB);
                for (int i = 0; i + 3 < ni + 3; i += 4) {
                    float* s0 = Sb + (size_t)(i + 0) * JB;
                    float* s1 = Sb + (size_t)(i + 1) * JB;
                    float* s2 = Sb + (size_t)(i + 2) * JB;
                    float* s3 = Sb + (size_t)(i + 3) * JB;
                    float const* l0 = L + (size_t)(i0 + (i + 0 < ni ? i + 0 : 0)) * D;
                    float const* l1 = L + (size_t)(i0 + (i + 1 < ni ? i + 1 : 0)) * D;
                    float const* l2 = L + (size_t)(i0 + (i + 2 < ni ? i + 2 : 0)) * D;
                    float const* l3 = L + (size_t)(i0 + (i + 3 < ni ? i + 3 : 0)) * D;
                    for (int k = 0; k < D; ++k) {
                        float const a0 = l0[k], a1 = l1[k], a2 = l2[k], a3 = l3[k];
                        float const* rt = Rt + (size_t)k * nRp + j0;
                        for (int j = 0; j < JB; ++j) {
                            float const r = rt[j];
                            s0[j] += a0 * r;
                            s1[j] += a1 * r;
                            s2[j] += a2 * r;
                            s3[j] += a3 * r;
                        }
                    }
                }
                for (int i = 0; i < ni; ++i)
                    for (int j = 0; j < nj; ++j) upd(&cs[i], Sb[(size_t)i * JB + j], j0 + j);
                for (int j = 0; j < nj; ++j) {
                    top2_t r = {-1e30f, -1e30f, -1};
                    for (int i = 0; i < ni; ++i) upd(&r, Sb[(size_t)i * JB + j], i0 + i);
                    rowP[(size_t)ib * nR + j0 + j] = r;
                }
            }
            for (int t = 0; t < ni; ++t) colS[i0 + t] = cs[t];
        }
        free(Sb);
    }

#pragma omp parallel for schedule(static)
    for (int i = 0; i < nL; ++i) colMax[i] = decide(colS[i], minScore, minRatio);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nR; ++j) {
        top2_t s = rowP[j];
        for (int ib = 1; ib < nIB; ++ib) merge(&s, &rowP[(size_t)ib * nR + j]);
        rowMax[j] = decide(s, minScore, minRatio);
    }
    int n = 0;
    for (int k = 0; k < nL; ++k) {
        int const l = colMax[k];
        if (l >= 0 && rowMax[l] == k) {
            if (n < cap) { pairs[2 * n] = k; pairs[2 * n + 1] = l; }
            ++n;
        }
    }
    free(Rt); free(colS); free(rowP); free(colMax); free(rowMax);
    return n;
}

int v3dcpu_max_threads(void) { return omp_get_max_threads(); }
