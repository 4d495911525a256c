/*
 * ORACLE for the putative matcher -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the CUDA SIFT putative matcher V3D_CUDA::SIFT_Matching
 * (V3D/CUDA/v3d_cudamatching.{h,cu}).  Only tests/, the smoke check and the CPU
 * baseline leg of the benchmark load it, and only to check the HIP library.
 * The library never calls it.
 *
 * Score definition: s_ij = fl32( sum_k (double)L_ik * (double)R_jk ), summed in FP64 in
 * ascending k.  Products of two floats are exact in FP64, so FMA contraction cannot
 * change the sum.  The reference computes S with cublasSgemm (cu:529-530) in an
 * unspecified FP32 order; that order is closed and not reproduced.
 *
 * Selection (per query, candidates scanned in ascending index):
 *   argMax  = (v > maxVal) ? idx : argMax
 *   nextVal = (v > maxVal) ? maxVal : fmaxf(v, nextVal)
 *   maxVal  = fmaxf(maxVal, v)                 init maxVal = nextVal = -1e30f, argMax = -1
 * follows _cuda_computeColumnMaximaWithRatioTest cu:133-162 and
 * _cuda_computeRowMaximaWithRatioTest cu:164-210 (ties keep the first index; NaN
 * scores are skipped by fmaxf).  Decision (cu:154-161):
 *   nextDist = (float)fmax(0.001, 2 - 2*nextVal);  bestDist = 2 - 2*maxVal
 *   accept  <=> maxVal > minScore && bestDist/nextDist < minRatio
 * with IEEE division (the reference is built with --use_fast_math, whose approximate
 * division is not reproduced).  minRatio < 0 selects the no-ratio rule
 * (_cuda_computeColumnMaxima cu:33-49, _cuda_computeRowMaxima cu:51-85).
 * Columns of S are left features, rows are right features (cu:31, cu:529-530):
 * colMax[i] is the choice of left i among right features, rowMax[j] the choice of right j.
 * Mutual check and output order: cu:546-551.
 *
 * Pins (tests/test_oracle.py): brute force with numpy on tiny inputs, hand-built score
 * matrices for ties / NaN / clamp / threshold cases, closed-form guided distances,
 * the symmetry S(L,R) = S(R,L)^T, and committed golden fixtures.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define D 128
#define QB 16 /* queries per vectorised block */

/* ------------------------------------------------------------------ selection rule */

/* Planted mistakes (tests/test_oracle.py builds these variants to show the checks fail
 * on them): V3DO_PLANT_TIE_LAST keeps the last index on ties, V3DO_PLANT_NO_CLAMP drops
 * the 0.001 floor of the ratio denominator, V3DO_PLANT_FP32_SUM sums in FP32. */
static inline void top2_update(float v, int idx, float* maxVal, float* nextVal, int* argMax)
{
    /* cu:145-147 */
#ifdef V3DO_PLANT_TIE_LAST
    *argMax = (v >= *maxVal) ? idx : *argMax;
#else
    *argMax = (v > *maxVal) ? idx : *argMax;
#endif
    *nextVal = (v > *maxVal) ? *maxVal : fmaxf(v, *nextVal);
    *maxVal = fmaxf(*maxVal, v);
}

/* cu:150-161 (ratio) and cu:48 (no ratio) */
int v3do_decide(float maxVal, float nextVal, int argMax, float minScore, float minRatio)
{
    if (minRatio < 0.0f) return (maxVal > minScore) ? argMax : -1;
#ifdef V3DO_PLANT_NO_CLAMP
    float const nextDist = 2.0f - 2.0f * nextVal;
#else
    float const nextDist = (float)fmax(0.001, (double)(2.0f - 2.0f * nextVal));
#endif
    float const bestDist = 2.0f - 2.0f * maxVal;
    float const ratio = bestDist / nextDist;
    return (maxVal > minScore && ratio < minRatio) ? argMax : -1;
}

/* ------------------------------------------------------------------ exact scores */

float v3do_exact_dot(float const* a, float const* b)
{
#ifdef V3DO_PLANT_FP32_SUM
    float s = 0.0f;
    for (int k = 0; k < D; ++k) s += a[k] * b[k];
    return s;
#else
    double s = 0.0;
    for (int k = 0; k < D; ++k) s += (double)a[k] * (double)b[k];
    return (float)s;
#endif
}

/* S[j*nL + i] = s_ij  (row-major nR x nL, the layout of the reference's score buffer) */
void v3do_scores(float const* L, int nL, float const* R, int nR, float* S)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nR; ++j)
        for (int i = 0; i < nL; ++i)
            S[(size_t)j * nL + i] = v3do_exact_dot(L + (size_t)i * D, R + (size_t)j * D);
}

/*
 * For every query q, scan candidates c = 0..nC-1 in ascending order with exact scores
 * and record (maxVal, nextVal, argMax).  Queries are processed 16 at a time with their
 * descriptors transposed to FP64 so that the k loop of each query stays sequential.
 */
void v3do_top2_scan(float const* Q, int nQ, float const* C, int nC,
                    float* best, float* next, int* arg)
{
    int const nBlocks = (nQ + QB - 1) / QB;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nBlocks; ++b) {
        double Qt[D][QB];
        int const q0 = b * QB;
        int const nq = (nQ - q0) < QB ? (nQ - q0) : QB;
        for (int k = 0; k < D; ++k)
            for (int t = 0; t < QB; ++t)
                Qt[k][t] = (t < nq) ? (double)Q[(size_t)(q0 + t) * D + k] : 0.0;
        float mx[QB], nx[QB];
        int am[QB];
        for (int t = 0; t < QB; ++t) { mx[t] = -1e30f; nx[t] = -1e30f; am[t] = -1; }
        for (int c = 0; c < nC; ++c) {
            float const* cc = C + (size_t)c * D;
            double acc[QB];
            for (int t = 0; t < QB; ++t) acc[t] = 0.0;
#ifdef V3DO_PLANT_FP32_SUM
            for (int t = 0; t < QB; ++t) {
                float f = 0.0f;
                for (int k = 0; k < D; ++k) f += cc[k] * (float)Qt[k][t];
                acc[t] = f;
            }
#else
            for (int k = 0; k < D; ++k) {
                double const a = (double)cc[k];
                for (int t = 0; t < QB; ++t) acc[t] += a * Qt[k][t];
            }
#endif
            for (int t = 0; t < QB; ++t) top2_update((float)acc[t], c, &mx[t], &nx[t], &am[t]);
        }
        for (int t = 0; t < nq; ++t) {
            best[q0 + t] = mx[t];
            next[q0 + t] = nx[t];
            arg[q0 + t] = am[t];
        }
    }
}

/* ------------------------------------------------------------------ selection on a given S */

/* per column (left feature) over rows (right features): cu:133-162 / cu:33-49 */
void v3do_column_top2(float const* S, int nRows, int nCols, float* best, float* next, int* arg)
{
    for (int col = 0; col < nCols; ++col) {
        float mx = -1e30f, nx = -1e30f;
        int am = -1;
        for (int row = 0; row < nRows; ++row) top2_update(S[(size_t)row * nCols + col], row, &mx, &nx, &am);
        best[col] = mx; next[col] = nx; arg[col] = am;
    }
}

/* per row (right feature) over columns (left features): cu:164-210 / cu:51-85 */
void v3do_row_top2(float const* S, int nRows, int nCols, float* best, float* next, int* arg)
{
    for (int row = 0; row < nRows; ++row) {
        float mx = -1e30f, nx = -1e30f;
        int am = -1;
        for (int col = 0; col < nCols; ++col) top2_update(S[(size_t)row * nCols + col], col, &mx, &nx, &am);
        best[row] = mx; next[row] = nx; arg[row] = am;
    }
}

void v3do_column_maxima(float const* S, int nRows, int nCols, float minScore, float minRatio, int* colMax)
{
    float* b = (float*)malloc(sizeof(float) * (size_t)(nCols + 1));
    float* n = (float*)malloc(sizeof(float) * (size_t)(nCols + 1));
    int* a = (int*)malloc(sizeof(int) * (size_t)(nCols + 1));
    v3do_column_top2(S, nRows, nCols, b, n, a);
    for (int c = 0; c < nCols; ++c) colMax[c] = v3do_decide(b[c], n[c], a[c], minScore, minRatio);
    free(b); free(n); free(a);
}

void v3do_row_maxima(float const* S, int nRows, int nCols, float minScore, float minRatio, int* rowMax)
{
    float* b = (float*)malloc(sizeof(float) * (size_t)(nRows + 1));
    float* n = (float*)malloc(sizeof(float) * (size_t)(nRows + 1));
    int* a = (int*)malloc(sizeof(int) * (size_t)(nRows + 1));
    v3do_row_top2(S, nRows, nCols, b, n, a);
    for (int r = 0; r < nRows; ++r) rowMax[r] = v3do_decide(b[r], n[r], a[r], minScore, minRatio);
    free(b); free(n); free(a);
}

/* cu:546-551: (k, colMax[k]) kept iff rowMax[colMax[k]] == k, in ascending k */
int v3do_mutual(int const* colMax, int nL, int const* rowMax, int nR, int32_t* pairs, int cap)
{
    int n = 0;
    for (int k = 0; k < nL; ++k) {
        int const l = colMax[k];
        if (l >= 0 && l < nR && rowMax[l] == k) {
            if (n < cap) { pairs[2 * n] = k; pairs[2 * n + 1] = l; }
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------ matcher on descriptors */

/* colMax (nL) and rowMax (nR) from descriptors, exact scores, no S stored. */
void v3do_maxima(float const* L, int nL, float const* R, int nR, float minScore, float minRatio,
                 int* colMax, int* rowMax)
{
    size_t const nmax = (size_t)(nL > nR ? nL : nR) + 1;
    float* b = (float*)malloc(sizeof(float) * nmax);
    float* n = (float*)malloc(sizeof(float) * nmax);
    int* a = (int*)malloc(sizeof(int) * nmax);
    v3do_top2_scan(L, nL, R, nR, b, n, a);
    for (int i = 0; i < nL; ++i) colMax[i] = v3do_decide(b[i], n[i], a[i], minScore, minRatio);
    v3do_top2_scan(R, nR, L, nL, b, n, a);
    for (int j = 0; j < nR; ++j) rowMax[j] = v3do_decide(b[j], n[j], a[j], minScore, minRatio);
    free(b); free(n); free(a);
}

/*
 * findPutativeMatches (cu:522-552) on nL x nR features that are already rounded.
 * Writes up to cap (left, right) pairs and returns the number of matches.
 */
int v3do_find_putative_matches(float const* L, int nL, float const* R, int nR,
                               float minScore, float minRatio, int32_t* pairs, int cap)
{
    if (nL <= 0 || nR <= 0) return 0;
    int* colMax = (int*)malloc(sizeof(int) * (size_t)nL);
    int* rowMax = (int*)malloc(sizeof(int) * (size_t)nR);
    v3do_maxima(L, nL, R, nR, minScore, minRatio, colMax, rowMax);
    int const n = v3do_mutual(colMax, nL, rowMax, nR, pairs, cap);
    free(colMax); free(rowMax);
    return n;
}

/* cu:444-460, cu:493-500: counts are rounded down to a multiple of 32 */
int v3do_round32(int n) { return n - (n & 31); }

/* cu:502, cu:518: the feature count used after setLeft/RightFeatures */
int v3do_effective_count(int n, int nMax)
{
    int const r = v3do_round32(n);
    int const m = v3do_round32(nMax);
    return r < m ? r : m;
}

/* ------------------------------------------------------------------ guided masks */

/*
 * _cuda_eliminateNonEpipolarScores (cu:277-311): S[row*nCols+col] = -1e20f where the
 * Sampson distance of (left col, right row) exceeds sqrThreshold.  F is row-major
 * double, rounded to float per element (cu:592-595).  Positions are AoS (x, y).
 * Arithmetic is plain FP32 without contraction.
 */
void v3do_mask_epipolar(float* S, int nRows, int nCols, float sqrThreshold, double const* F,
                        float const* leftXY, float const* rightXY)
{
    float f[9];
    for (int t = 0; t < 9; ++t) f[t] = (float)F[t];
    for (int col = 0; col < nCols; ++col) {
        float const px = leftXY[2 * col], py = leftXY[2 * col + 1];
        float const Fpx = f[0] * px + f[1] * py + f[2];
        float const Fpy = f[3] * px + f[4] * py + f[5];
        float const Fpz = f[6] * px + f[7] * py + f[8];
        for (int row = 0; row < nRows; ++row) {
            float const qx = rightXY[2 * row], qy = rightXY[2 * row + 1];
            float const Ftqx = f[0] * qx + f[3] * qy + f[6];
            float const Ftqy = f[1] * qx + f[4] * qy + f[7];
            float num = qx * Fpx + qy * Fpy + Fpz;
            num *= num;
            float const denom = Fpx * Fpx + Fpy * Fpy + Ftqx * Ftqx + Ftqy * Ftqy;
            float const sampson = num / denom;
            if (sampson > sqrThreshold) S[(size_t)row * nCols + col] = -1e20f;
        }
    }
}

/* _cuda_eliminateNonHomographyScores (cu:313-343) */
void v3do_mask_homography(float* S, int nRows, int nCols, float sqrThreshold, double const* H,
                          float const* leftXY, float const* rightXY)
{
    float h[9];
    for (int t = 0; t < 9; ++t) h[t] = (float)H[t];
    for (int col = 0; col < nCols; ++col) {
        float const px = leftXY[2 * col], py = leftXY[2 * col + 1];
        float Hpx = h[0] * px + h[1] * py + h[2];
        float Hpy = h[3] * px + h[4] * py + h[5];
        float const Hpz = h[6] * px + h[7] * py + h[8];
        Hpx /= Hpz;
        Hpy /= Hpz;
        for (int row = 0; row < nRows; ++row) {
            float const qx = rightXY[2 * row], qy = rightXY[2 * row + 1];
            float const sqrDist = (Hpx - qx) * (Hpx - qx) + (Hpy - qy) * (Hpy - qy);
            if (sqrDist > sqrThreshold) S[(size_t)row * nCols + col] = -1e20f;
        }
    }
}

/*
 * runGuidedMatching / runGuidedHomographyMatching (cu:554-694) on a stored exact S
 * (nR x nL, modified in place -- repeated guided calls compound, as in the reference).
 * kind 0 = epipolar (F), 1 = homography (H).
 */
int v3do_guided(float* S, int nL, int nR, int kind, double const* M, float distanceThreshold,
                float const* leftXY, float const* rightXY, float minScore, float minRatio,
                int32_t* pairs, int cap)
{
    if (nL <= 0 || nR <= 0) return 0;
    float const sqrThreshold = distanceThreshold * distanceThreshold;
    if (kind == 0) v3do_mask_epipolar(S, nR, nL, sqrThreshold, M, leftXY, rightXY);
    else v3do_mask_homography(S, nR, nL, sqrThreshold, M, leftXY, rightXY);
    int* colMax = (int*)malloc(sizeof(int) * (size_t)nL);
    int* rowMax = (int*)malloc(sizeof(int) * (size_t)nR);
    v3do_column_maxima(S, nR, nL, minScore, minRatio, colMax);
    v3do_row_maxima(S, nR, nL, minScore, minRatio, rowMax);
    int const n = v3do_mutual(colMax, nL, rowMax, nR, pairs, cap);
    free(colMax); free(rowMax);
    return n;
}

/* ------------------------------------------------------------------ vocabulary tree */

/*
 * VocabularyTree<Depth, BranchFactor, 128>::findBestLeafIndex (V3D/Base/v3d_voctree.h:63-99)
 * on the raw node array of the tree file (v3d_voctree.h:21-51: levels one after the other,
 * the children of node x at (x + 1) * branch, the root's children at 0).  Per level the
 * child with the largest score wins, score = sequential FP32 sum of the 128 products
 * (v3d_voctree.h:90, no contraction), strict '>' against a running maximum that starts at
 * -1e30 (first index on ties, NaN never wins, no winner = -1).  globalId = 0 returns the
 * visual word leafId - levelStart[depth - 1] (v3d_voctree.h:78); 1 returns leafId, what
 * VocabularyTreeTraversal::traverseTree writes (cu:784-791).  A child range outside the
 * node array (the reference would read past its buffer) returns INT32_MIN.
 */
int32_t v3do_voctree_leaf(float const* tree, int branch, int depth, float const* key, int globalId)
{
    long long nNodes = 0, len = 1, lastStart = 0;
    for (int d = 0; d < depth; ++d) {
        lastStart = nNodes;
        len *= branch;
        nNodes += len;
    }
    long long leafId = 0, next = 0;
    for (int d = 0; d < depth; ++d) {
        if (next < 0 || next + branch > nNodes) return INT32_MIN;
        float const* nodes = tree + next * D;
        int best = -1;
        float maxScore = -1e30f;
        for (int k = 0; k < branch; ++k) {
            float const* c = nodes + (size_t)k * D;
            float score = 0.0f;
            for (int i = 0; i < D; ++i) score += key[i] * c[i];
            if (score > maxScore) {
                best = k;
                maxScore = score;
            }
        }
        leafId = next + best;
        next = (leafId + 1) * branch;
    }
    return (int32_t)(globalId ? leafId : leafId - lastStart);
}

void v3do_voctree_leaves(float const* tree, int branch, int depth, float const* keys, int n, int32_t* out,
                         int globalId)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int q = 0; q < n; ++q) out[q] = v3do_voctree_leaf(tree, branch, depth, keys + (size_t)q * D, globalId);
}
