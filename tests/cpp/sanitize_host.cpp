// Host-code run under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitizers.py with g++ -fsanitize=address,undefined): the oracle's matcher,
// guided masks and tree descent (oracle/v3dm_oracle.c, compiled as C beside this file) and
// the library's robust-orientation host code (csrc/v3dm_ransac.cpp), on ragged and
// degenerate sizes -- empty sets, a single feature, counts around the 32-row granularity,
// caps smaller than the result, all-zero and NaN rows, too few correspondences.  The
// sanitizers abort on the first error; the exit status is 0 when every call returned what
// the checks below expect.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/v3dm.h"
#include "../../include/v3dm_ransac.h"

extern "C" {
int v3do_find_putative_matches(float const* L, int nL, float const* R, int nR, float minScore, float minRatio,
                               int32_t* pairs, int cap);
void v3do_maxima(float const* L, int nL, float const* R, int nR, float minScore, float minRatio, int* colMax,
                 int* rowMax);
int v3do_guided(float* S, int nL, int nR, int kind, double const* M, float distanceThreshold, float const* leftXY,
                float const* rightXY, float minScore, float minRatio, int32_t* pairs, int cap);
void v3do_scores(float const* L, int nL, float const* R, int nR, float* S);
void v3do_voctree_leaves(float const* tree, int branch, int depth, float const* keys, int n, int32_t* out,
                         int globalId);
int v3do_effective_count(int n, int nMax);
}

namespace {

int failures = 0;

void expect(bool ok, char const* what)
{
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++failures;
    }
}

uint64_t rng_state = 0x243f6a8885a308d3ull;

double uniform()
{
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

// unit rows, with every third right row a noisy copy of a left row
void features(int nL, int nR, std::vector<float>& L, std::vector<float>& R)
{
    L.assign((size_t)nL * 128, 0.0f);
    R.assign((size_t)nR * 128, 0.0f);
    for (int i = 0; i < nL; ++i) {
        double s = 0;
        for (int k = 0; k < 128; ++k) s += (L[(size_t)i * 128 + k] = (float)uniform()) * L[(size_t)i * 128 + k];
        for (int k = 0; k < 128; ++k) L[(size_t)i * 128 + k] /= (float)std::sqrt(s);
    }
    for (int j = 0; j < nR; ++j) {
        double s = 0;
        for (int k = 0; k < 128; ++k) {
            float v = (float)uniform();
            if (nL > 0 && j % 3 == 0) v = L[(size_t)(j % nL) * 128 + k] + 0.02f * (float)uniform();
            R[(size_t)j * 128 + k] = v;
            s += (double)v * v;
        }
        for (int k = 0; k < 128; ++k) R[(size_t)j * 128 + k] /= (float)std::sqrt(s);
    }
}

void matcher_cases()
{
    int const sizes[] = {0, 1, 5, 31, 32, 33, 97};
    for (int nL : sizes)
        for (int nR : sizes) {
            std::vector<float> L, R;
            features(nL, nR, L, R);
            int const cap = 3;  // smaller than most results: the count is still the full one
            std::vector<int32_t> pairs(2 * cap);
            int const n = v3do_find_putative_matches(L.data(), nL, R.data(), nR, 0.9f, 0.8f, pairs.data(), cap);
            std::vector<int32_t> all(2 * (size_t)(nL > 0 ? nL : 1));
            int const n2 = v3do_find_putative_matches(L.data(), nL, R.data(), nR, 0.9f, 0.8f, all.data(), nL);
            expect(n == n2 && n >= 0 && n <= (nL < nR ? nL : nR), "match count independent of the cap");
            for (int k = 0; k < (n < cap ? n : cap); ++k)
                expect(pairs[2 * k] == all[2 * k] && pairs[2 * k + 1] == all[2 * k + 1], "capped pairs are a prefix");
            if (nL > 0 && nR > 0) {
                std::vector<int> colMax(nL), rowMax(nR);
                v3do_maxima(L.data(), nL, R.data(), nR, 0.9f, -1.0f, colMax.data(), rowMax.data());
                for (int i = 0; i < nL; ++i) expect(colMax[i] >= -1 && colMax[i] < nR, "column decision in range");
                // guided: scores, a homography mask, then an epipolar mask on the same matrix
                std::vector<float> S((size_t)nL * nR);
                v3do_scores(L.data(), nL, R.data(), nR, S.data());
                std::vector<float> xyL(2 * (size_t)nL), xyR(2 * (size_t)nR);
                for (float& v : xyL) v = (float)(640.0 * uniform());
                for (float& v : xyR) v = (float)(640.0 * uniform());
                double const H[9] = {1.0, 0.01, 3.0, -0.01, 1.0, -2.0, 1e-5, 0.0, 1.0};
                double const F[9] = {0.0, -1e-4, 0.02, 1e-4, 0.0, -0.03, -0.02, 0.03, 1.0};
                int const g1 = v3do_guided(S.data(), nL, nR, 1, H, 50.0f, xyL.data(), xyR.data(), 0.9f, 0.8f,
                                           pairs.data(), cap);
                int const g2 = v3do_guided(S.data(), nL, nR, 0, F, 5.0f, xyL.data(), xyR.data(), 0.9f, 0.8f,
                                           pairs.data(), cap);
                expect(g1 >= 0 && g2 >= 0, "guided counts");
            }
        }
    // an all-zero row and a NaN row on each side
    std::vector<float> L, R;
    features(40, 40, L, R);
    for (int k = 0; k < 128; ++k) L[3 * 128 + k] = 0.0f, R[7 * 128 + k] = NAN;
    std::vector<int32_t> pairs(80);
    int const n = v3do_find_putative_matches(L.data(), 40, R.data(), 40, 0.9f, 0.8f, pairs.data(), 40);
    expect(n >= 0 && n <= 40, "degenerate rows");
    expect(v3do_effective_count(33, 32) == 32 && v3do_effective_count(31, 64) == 0, "rounding of counts");
}

void tree_cases()
{
    int const branch = 3, depth = 3;
    int const nNodes = 3 + 9 + 27;
    std::vector<float> tree((size_t)nNodes * 128);
    for (float& v : tree) v = (float)uniform() - 0.5f;
    for (int n : {0, 1, 64, 129}) {
        std::vector<float> keys((size_t)(n > 0 ? n : 1) * 128);
        for (float& v : keys) v = (float)uniform() - 0.5f;
        if (n > 1)
            for (int k = 0; k < 128; ++k) keys[128 + k] = NAN;  // NaN scores never win: no child
        std::vector<int32_t> out(n > 0 ? n : 1, -7);
        v3do_voctree_leaves(tree.data(), branch, depth, keys.data(), n, out.data(), 0);
        for (int q = 0; q < n; ++q) expect(out[q] >= -28 && out[q] < 27, "visual word in range");
    }
}

void ransac_cases()
{
    double const K[9] = {800.0, 0.0, 320.0, 0.0, 800.0, 240.0, 0.0, 0.0, 1.0};
    // a scene: x2 ~ R x1 + T with R a small rotation about y, T mostly along x
    double const c = std::cos(0.1), s = std::sin(0.1);
    double const Rt[9] = {c, 0.0, s, 0.0, 1.0, 0.0, -s, 0.0, c};
    double const Tt[3] = {0.98, 0.1, 0.17};
    for (int n : {5, 6, 11, 300}) {
        std::vector<double> corrs(4 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            double const X[3] = {2.0 * uniform() - 1.0, 2.0 * uniform() - 1.0, 3.0 + 5.0 * uniform()};
            double Y[3];
            for (int r = 0; r < 3; ++r) Y[r] = Rt[3 * r] * X[0] + Rt[3 * r + 1] * X[1] + Rt[3 * r + 2] * X[2] + Tt[r];
            corrs[4 * i + 0] = 800.0 * X[0] / X[2] + 320.0;
            corrs[4 * i + 1] = 800.0 * X[1] / X[2] + 240.0;
            corrs[4 * i + 2] = 800.0 * Y[0] / Y[2] + 320.0;
            corrs[4 * i + 3] = 800.0 * Y[1] / Y[2] + 240.0;
            if (i % 4 == 3) corrs[4 * i + 2] = 640.0 * uniform();  // outliers
        }
        for (int refine = 0; refine < 2; ++refine) {
            v3dm_orientation res;
            std::vector<int32_t> inl(n);
            int const rc = v3dm_robust_orientation(corrs.data(), n, K, K, 2.0, 200, 1, refine, 7, &res, inl.data());
            expect(rc == V3DM_OK && res.nInliers >= 0 && res.nInliers <= n, "robust orientation");
            for (int k = 1; k < res.nInliers; ++k) expect(inl[k - 1] < inl[k], "inliers ascending");
        }
        std::vector<double> err(n);
        double const F[9] = {0.0, -1e-4, 0.02, 1e-4, 0.0, -0.03, -0.02, 0.03, 1.0};
        expect(v3dm_sampson_errors(F, corrs.data(), n, err.data()) == V3DM_OK, "sampson errors");
    }
    v3dm_orientation res;
    double const few[16] = {0};
    expect(v3dm_robust_orientation(few, 4, K, K, 2.0, 100, 0, 0, 1, &res, nullptr) == V3DM_ERR_INVALID,
           "four correspondences are refused");
    expect(v3dm_sampson_errors(K, nullptr, 0, nullptr) == V3DM_OK, "no correspondences");
    // five-point: random samples and a degenerate one (five copies of one point)
    for (int t = 0; t < 200; ++t) {
        double x1[5], y1[5], x2[5], y2[5], Es[90];
        for (int k = 0; k < 5; ++k) {
            x1[k] = uniform() - 0.5, y1[k] = uniform() - 0.5, x2[k] = uniform() - 0.5, y2[k] = uniform() - 0.5;
            if (t == 0) x1[k] = y1[k] = x2[k] = y2[k] = 0.1;
        }
        int nE = -1;
        expect(v3dm_essentials_from_five(x1, y1, x2, y2, Es, &nE) == V3DM_OK && nE >= 0 && nE <= 10, "five-point");
        for (int e = 0; e < nE; ++e) {
            double R[9], T[3];
            int ok = -1;
            expect(v3dm_relative_pose_from_essential(Es + 9 * e, 5, x1, y1, x2, y2, R, T, &ok) == V3DM_OK,
                   "pose from E");
        }
    }
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, T[3] = {1, 0, 0};
    expect(v3dm_refine_relative_pose(nullptr, 0, K, K, R, T, 2.0) == V3DM_OK, "refinement without correspondences");
}

}  // namespace

int main()
{
    matcher_cases();
    tree_cases();
    ransac_cases();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized host run: ok\n");
    return 0;
}
