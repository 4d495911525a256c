// C entry points over the reference's own two-view pose utilities -- TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/build.py (build_ref) together with the reference's sources where they
// lie (V3D/Geometry/v3d_poseutilities.cpp, V3D/Math/v3d_mathutilities.cpp,
// V3D/Math/v3d_optimization.cpp) into oracle/_ref/libref_pose.so; loaded only by tests, to
// check the library's robust-orientation building blocks against the reference's:
//
//   ref_essentials_from_five          computeEssentialsFromFiveCorrs (v3d_poseutilities.cpp:1015-1060)
//   ref_relative_pose_from_essential  relativePoseFromEssential (v3d_poseutilities.cpp:1263-1291)
//   ref_refine_relative_pose          refineRelativePose (v3d_poseutilities.cpp:1063-1099)
//
// Matrices cross this boundary row-major, correspondences as rows (left u, left v, right u,
// right v), the layout of the library's include/v3dm_ransac.h.
#include "Geometry/v3d_poseutilities.h"

#include <vector>

using namespace V3D;

namespace {

Matrix3x3d mat_in(double const* a)
{
    Matrix3x3d m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = a[3 * r + c];
    return m;
}

void mat_out(Matrix3x3d const& m, double* a)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[3 * r + c] = m[r][c];
}

}  // namespace

extern "C" {

// Every solution, row-major in Es[9 k ..]; returns the count (at most 10), -1 on an exception.
int ref_essentials_from_five(double const* x1, double const* y1, double const* x2, double const* y2, double* Es)
{
    std::vector<Matrix3x3d> sols;
    try {
        computeEssentialsFromFiveCorrs(x1, y1, x2, y2, sols);
    } catch (...) {
        return -1;
    }
    int const n = (int)sols.size() < 10 ? (int)sols.size() : 10;
    for (int k = 0; k < n; ++k) mat_out(sols[k], Es + 9 * k);
    return n;
}

// 1 and (R, T) when every correspondence gives the same configuration, else 0.
int ref_relative_pose_from_essential(double const* E, int n, double const* x1, double const* y1, double const* x2,
                                     double const* y2, double* R, double* T)
{
    Matrix3x3d Rm;
    Vector3d Tv;
    bool const ok = relativePoseFromEssential(mat_in(E), n, x1, y1, x2, y2, Rm, Tv);
    mat_out(Rm, R);
    for (int i = 0; i < 3; ++i) T[i] = Tv[i];
    return ok ? 1 : 0;
}

// (R, T) refined in place on n pixel correspondences.
int ref_refine_relative_pose(double const* corrs, int n, double const* K1, double const* K2, double* R, double* T,
                             double sampsonInlierThreshold)
{
    std::vector<Vector2d> left(n), right(n);
    for (int i = 0; i < n; ++i) {
        left[i][0] = corrs[4 * i + 0];
        left[i][1] = corrs[4 * i + 1];
        right[i][0] = corrs[4 * i + 2];
        right[i][1] = corrs[4 * i + 3];
    }
    Matrix3x3d Rm = mat_in(R);
    Vector3d Tv;
    for (int i = 0; i < 3; ++i) Tv[i] = T[i];
    bool const ok = refineRelativePose(left, right, mat_in(K1), mat_in(K2), Rm, Tv, sampsonInlierThreshold);
    mat_out(Rm, R);
    for (int i = 0; i < 3; ++i) T[i] = Tv[i];
    return ok ? 1 : 0;
}

}  // extern "C"
