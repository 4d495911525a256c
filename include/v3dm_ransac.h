/* Robust relative orientation from putative matches: the downstream consumer of the match
 * lists (host code, double precision).
 *
 * Reference: computeRobustOrientationMLE (V3D/Apps/SfM/reconstruction_common.cpp:139-340,
 * declared at V3D/Apps/SfM/reconstruction_common.h:682-687) and its building blocks:
 * sampsonEpipolarError (reconstruction_common.h:588-608), ransacNSamples
 * (reconstruction_common.h:642-650), getShuffleCount (V3D/Geometry/v3d_homography.cpp:14-18),
 * computeEssentialsFromFiveCorrs (V3D/Geometry/v3d_poseutilities.cpp:1015-1060),
 * relativePoseFromEssential (v3d_poseutilities.cpp:1263-1291) and refineRelativePose
 * (v3d_poseutilities.cpp:1063-1099, minimiser at :812-1006 over
 * V3D/Math/v3d_optimization.h:82-277).
 *
 * Conventions (the reference's): x2 ~ R x1 + T for normalised image points x = K^-1 (u, v,
 * 1); E = [T]x R; F = K2^-T E K1^-1; matrices row-major; T of unit length.  Correspondences
 * are rows of four doubles (left u, left v, right u, right v) in pixels; where the reference
 * takes them as PointCorrespondence (v3dm_sampson_errors, v3dm_robust_orientation) they are
 * used as it holds them, rounded to float (PointMeasurement::pos is a Vector2f,
 * V3D/Geometry/v3d_mviewutilities.h:14-17), and the arithmetic follows the reference's
 * operation order, so that the results are the reference's to rounding for the same
 * samples.
 *
 * Return codes are the V3DM_* codes of v3dm.h; messages are not kept (no handle). */
#ifndef V3DM_RANSAC_H
#define V3DM_RANSAC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct v3dm_orientation {
    double essential[9];
    double fundamental[9];
    double rotation[9];
    double translation[3];
    double error;  /* sum over all correspondences of min(Sampson error, threshold^2) / N of
                      the best hypothesis (RobustOrientationResult::error); +inf if none */
    int nInliers;  /* correspondences with Sampson error < threshold^2 under `fundamental` */
    int nSamplesDrawn;
    int nHypotheses;
} v3dm_orientation;

/* sampsonEpipolarError(corr, F) (reconstruction_common.h:588-608) of n correspondences:
 * (q^T F p)^2 / ((Fp)_1^2 + (Fp)_2^2 + (F^T q)_1^2 + (F^T q)_2^2). */
int v3dm_sampson_errors(const double F[9], const double* corrs, int n, double* out);

/* computeEssentialsFromFiveCorrs (v3d_poseutilities.cpp:1015-1060): every real essential
 * matrix (at most 10, row-major in Es[9 k .. 9 k + 8]) compatible with five correspondences
 * of normalised points (x1, y1) <-> (x2, y2).  Solved here as the eigenproblem of the action
 * matrix of multiplication by x on the quotient ring of the ten cubic constraints
 * (det E = 0, 2 E E^T E - tr(E E^T) E = 0) over E = x X + y Y + z Z + W, {X, Y, Z, W} the
 * null space of the five epipolar constraints. */
int v3dm_essentials_from_five(const double x1[5], const double y1[5], const double x2[5], const double y2[5],
                              double Es[90], int* nSolutions);

/* relativePoseFromEssential(E, n, x1, y1, x2, y2, R, T) (v3d_poseutilities.cpp:1263-1291):
 * the decomposition of E whose triangulated points lie in front of both cameras for all n
 * correspondences, decided point by point as the reference does (Nister's singular vectors
 * from cross products of the rows of E, exact triangulation and the twisted-pair test).
 * *ok = 0 (R, T zero) when the points disagree on the configuration. */
int v3dm_relative_pose_from_essential(const double E[9], int n, const double* x1, const double* y1,
                                      const double* x2, const double* y2, double R[9], double T[3], int* ok);

/* refineRelativePose (v3d_poseutilities.cpp:1063-1099): Levenberg-Marquardt over (R, T) on
 * the signed Sampson residuals of the normalised points, reweighted by the pixel-space
 * Sampson error against sampsonInlierThreshold; R and T are updated in place.  The
 * reference's optimiser schedule, including its Rodrigues map, which takes a rotation step
 * of at most 1e-6 rad as the identity (V3D/Math/v3d_mathutilities.h:170-190).  Positions as
 * given (the reference takes them as Vector2d here). */
int v3dm_refine_relative_pose(const double* corrs, int n, const double K1[9], const double K2[9], double R[9],
                              double T[3], double sampsonInlierThreshold);

/* computeRobustOrientationMLE(corrs, K1, K2, inlierThreshold, nSamples, res, reportInliers,
 * mode) (reconstruction_common.cpp:139-340): RANSAC over five-point samples (consecutive
 * quintuples of shuffled index lists), truncated Sampson cost, pose from the best E,
 * optional iterative refinement, adaptive sample count (confidence 1 - 1e-9).  The shuffles
 * come from a seeded generator of this library (the reference uses std::random_shuffle):
 * the same seed gives the same result.  inliers (may be NULL when !reportInliers): the
 * indices with Sampson error < inlierThreshold^2 under the result's F, in ascending order,
 * res->nInliers of them (capacity n).  V3DM_ERR_INVALID for n < 5 (the reference throws). */
int v3dm_robust_orientation(const double* corrs, int n, const double K1[9], const double K2[9],
                            double inlierThreshold, int nSamples, int reportInliers, int iterativeRefinement,
                            uint64_t seed, v3dm_orientation* res, int32_t* inliers);

#ifdef __cplusplus
}
#endif

#endif /* V3DM_RANSAC_H */
