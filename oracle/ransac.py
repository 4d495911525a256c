"""Two-view geometry checks for the robust-orientation consumer -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and by the CPU measurement leg of bench.py, never by the library.  Each
function restates in numpy (float64) what the reference computes, citing it, or states a
property that the mathematics fixes (the epipolar and essential-matrix constraints, the
pose of a synthetic scene), so that the library's results are checked against something
it does not share code with:

* sampson_error: sampsonEpipolarError (V3D/Apps/SfM/reconstruction_common.h:588-608).
* shuffle_count: getShuffleCount (V3D/Geometry/v3d_homography.cpp:14-18).
* ransac_n_samples: ransacNSamples(nNeededSamples, outlierRatio, wantedPsucces)
  (reconstruction_common.h:642-650).
* essential_residuals: det E, 2 E E^T E - tr(E E^T) E and the epipolar constraints (the ten
  cubic constraints and the five linear ones that define computeEssentialsFromFiveCorrs'
  solutions, v3d_poseutilities.cpp:1015-1060).
* synthetic_two_view: a scene with known K, R, T (x2 ~ R x1 + T), Gaussian pixel noise
  and uniform outliers; rotation_error_deg / direction_error_deg measure a recovered pose.
"""
import numpy as np


def sampson_error(F, corrs):
    """Element by element in the reference's order (multiply_A_v_affine, then the numerator
    and denominator sums left to right), so that the float64 roundings are the same."""
    F = np.asarray(F, np.float64).reshape(3, 3)
    c = np.asarray(corrs, np.float64).reshape(-1, 4)
    px, py, qx, qy = c[:, 0], c[:, 1], c[:, 2], c[:, 3]
    fp0 = F[0, 0] * px + F[0, 1] * py + F[0, 2]
    fp1 = F[1, 0] * px + F[1, 1] * py + F[1, 2]
    fp2 = F[2, 0] * px + F[2, 1] * py + F[2, 2]
    ftq0 = F[0, 0] * qx + F[1, 0] * qy + F[2, 0]
    ftq1 = F[0, 1] * qx + F[1, 1] * qy + F[2, 1]
    tmp = qx * fp0 + qy * fp1 + fp2
    return (tmp * tmp) / (fp0 * fp0 + fp1 * fp1 + ftq0 * ftq0 + ftq1 * ftq1)


def shuffle_count(n, n_rounds, min_sample_size):
    return max(1, int(np.float32(n_rounds) * np.float32(min_sample_size) / np.float32(n) + np.float32(0.5)))


def ransac_n_samples(n_needed, outlier_ratio, wanted):
    max_n = 1e8
    if outlier_ratio >= 1.0:
        return int(max_n)
    with np.errstate(divide="ignore"):
        n = np.log(1.0 - wanted) / np.log(1.0 - (1.0 - outlier_ratio) ** n_needed)
    return int(max(0.0, min(max_n, n)))


def skew(t):
    return np.array([[0.0, -t[2], t[1]], [t[2], 0.0, -t[0]], [-t[1], t[0], 0.0]])


def essential_from_pose(R, T):
    return skew(T) @ R


def essential_residuals(E, x1, y1, x2, y2):
    """(|det E|, |2 E E^T E - tr(E E^T) E|_max, max |x2^T E x1|) for unit-norm E."""
    E = np.asarray(E, np.float64) / np.linalg.norm(E)
    EEt = E @ E.T
    cubic = 2.0 * EEt @ E - np.trace(EEt) * E
    p = np.c_[x1, y1, np.ones(len(x1))]
    q = np.c_[x2, y2, np.ones(len(x2))]
    epi = np.abs((q @ E * p).sum(1)).max()
    return abs(np.linalg.det(E)), np.abs(cubic).max(), epi


def rotation(w):
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    if th == 0.0:
        return np.eye(3)
    K = skew(w / th)
    return np.eye(3) + np.sin(th) * K + (1.0 - np.cos(th)) * K @ K


def synthetic_two_view(rng, n_in, n_out, K, R, T, noise_px, depth=(4.0, 12.0), size=(640, 480)):
    """Correspondences (n, 4) in pixels, shuffled, and the inlier mask."""
    X = np.c_[rng.uniform(-3, 3, (n_in, 2)), rng.uniform(depth[0], depth[1], n_in)]
    u1 = X @ K.T
    u1 = u1[:, :2] / u1[:, 2:3]
    X2 = X @ R.T + T
    u2 = X2 @ K.T
    u2 = u2[:, :2] / u2[:, 2:3]
    u1 = u1 + rng.normal(0, noise_px, u1.shape)
    u2 = u2 + rng.normal(0, noise_px, u2.shape)
    o = rng.uniform([0, 0, 0, 0], [size[0], size[1], size[0], size[1]], (n_out, 4))
    corrs = np.r_[np.c_[u1, u2], o]
    perm = rng.permutation(len(corrs))
    return corrs[perm], perm < n_in


def rotation_error_deg(Ra, Rb):
    c = (np.trace(np.asarray(Ra).T @ np.asarray(Rb)) - 1.0) / 2.0
    return float(np.degrees(np.arccos(np.clip(c, -1.0, 1.0))))


def direction_error_deg(ta, tb):
    ta = np.asarray(ta) / np.linalg.norm(ta)
    tb = np.asarray(tb) / np.linalg.norm(tb)
    return float(np.degrees(np.arccos(np.clip(ta @ tb, -1.0, 1.0))))
