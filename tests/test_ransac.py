"""Robust relative orientation from match lists (computeRobustOrientationMLE,
V3D/Apps/SfM/reconstruction_common.cpp:139-340) over the C ABI (include/v3dm_ransac.h):
host code, so these tests run without a GPU.

Pinned by: the reference's Sampson formula restated in numpy; the constraints that define
an essential matrix and the known pose of synthetic scenes (the five-point solutions must
contain the true E, the decomposition must return the true R, T); the reference's sample
count formulas (getShuffleCount, ransacNSamples) through the number of samples the loop
draws.  The RANSAC result itself depends on the shuffles (the reference's come from
std::random_shuffle): parity is on the components and on the recovered pose of scenes
with a known answer, not on the sequence of samples.
"""
import numpy as np
import pytest

from oracle import ransac as OR
from v3dmatch import ransac
from v3dmatch._native import V3DMError

K = np.array([[800.0, 0.0, 320.0], [0.0, 800.0, 240.0], [0.0, 0.0, 1.0]])


def _scene(rng, n_pts=5):
    R = OR.rotation(rng.normal(size=3) * 0.3)
    T = rng.normal(size=3)
    T /= np.linalg.norm(T)
    X = np.c_[rng.uniform(-1, 1, (n_pts, 2)), rng.uniform(2, 6, n_pts)]
    x1 = X[:, :2] / X[:, 2:3]
    X2 = X @ R.T + T
    x2 = X2[:, :2] / X2[:, 2:3]
    return R, T, x1, x2


def test_sampson_errors_match_reference_formula():
    rng = np.random.default_rng(0)
    F = rng.normal(size=(3, 3))
    corrs = rng.uniform(0, 640, (1000, 4))
    got = ransac.sampson_errors(F, corrs)
    want = OR.sampson_error(F, corrs)
    assert np.array_equal(got, want)  # same operations in the same order
    # a correspondence on its epipolar line scores 0, up to the single-precision rounding of
    # the positions (the reference's PointMeasurement holds floats)
    R, T, x1, x2 = _scene(rng, 3)
    E = OR.essential_from_pose(R, T)
    c = np.c_[x1, x2]
    assert np.abs(ransac.sampson_errors(E, c)).max() < 1e-14
    assert np.abs(ransac.sampson_errors(E, c.astype(np.float32).astype(np.float64))).max() < 1e-14


def test_five_point_contains_true_essential():
    rng = np.random.default_rng(1)
    counts = []
    for _ in range(100):
        R, T, x1, x2 = _scene(rng)
        Es = ransac.essentials_from_five(x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
        counts.append(len(Es))
        assert 1 <= len(Es) <= 10
        Et = OR.essential_from_pose(R, T)
        Et /= np.linalg.norm(Et)
        d = min(min(np.linalg.norm(E / np.linalg.norm(E) - Et), np.linalg.norm(E / np.linalg.norm(E) + Et))
                for E in Es)
        assert d < 1e-7
        for E in Es:
            det, cubic, epi = OR.essential_residuals(E, x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
            assert det < 1e-7 and cubic < 1e-7 and epi < 1e-9
    assert max(counts) > 2  # several real solutions occur


def test_five_point_degenerate_sample():
    # five copies of one correspondence: rank-1 constraints, no solution is returned
    x = np.full(5, 0.1)
    assert len(ransac.essentials_from_five(x, x, x, x)) == 0


def test_pose_from_essential():
    rng = np.random.default_rng(2)
    for _ in range(50):
        R, T, x1, x2 = _scene(rng)
        E = OR.essential_from_pose(R, T) * rng.uniform(0.5, 3.0) * rng.choice([-1.0, 1.0])
        Rr, Tr, ok = ransac.relative_pose_from_essential(E, x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
        assert ok
        assert np.allclose(Rr, R, atol=1e-9) and np.allclose(Tr, T, atol=1e-9)
    # points that disagree on the configuration (one seen behind the cameras): no pose
    R, T, x1, x2 = _scene(rng)
    E = OR.essential_from_pose(R, T)
    X = np.r_[x1[0], 1.0] * -3.0  # a point behind camera 1 on the same ray
    X2 = R @ X + T
    x1b = np.r_[x1[:4], [X[:2] / X[2]]]
    x2b = np.r_[x2[:4], [X2[:2] / X2[2]]]
    Rr, Tr, ok = ransac.relative_pose_from_essential(E, x1b[:, 0], x1b[:, 1], x2b[:, 0], x2b[:, 1])
    assert not ok and not Rr.any() and not Tr.any()


def test_refine_relative_pose_reduces_error():
    rng = np.random.default_rng(3)
    R = OR.rotation([0.05, -0.2, 0.03])
    T = np.array([1.0, 0.1, 0.2]) / np.linalg.norm([1.0, 0.1, 0.2])
    corrs, inl = OR.synthetic_two_view(rng, 800, 0, K, R, T, 0.3)
    R0 = R @ OR.rotation([0.01, -0.008, 0.012])
    T0 = T + np.array([0.03, -0.02, 0.01])
    T0 /= np.linalg.norm(T0)
    Rr, Tr = ransac.refine_relative_pose(corrs, K, K, R0, T0, 2.0)
    assert OR.rotation_error_deg(Rr, R) < 0.2 * OR.rotation_error_deg(R0, R)
    assert OR.direction_error_deg(Tr, T) < 0.2 * OR.direction_error_deg(T0, T)
    assert abs(np.linalg.norm(Tr) - 1.0) < 1e-12 and abs(np.linalg.det(Rr) - 1.0) < 1e-12


@pytest.mark.parametrize("refine", [False, True])
def test_robust_orientation_recovers_scene(refine):
    rng = np.random.default_rng(4)
    R = OR.rotation([0.05, -0.2, 0.03])
    T = np.array([1.0, 0.1, 0.2]) / np.linalg.norm([1.0, 0.1, 0.2])
    corrs, is_in = OR.synthetic_two_view(rng, 1500, 700, K, R, T, 0.5)
    res = ransac.robust_orientation(corrs, K, K, 2.0, 1000, True, refine, seed=7)
    got = np.zeros(len(corrs), bool)
    got[res["inliers"]] = True
    recall = (got & is_in).sum() / is_in.sum()
    precision = (got & is_in).sum() / got.sum()
    rot, tdir = OR.rotation_error_deg(res["rotation"], R), OR.direction_error_deg(res["translation"], T)
    print(f"refine={refine}: rotation {rot:.3f} deg, direction {tdir:.3f} deg, recall {recall:.4f}, "
          f"precision {precision:.4f}, samples {res['samples']}")
    assert rot < (0.2 if refine else 1.0) and tdir < (0.5 if refine else 3.0)
    assert recall > (0.99 if refine else 0.85) and precision > 0.99
    # the reported matrices are consistent with the pose, and the inliers with F
    E = OR.essential_from_pose(res["rotation"], res["translation"])
    Er = res["essential"] / np.linalg.norm(res["essential"])
    assert min(np.linalg.norm(Er - E / np.linalg.norm(E)), np.linalg.norm(Er + E / np.linalg.norm(E))) < 1e-6
    Kinv = np.linalg.inv(K)
    F = Kinv.T @ res["essential"] @ Kinv
    assert np.allclose(res["fundamental"] / np.linalg.norm(res["fundamental"]), F / np.linalg.norm(F), atol=1e-9)
    want = np.flatnonzero(OR.sampson_error(res["fundamental"], corrs) < 4.0)
    assert np.array_equal(res["inliers"], want)
    # the same seed gives the same result
    again = ransac.robust_orientation(corrs, K, K, 2.0, 1000, True, refine, seed=7)
    assert np.array_equal(again["rotation"], res["rotation"]) and np.array_equal(again["inliers"], res["inliers"])


def test_sample_counts_follow_reference_formulas():
    rng = np.random.default_rng(5)
    R = OR.rotation([0.02, 0.1, -0.05])
    T = np.array([0.2, 1.0, 0.1]) / np.linalg.norm([0.2, 1.0, 0.1])
    # exact inliers only: the first pose leaves no outliers, ransacNSamples(5, 0, 1 - 1e-9)
    # is 0 and the loop stops after that sample
    assert OR.ransac_n_samples(5, 0.0, 1.0 - 10e-10) == 0
    corrs, _ = OR.synthetic_two_view(rng, 200, 0, K, R, T, 0.0)
    res = ransac.robust_orientation(corrs, K, K, 1.0, 1000, True, False, seed=1)
    # (the error is what rounding the positions to float, as the reference stores them, leaves)
    assert res["samples"] == 1 and res["n_inliers"] == 200 and res["error"] < 1e-8
    # random correspondences: no early stop; getShuffleCount shuffles of floor((N - 1) / 5)
    # quintuples each, or nSamples + 1 samples, whichever is smaller
    for n, n_samples in ((12, 1000), (100, 7), (37, 40)):
        corrs = rng.uniform(0, 640, (n, 4))
        res = ransac.robust_orientation(corrs, K, K, 0.5, n_samples, False, False, seed=2)
        per_shuffle = len(range(0, n - 5, 5))
        want = min(OR.shuffle_count(n, n_samples, 5) * per_shuffle, n_samples + 1)
        assert res["samples"] == want, (n, n_samples, res["samples"], want)


def test_too_few_correspondences():
    with pytest.raises(V3DMError):
        ransac.robust_orientation(np.zeros((4, 4)), K, K, 1.0)
    with pytest.raises(V3DMError):
        ransac.robust_orientation(np.ones((10, 4)), np.zeros((3, 3)), K, 1.0)
