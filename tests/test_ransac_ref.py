"""The robust-orientation building blocks against the reference's own code, compiled from its
sources into oracle/_ref/libref_pose.so (oracle/build.py, build_ref): the five-point solver
(computeEssentialsFromFiveCorrs, V3D/Geometry/v3d_poseutilities.cpp:1015-1060), the pose
from an essential matrix (relativePoseFromEssential, :1263-1291) and the refinement
(refineRelativePose, :1063-1099).  Host code on both sides, so these tests run without a
GPU; they are skipped where the reference tree was not there to build from.

The library solves the five-point problem by another method (the action matrix of the
quotient ring, where the reference eliminates to a tenth-degree polynomial), so the two
solution lists are compared as sets of unit-norm matrices up to sign.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import ransac as OR
from v3dmatch import ransac

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libref_pose.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference pose library not built "
                                "(the reference sources were absent at build time)")

K = np.array([[800.0, 0.0, 320.0], [0.0, 800.0, 240.0], [0.0, 0.0, 1.0]])
D = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def ref():
    lib = C.CDLL(REF)
    lib.ref_essentials_from_five.restype = C.c_int
    lib.ref_essentials_from_five.argtypes = [D, D, D, D, D]
    lib.ref_relative_pose_from_essential.restype = C.c_int
    lib.ref_relative_pose_from_essential.argtypes = [D, C.c_int, D, D, D, D, D, D]
    lib.ref_refine_relative_pose.restype = C.c_int
    lib.ref_refine_relative_pose.argtypes = [D, C.c_int, D, D, D, D, C.c_double]
    return lib


def _p(a):
    return a.ctypes.data_as(D)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def ref_five(lib, x1, y1, x2, y2):
    args = [_f64(v) for v in (x1, y1, x2, y2)]
    out = np.zeros(90)
    n = lib.ref_essentials_from_five(*[_p(a) for a in args], _p(out))
    assert n >= 0
    return [out[9 * k:9 * k + 9].reshape(3, 3) for k in range(n)]


def ref_pose(lib, E, x1, y1, x2, y2):
    args = [_f64(v) for v in (x1, y1, x2, y2)]
    R, T = np.zeros(9), np.zeros(3)
    ok = lib.ref_relative_pose_from_essential(_p(_f64(E).ravel()), len(args[0]), *[_p(a) for a in args], _p(R),
                                              _p(T))
    return R.reshape(3, 3), T, bool(ok)


def _unit(E):
    E = np.asarray(E, np.float64) / np.linalg.norm(E)
    return E * np.sign(E.ravel()[np.argmax(np.abs(E.ravel()))])


def _scene(rng, n_pts=5):
    R = OR.rotation(rng.normal(size=3) * 0.3)
    T = rng.normal(size=3)
    T /= np.linalg.norm(T)
    X = np.c_[rng.uniform(-1, 1, (n_pts, 2)), rng.uniform(2, 6, n_pts)]
    x1 = X[:, :2] / X[:, 2:3]
    X2 = X @ R.T + T
    x2 = X2[:, :2] / X2[:, 2:3]
    return R, T, x1, x2


def test_five_point_same_solutions_as_reference(ref):
    """Same real solutions as the reference's solver, sample by sample: well-separated
    solutions of both lists pair up to 1e-6; a count may differ only where the reference's
    polynomial route finds a root pair near the real axis that the other method does not
    (or the reverse), which happens on a small fraction of samples."""
    rng = np.random.default_rng(11)
    n_same = n_total = 0
    for trial in range(300):
        if trial % 3 == 2:  # samples of random points (outlier quintuples of a RANSAC run)
            x1, x2 = rng.uniform(-0.5, 0.5, (5, 2)), rng.uniform(-0.5, 0.5, (5, 2))
        else:
            _, _, x1, x2 = _scene(rng)
        mine = [_unit(E) for E in ransac.essentials_from_five(x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])]
        theirs = [_unit(E) for E in ref_five(ref, x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])]
        n_total += 1
        matched = sum(1 for E in theirs if mine and min(np.linalg.norm(E - M) for M in mine) < 1e-6)
        if len(mine) == len(theirs) == matched:
            n_same += 1
        # every solution of the reference that satisfies the essential constraints well is
        # one of the library's
        for E in theirs:
            det, cubic, epi = OR.essential_residuals(E, x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
            if det < 1e-9 and cubic < 1e-9 and epi < 1e-10:
                assert mine and min(np.linalg.norm(E - M) for M in mine) < 1e-6, trial
    print(f"identical solution sets on {n_same} of {n_total} samples")
    assert n_same >= 0.97 * n_total


def test_pose_from_essential_same_as_reference(ref):
    """The decomposition of the reference for the true E and for every five-point solution of
    random samples: the same decision (one configuration for all points or not) and the
    same R, T."""
    rng = np.random.default_rng(12)
    n_ok = 0
    for trial in range(200):
        R, T, x1, x2 = _scene(rng, 8)
        Es = [OR.essential_from_pose(R, T) * rng.uniform(0.5, 3.0) * rng.choice([-1.0, 1.0])]
        Es += list(ref_five(ref, x1[:5, 0], x1[:5, 1], x2[:5, 0], x2[:5, 1]))
        for E in Es:
            Rm, Tm, okm = ransac.relative_pose_from_essential(E, x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
            Rr, Tr, okr = ref_pose(ref, E, x1[:, 0], x1[:, 1], x2[:, 0], x2[:, 1])
            assert okm == okr, trial
            if okr:
                n_ok += 1
                assert np.allclose(Rm, Rr, atol=1e-9) and np.allclose(Tm, Tr, atol=1e-9), trial
            else:
                assert not Rm.any() and not Tm.any()
    assert n_ok >= 200


def test_refine_relative_pose_same_as_reference(ref):
    """The same refined pose as the reference's Levenberg-Marquardt from the same start, on
    inliers with noise and on a set with outliers.  With outliers the reweighted problem
    makes the iteration creep (steps accepted and rejected in turn up to the 50-iteration
    limit), where the reference's Rodrigues map drops rotation steps of at most 1e-6 rad;
    the library's refinement does the same, so even there the end points agree to
    rounding."""
    rng = np.random.default_rng(13)
    for n_out in (0, 150):
        R = OR.rotation([0.05, -0.2, 0.03])
        T = np.array([1.0, 0.1, 0.2]) / np.linalg.norm([1.0, 0.1, 0.2])
        corrs, _ = OR.synthetic_two_view(rng, 600, n_out, K, R, T, 0.4)
        R0 = R @ OR.rotation([0.01, -0.008, 0.012])
        T0 = T + np.array([0.03, -0.02, 0.01])
        T0 /= np.linalg.norm(T0)
        Rm, Tm = ransac.refine_relative_pose(corrs, K, K, R0, T0, 2.0)
        Rr, Tr = _f64(R0).ravel().copy(), _f64(T0).copy()
        c = _f64(corrs).ravel()
        assert ref.ref_refine_relative_pose(_p(c), len(corrs), _p(_f64(K).ravel()), _p(_f64(K).ravel()), _p(Rr),
                                            _p(Tr), 2.0) == 1
        Rr = Rr.reshape(3, 3)
        rot, tdir = OR.rotation_error_deg(Rm, Rr), OR.direction_error_deg(Tm, Tr)
        Kinv = np.linalg.inv(K)
        cost = [np.minimum(OR.sampson_error(Kinv.T @ OR.essential_from_pose(Rx, Tx) @ Kinv, corrs), 4.0).sum()
                for Rx, Tx in ((Rm, Tm), (Rr, Tr))]
        print(f"outliers {n_out}: library vs reference refinement: rotation {rot:.2e} deg, direction {tdir:.2e} deg, "
              f"cost {cost[0]:.9g} vs {cost[1]:.9g}")
        assert np.abs(Rm - Rr).max() < 1e-12 and np.abs(Tm - Tr).max() < 1e-12
        assert abs(cost[0] - cost[1]) <= 1e-12 * cost[1]


@pytest.mark.parametrize("refine", [False, True])
def test_robust_orientation_loop_same_as_reference(ref, refine):
    """The library's RANSAC loop against computeRobustOrientationMLE restated step by step
    (oracle.ransac.robust_orientation_mle) on the reference's own five-point solver, pose
    decomposition and refinement, with the same shuffles: same samples drawn, same best
    hypothesis (E, F, R, T to rounding), same error and the same inlier set, on a scene of
    inliers with noise and uniform outliers.  The hypothesis counts may differ by the few
    samples whose solution counts differ."""
    rng = np.random.default_rng(21)
    R = OR.rotation([0.05, -0.2, 0.03])
    T = np.array([1.0, 0.1, 0.2]) / np.linalg.norm([1.0, 0.1, 0.2])
    corrs, _ = OR.synthetic_two_view(rng, 900, 500, K, R, T, 0.5)

    def five(x1, y1, x2, y2):
        return ref_five(ref, x1, y1, x2, y2)

    def pose(E, x1, y1, x2, y2):
        return ref_pose(ref, E, x1, y1, x2, y2)

    def refine_pose(c, K1, K2, R0, T0, thr):
        Rr, Tr = _f64(R0).ravel().copy(), _f64(T0).copy()
        assert ref.ref_refine_relative_pose(_p(_f64(c).ravel()), len(c), _p(_f64(K1).ravel()), _p(_f64(K2).ravel()),
                                            _p(Rr), _p(Tr), thr) == 1
        return Rr.reshape(3, 3), Tr

    for seed in (3, 8):
        want = OR.robust_orientation_mle(corrs, K, K, 2.0, 500, refine, seed, five, pose, refine_pose)
        got = ransac.robust_orientation(corrs, K, K, 2.0, 500, True, refine, seed=seed)
        print(f"refine={refine} seed={seed}: samples {got['samples']} / {want['samples']}, hypotheses "
              f"{got['hypotheses']} / {want['hypotheses']}, inliers {len(got['inliers'])} / {len(want['inliers'])}, "
              f"error {got['error']:.12g} / {want['error']:.12g}")
        assert got["samples"] == want["samples"]
        assert abs(got["hypotheses"] - want["hypotheses"]) <= 0.02 * want["hypotheses"]
        assert np.array_equal(got["inliers"], want["inliers"])
        assert abs(got["error"] - want["error"]) <= 1e-9 * want["error"]
        for k in ("rotation", "translation"):
            assert np.abs(got[k] - want[k]).max() < 1e-9, k
        for k in ("essential", "fundamental"):
            a, b = got[k] / np.linalg.norm(got[k]), want[k] / np.linalg.norm(want[k])
            assert min(np.abs(a - b).max(), np.abs(a + b).max()) < 1e-9, k
