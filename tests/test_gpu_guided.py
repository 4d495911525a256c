"""GPU parity of guided matching (runGuidedMatching / runGuidedHomographyMatching,
V3D/CUDA/v3d_cudamatching.cu:554-694) against the CPU oracle.

The oracle masks the exact score matrix of the last findPutativeMatches in place
(-1e20 for pairs outside the distance threshold, cu:308 / cu:340) and selects on it;
repeated guided calls compound on the same matrix.  The GPU path must give the identical
match list and identical per-feature decisions.  Positions come from a synthetic two-view
geometry: matched copies are projections of the same 3-D point (plus pixel noise), the
distractors sit anywhere in the image.
"""
import os
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import learned, synth

pytestmark = pytest.mark.gpu

NMAX = 16384


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def matcher():
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(NMAX, NMAX, True)
    yield m
    m.close()


def _skew(t):
    return np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]], np.float64)


def _rodrigues(a):
    th = np.linalg.norm(a)
    k = a / th
    Kx = _skew(k)
    return np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx


def _geometry(nL, nR, truth, seed, noise=0.7):
    """Left / right positions (float32 AoS) and (F, H): F of a general two-view pair,
    H of a plane.  Right copies (truth >= 0) project the left feature's 3-D point."""
    rng = np.random.default_rng(seed)
    K = np.array([[800.0, 0, 512], [0, 800.0, 384], [0, 0, 1]])
    R = _rodrigues(rng.standard_normal(3) * 0.05)
    t = np.array([1.0, 0.1, 0.05]) + rng.standard_normal(3) * 0.02
    # 3-D points at random depths: F holds for every copy, the plane homography H (the
    # fronto-parallel plane at the median depth) only for copies near that depth
    uv = np.stack([rng.uniform(0, 1024, nL), rng.uniform(0, 768, nL)], 1)
    depth = rng.uniform(6, 20, nL)
    X = (np.linalg.inv(K) @ np.c_[uv, np.ones(nL)].T).T * depth[:, None]
    Xr = (R @ X.T).T + t
    pr = (K @ Xr.T).T
    pr = pr[:, :2] / pr[:, 2:3]
    left = uv.astype(np.float32)
    right = np.stack([rng.uniform(0, 1024, nR), rng.uniform(0, 768, nR)], 1)
    cp = truth >= 0
    right[cp] = pr[truth[cp]] + rng.standard_normal((cp.sum(), 2)) * noise
    right = right.astype(np.float32)
    Kinv = np.linalg.inv(K)
    F = Kinv.T @ _skew(t) @ R @ Kinv
    F /= np.linalg.norm(F)
    d = float(np.median(depth))
    H = K @ (R + np.outer(t, [0, 0, 1.0]) / d) @ Kinv
    H /= H[2, 2]
    return left, right, F, H


def _case(nL, nR, seed, frac=0.4):
    qL, qR, truth = synth.make_pair(nL, nR, frac, config=1, pair=seed)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    lxy, rxy, F, H = _geometry(nL, nR, truth, seed)
    return L, R, lxy, rxy, F, H


def _run(m, L, R, lxy, rxy, calls, min_score=0.9, min_ratio=0.8, precision=v3dmatch.PREC_FP32):
    """find, then the guided calls [(kind, M, thr), ...]; returns the GPU result of every
    guided call and the oracle's."""
    nL, nR = L.shape[0], R.shape[0]
    m.setLeftFeatures(nL, L)
    m.setRightFeatures(nR, R)
    m.set_precision(precision)
    m.find_putative_matches_array(min_score, min_ratio)
    S = oracle.scores(L, R)
    out = []
    for kind, M, thr in calls:
        if kind == 0:
            got = m.runGuidedMatching(M, thr, lxy, rxy, min_score, min_ratio)
        else:
            got = m.runGuidedHomographyMatching(M, thr, lxy, rxy, min_score, min_ratio)
        got = np.array(got, np.int32).reshape(-1, 2)
        want = oracle.guided(S, kind, M, np.float32(thr), lxy, rxy, min_score, min_ratio)
        cm, rm = m.get_maxima()
        ocm = oracle.column_maxima(S, min_score, min_ratio)
        orm = oracle.row_maxima(S, min_score, min_ratio)
        out.append((got, want, cm, rm, ocm, orm))
    m.set_precision(v3dmatch.PREC_FP32)
    return out


def _assert_same(res, label):
    for k, (got, want, cm, rm, ocm, orm) in enumerate(res):
        assert got.shape == want.shape, (label, k, got.shape, want.shape)
        assert np.array_equal(got, want), (label, k)
        assert np.array_equal(cm, ocm), (label, k)
        assert np.array_equal(rm, orm), (label, k)


@pytest.mark.parametrize("nL,nR,seed,kind,thr", [
    (1024, 1024, 1, 0, 2.0), (1024, 1024, 1, 1, 4.0), (2048, 1536, 2, 0, 1.0), (992, 1504, 3, 1, 8.0),
    (4096, 4096, 4, 0, 3.0), (4096, 4096, 5, 1, 3.0), (160, 4096, 6, 0, 2.0),
])
def test_guided_matches_oracle(matcher, nL, nR, seed, kind, thr):
    say(f"case guided kind={kind} nL={nL} nR={nR} seed={seed} thr={thr}")
    L, R, lxy, rxy, F, H = _case(nL, nR, seed)
    M = F if kind == 0 else H
    res = _run(matcher, L, R, lxy, rxy, [(kind, M, thr)])
    got, want = res[0][0], res[0][1]
    plain = oracle.find_putative_matches(L, R, 0.9, 0.8)
    say(f"  putative={len(plain)} guided={len(got)} oracle={len(want)}")
    _assert_same(res, "guided")


def test_guided_no_ratio(matcher):
    say("case guided no-ratio nL=nR=2048")
    L, R, lxy, rxy, F, H = _case(2048, 2048, 7)
    _assert_same(_run(matcher, L, R, lxy, rxy, [(0, F, 2.0)], 0.8, -1.0), "no-ratio")


def test_guided_masks_compound(matcher):
    # F then H then F again with a tighter threshold: every call masks the same stored
    # scores further (cu:592-605 writes into the score matrix in place)
    say("case guided compound nL=nR=4096")
    L, R, lxy, rxy, F, H = _case(4096, 4096, 8)
    res = _run(matcher, L, R, lxy, rxy, [(0, F, 4.0), (1, H, 30.0), (0, F, 1.0)])
    say(f"  matches per call: {[len(r[0]) for r in res]}")
    _assert_same(res, "compound")
    # a new findPutativeMatches starts from unmasked scores again
    m = matcher
    got = m.find_putative_matches_array(0.9, 0.8)
    assert np.array_equal(got, oracle.find_putative_matches(L, R, 0.9, 0.8))


def test_guided_edge_positions(matcher):
    # NaN positions (NaN distance keeps the pair), points mapped to infinity by H (third
    # component 0), an all-removing threshold and a threshold that removes nothing
    say("case guided edge positions")
    L, R, lxy, rxy, F, H = _case(1024, 1024, 9)
    lxy = lxy.copy()
    rxy = rxy.copy()
    lxy[3] = [np.nan, 5.0]
    rxy[7] = [np.nan, np.nan]
    Hz = np.array([[1.0, 0, 0], [0, 1.0, 0], [1.0, 0, -5.0]])
    lxy[10] = [5.0, 0.0]   # Hp = (5, 0, 0): 5/0 = inf, 0/0 = NaN
    lxy[11] = [5.0, 2.0]   # (inf, inf)
    res = _run(matcher, L, R, lxy, rxy, [(1, Hz, 1e6)])
    _assert_same(res, "H with z = 0")
    res = _run(matcher, L, R, lxy, rxy, [(0, F, 0.0)])
    _assert_same(res, "threshold 0")
    res = _run(matcher, L, R, lxy, rxy, [(0, F, float("inf"))])
    _assert_same(res, "threshold inf")
    assert np.array_equal(res[0][0], oracle.find_putative_matches(L, R, 0.9, 0.8))


def test_guided_full_size_properties(matcher):
    # 16384 x 16384: a threshold that removes nothing reproduces the putative set exactly;
    # a real threshold keeps only geometrically consistent pairs, every one of which
    # passes the oracle's closed-form distance check, and the list stays sorted
    say("case guided full size 16384x16384")
    L, R, lxy, rxy, F, H = _case(16384, 16384, 10)
    m = matcher
    m.setLeftFeatures(16384, L)
    m.setRightFeatures(16384, R)
    put = m.find_putative_matches_array(0.9, 0.8)
    same = np.array(m.runGuidedMatching(F, float("inf"), lxy, rxy), np.int32).reshape(-1, 2)
    assert np.array_equal(same, put)
    m.find_putative_matches_array(0.9, 0.8)
    g = np.array(m.runGuidedMatching(F, 2.0, lxy, rxy), np.int32).reshape(-1, 2)
    say(f"  putative={len(put)} guided={len(g)}")
    assert len(g) > 0 and np.all(np.diff(g[:, 0]) > 0)
    # every guided match is within the threshold (the oracle's mask on a 1-entry matrix)
    for i, j in g[:: max(1, len(g) // 200)]:
        s = oracle.mask_epipolar(np.zeros((1, 1), np.float32), np.float32(4.0), F, lxy[i:i + 1], rxy[j:j + 1])
        assert s[0, 0] == 0.0


def test_guided_after_learned(matcher):
    # a guided call after a learned-shortlist find scores exactly (FP32) and supersedes the
    # learned result and its pending sampled check
    assert os.path.exists(learned.WEIGHTS_PATH)
    say("case guided after learned find")
    m = matcher
    learned.Network(m).load(learned.WEIGHTS_PATH)
    L, R, lxy, rxy, F, H = _case(4096, 4096, 11)
    res = _run(m, L, R, lxy, rxy, [(1, H, 3.0)], precision=v3dmatch.PREC_LEARNED)
    _assert_same(res, "after learned")
    st = m.stats()
    assert st["precision"] == v3dmatch.PREC_FP32 and st["nFallback"] == 0


def test_guided_requires_allocation_flag():
    m = v3dmatch.SIFT_Matching(0)
    try:
        m.enableRoundingWarning(False)
        m.allocate(256, 256)
        L, R, lxy, rxy, F, H = _case(256, 256, 12)
        m.setLeftFeatures(256, L)
        m.setRightFeatures(256, R)
        m.findPutativeMatches()
        with pytest.raises(v3dmatch.V3DMError):
            m.runGuidedMatching(F, 2.0, lxy, rxy)
    finally:
        m.close()
