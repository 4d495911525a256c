"""Pins of the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors for its matcher (no test
directory; Data/Example of README.md:78-85 is not shipped), so the oracle is pinned by
checks that do not retype its code:
  * scores against math.fsum (exactly rounded FP64 sums) and numpy, on small inputs;
  * the selection rule of cu:133-210 against a numpy formulation built from textbook
    routines (nanargmax for the first maximal index, a sort for the multiset second best);
  * hand-built score matrices for ties, NaN, the 0.001 clamp, thresholds, -1e30;
  * the mutual check of cu:546-551 by brute force over all pairs;
  * invariants: S(L, R) = S(R, L)^T, unit self-scores;
  * planted mistakes: variants of the oracle built with -DV3DO_PLANT_* must fail these
    checks.
"""
import math
import os
import tempfile

import numpy as np
import pytest

from oracle import build as obuild
from oracle import oracle
from v3dmatch import synth


# ------------------------------------------------------------------ independent formulation
def np_decide(best, nxt, arg, min_score, min_ratio):
    """Decision of cu:150-161 evaluated in numpy float32 / float64 scalars."""
    best, nxt = np.float32(best), np.float32(nxt)
    if min_ratio < 0:
        return int(arg) if best > np.float32(min_score) else -1
    nd = np.float32(np.float32(2.0) - np.float32(2.0) * nxt)
    next_dist = np.float32(max(0.001, float(nd)))
    best_dist = np.float32(np.float32(2.0) - np.float32(2.0) * best)
    ratio = np.float32(best_dist / next_dist)
    return int(arg) if (best > np.float32(min_score) and ratio < np.float32(min_ratio)) else -1


def np_top2(v):
    """(best, second, first index of best) of a vector with the reference's conventions:
    NaN entries are ignored, the state starts at -1e30f (values <= -1e30 never become the
    best), the second best is the second largest of the multiset."""
    v = np.asarray(v, dtype=np.float32)
    floor = np.float32(-1e30)
    vals = np.concatenate([v[~np.isnan(v)], [floor, floor]])
    vals = np.maximum(vals, floor)
    srt = np.sort(vals)[::-1]
    best, second = srt[0], srt[1]
    if best > floor:
        w = np.where(v == best)[0]
        arg = int(w[0])
    else:
        arg = -1
    return np.float32(best), np.float32(second), arg


def np_maxima(S, min_score, min_ratio):
    """colMax (per column = left) and rowMax (per row = right) of a score matrix S[nR, nL]."""
    nR, nL = S.shape
    col = np.array([np_decide(*np_top2(S[:, i]), min_score, min_ratio) for i in range(nL)], np.int32)
    row = np.array([np_decide(*np_top2(S[j, :]), min_score, min_ratio) for j in range(nR)], np.int32)
    return col, row


def np_mutual(col, row):
    out = []
    for k in range(len(col)):
        for l in range(len(row)):
            if col[k] == l and row[l] == k:
                out.append((k, l))
    return np.array(out, np.int32).reshape(-1, 2)


def fsum_scores(L, R):
    S = np.empty((R.shape[0], L.shape[0]), np.float32)
    for j in range(R.shape[0]):
        for i in range(L.shape[0]):
            S[j, i] = np.float32(math.fsum(float(a) * float(b) for a, b in zip(L[i], R[j])))
    return S


def small_pair(nL, nR, seed):
    qL, qR, _ = synth.make_pair(nL, nR, 0.5, config=0, pair=seed)
    return synth.dequantise(qL), synth.dequantise(qR)


# ------------------------------------------------------------------ scores
def test_scores_match_exactly_rounded_sums():
    L, R = small_pair(24, 20, 1)
    S = oracle.scores(L, R)
    F = fsum_scores(L, R)
    # the oracle rounds an FP64 sum, math.fsum rounds the exact sum: equal except for
    # double-rounding cases, which would differ by one ulp at most
    diff = np.abs(S.view(np.int32).astype(np.int64) - F.view(np.int32).astype(np.int64))
    assert diff.max() <= 1
    assert (diff == 0).mean() >= 0.99


def test_scores_close_to_numpy_and_symmetric():
    L, R = small_pair(96, 64, 2)
    S = oracle.scores(L, R)
    N = (R.astype(np.float64) @ L.astype(np.float64).T).astype(np.float32)
    assert np.max(np.abs(S - N)) <= 2 ** -23
    assert np.array_equal(oracle.scores(R, L), S.T)


def test_unit_self_scores():
    L, _ = small_pair(64, 32, 3)
    s = np.array([oracle.exact_dot(L[i], L[i]) for i in range(64)])
    assert np.all(np.abs(s - 1.0) < 1e-6)


def test_top2_scan_equals_scores_then_selection():
    L, R = small_pair(70, 90, 4)
    S = oracle.scores(L, R)
    b, s, a = oracle.top2_scan(L, R)
    for i in range(70):
        nb, ns, na = np_top2(S[:, i])
        assert (b[i], s[i], a[i]) == (nb, ns, na)
    b, s, a = oracle.top2_scan(R, L)
    for j in range(90):
        nb, ns, na = np_top2(S[j, :])
        assert (b[j], s[j], a[j]) == (nb, ns, na)


# ------------------------------------------------------------------ selection rule on hand-built S
def hand_cases():
    f = np.float32
    cases = []
    # exact tie for best -> first index; second equals best -> ratio 1 -> reject
    cases.append(np.array([[0.5, 0.95, 0.3], [0.95, 0.2, 0.3], [0.1, 0.95, 0.3]], f))
    # near-duplicates above 0.9995: the 0.001 clamp lets a tie pass (cu:154)
    cases.append(np.array([[0.99995, 0.1], [0.99993, 0.2], [0.1, 0.3]], f))
    cases.append(np.array([[0.99998, 0.9998], [0.999979, 0.9998], [0.1, 0.3]], f))  # accepted only via the clamp
    # NaN entries are skipped
    cases.append(np.array([[np.nan, 0.95], [0.96, np.nan], [0.5, 0.2], [np.nan, np.nan]], f))
    # everything below the threshold
    cases.append(np.full((5, 4), 0.5, f))
    # values at and below the -1e30 floor, and -inf
    cases.append(np.array([[-1e30, -2e30], [-np.inf, -1e30], [0.95, -5.0]], f))
    # threshold boundary: exactly minScore is rejected (strict >)
    cases.append(np.array([[0.9, 0.1], [0.1, 0.9000001]], f))
    # single candidate row / column
    cases.append(np.array([[0.97, 0.2, 0.93]], f))
    cases.append(np.array([[0.97], [0.2], [0.93]], f))
    rng = np.random.default_rng(5)
    for _ in range(20):
        S = rng.uniform(0.6, 1.0, size=(rng.integers(1, 9), rng.integers(1, 9))).astype(f)
        S[rng.random(S.shape) < 0.2] = S.flat[0]  # plant ties
        cases.append(S)
    return cases


@pytest.mark.parametrize("min_score,min_ratio", [(0.9, 0.8), (0.5, 0.95), (0.9, -1.0), (0.0, 1.5), (0.95, 0.0)])
def test_selection_rule_hand_cases(min_score, min_ratio):
    for S in hand_cases():
        col, row = np_maxima(S, min_score, min_ratio)
        assert np.array_equal(oracle.column_maxima(S, min_score, min_ratio), col), S
        assert np.array_equal(oracle.row_maxima(S, min_score, min_ratio), row), S


def test_decide_clamp_and_ratio_edges():
    # next >= 0.9995 clamps nextDist at 0.001
    assert oracle.decide(0.9997, 0.9999, 4, 0.9, 0.8) == 4
    assert oracle.decide(0.9995, 0.9999, 4, 0.9, 0.8) == -1
    # ratio computed against 2 - 2*next
    assert oracle.decide(0.95, 0.7, 1, 0.9, 0.8) == 1     # 0.1 / 0.6
    assert oracle.decide(0.95, 0.94, 1, 0.9, 0.8) == -1   # 0.1 / 0.12
    assert oracle.decide(0.95, 0.94, 1, 0.9, -1.0) == 1   # no ratio test
    assert oracle.decide(-1e30, -1e30, -1, 0.9, 0.8) == -1


# ------------------------------------------------------------------ full matcher
@pytest.mark.parametrize("nL,nR,seed", [(32, 32, 6), (96, 64, 7), (64, 96, 8)])
def test_find_putative_matches_brute_force(nL, nR, seed):
    L, R = small_pair(nL, nR, seed)
    S = fsum_scores(L, R)
    for ms, mr in [(0.9, 0.8), (0.8, 0.9), (0.9, -1.0)]:
        col, row = np_maxima(S, ms, mr)
        want = np_mutual(col, row)
        got = oracle.find_putative_matches(L, R, ms, mr)
        assert np.array_equal(got, want)
        gc, gr = oracle.maxima(L, R, ms, mr)
        assert np.array_equal(gc, col) and np.array_equal(gr, row)


def test_mutual_brute_force():
    rng = np.random.default_rng(9)
    for _ in range(50):
        nL, nR = rng.integers(1, 12, size=2)
        col = rng.integers(-1, nR, size=nL).astype(np.int32)
        row = rng.integers(-1, nL, size=nR).astype(np.int32)
        assert np.array_equal(oracle.mutual(col, row), np_mutual(col, row))


def test_rounding_to_multiples_of_32():
    # cu:444-460, cu:493-502
    assert [oracle.round32(n) for n in (0, 31, 32, 33, 95, 96, 1000)] == [0, 0, 32, 32, 64, 96, 992]
    assert oracle.effective_count(1000, 512) == 512
    assert oracle.effective_count(100, 4096) == 96
    assert oracle.effective_count(4096, 4100) == 4096


# ------------------------------------------------------------------ planted mistakes
@pytest.fixture(scope="module")
def planted_dir():
    d = tempfile.mkdtemp(prefix="v3do_planted_")
    yield d


def _checks_pass():
    """Run the selection and score checks above against the currently loaded oracle."""
    try:
        for ms, mr in [(0.9, 0.8), (0.0, 1.5)]:
            for S in hand_cases():
                col, row = np_maxima(S, ms, mr)
                if not np.array_equal(oracle.column_maxima(S, ms, mr), col):
                    return False
                if not np.array_equal(oracle.row_maxima(S, ms, mr), row):
                    return False
        L, R = small_pair(24, 20, 1)
        if not np.array_equal(oracle.scores(L, R), fsum_scores(L, R)):
            return False
        return True
    finally:
        pass


@pytest.mark.parametrize("plant", ["V3DO_PLANT_TIE_LAST", "V3DO_PLANT_NO_CLAMP", "V3DO_PLANT_FP32_SUM"])
def test_planted_mistakes_are_caught(planted_dir, plant):
    path = obuild.build_variant(plant, planted_dir)
    good = os.path.join(os.path.dirname(os.path.abspath(oracle.__file__)), "liboracle.so")
    try:
        oracle.use_library(path)
        assert not _checks_pass(), f"{plant} went unnoticed"
    finally:
        oracle.use_library(good)
    assert _checks_pass()


# ------------------------------------------------------------------ guided masks
def test_guided_epipolar_closed_form():
    # F of a pure horizontal translation: epipolar lines are rows, Sampson distance of
    # (p, q) is (qy - py)^2 / 2
    F = np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]], np.float64)
    lxy = np.array([[10, 5], [3, 8], [7, 20]], np.float32)
    rxy = np.array([[1, 5.5], [2, 8], [3, 30], [4, 21]], np.float32)
    S = np.full((4, 3), 0.95, np.float32)
    M = oracle.mask_epipolar(S, np.float32(1.0), F, lxy, rxy)
    for j in range(4):
        for i in range(3):
            d = (rxy[j, 1] - lxy[i, 1]) ** 2 / 2.0
            assert (M[j, i] == np.float32(-1e20)) == (d > 1.0)


def test_guided_homography_closed_form():
    H = np.array([[2, 0, 1], [0, 2, -1], [0, 0, 1]], np.float64)
    lxy = np.array([[1, 1], [2, 3]], np.float32)
    rxy = np.array([[3, 1], [5, 5], [3.5, 1.5]], np.float32)
    S = np.full((3, 2), 0.95, np.float32)
    M = oracle.mask_homography(S, np.float32(0.5), H, lxy, rxy)
    for j in range(3):
        for i in range(2):
            px, py = 2 * lxy[i, 0] + 1, 2 * lxy[i, 1] - 1
            d = (px - rxy[j, 0]) ** 2 + (py - rxy[j, 1]) ** 2
            assert (M[j, i] == np.float32(-1e20)) == (d > 0.5)


# ------------------------------------------------------------------ vocabulary tree
def _tree_levels(branch, depth):
    starts, s, ln = [], 0, 1
    for _ in range(depth):
        starts.append(s)
        ln *= branch
        s += ln
    return starts, s


def test_voctree_oracle_follows_planted_paths():
    # nodes are distinct scaled basis directions; a key built as the sum of the planted
    # path's directions (weights falling with the level) descends exactly along that path
    branch, depth = 3, 3
    starts, n_nodes = _tree_levels(branch, depth)
    rng = np.random.default_rng(0)
    dims = rng.permutation(128)
    tree = np.zeros((n_nodes, 128), np.float32)
    for x in range(n_nodes):
        tree[x, dims[x % 128]] = 1.0
        tree[x, dims[(x * 7 + 3) % 128]] += 0.01 * (x % 5)
    for leaf in range(branch ** depth):
        path, x = [], leaf
        for d in range(depth):
            path.append(x % branch)
            x //= branch
        path = path[::-1]
        key = np.zeros(128, np.float32)
        node, nxt = 0, 0
        for d, k in enumerate(path):
            node = nxt + k
            key += np.float32(4.0 ** (depth - d)) * tree[node]
            nxt = (node + 1) * branch
        got = oracle.voctree_leaves(tree, branch, depth, key[None], global_ids=False)[0]
        glob = oracle.voctree_leaves(tree, branch, depth, key[None], global_ids=True)[0]
        assert glob == node and got == node - starts[-1], (leaf, path, got, glob)


def test_voctree_oracle_against_float64_descent():
    # well-separated random cases: a float64 descent (children at (x + 1) * branch) gives the
    # same leaves whenever every level's best leads its second by a clear margin
    rng = np.random.default_rng(1)
    for branch, depth in [(10, 3), (4, 5), (30, 2)]:
        starts, n_nodes = _tree_levels(branch, depth)
        tree = rng.standard_normal((n_nodes, 128)).astype(np.float32)
        keys = rng.standard_normal((200, 128)).astype(np.float32)
        got = oracle.voctree_leaves(tree, branch, depth, keys, global_ids=True)
        checked = 0
        for q in range(keys.shape[0]):
            nxt, node, clear = 0, 0, True
            for d in range(depth):
                s = tree[nxt:nxt + branch].astype(np.float64) @ keys[q].astype(np.float64)
                o = np.sort(s)
                clear &= (o[-1] - o[-2]) > 1e-3
                node = nxt + int(np.argmax(s))
                nxt = (node + 1) * branch
            if clear:
                assert got[q] == node
                checked += 1
        assert checked > 150


def test_voctree_oracle_ties_nan_and_floor():
    branch, depth = 4, 2
    starts, n_nodes = _tree_levels(branch, depth)
    tree = np.zeros((n_nodes, 128), np.float32)
    tree[:, 0] = 1.0  # every child identical: ties everywhere -> first child at each level
    key = np.zeros(128, np.float32)
    key[0] = 1.0
    assert oracle.voctree_leaves(tree, branch, depth, key[None], True)[0] == (0 + 1) * branch + 0
    # NaN key: no child ever beats -1e30 -> leafId = next - 1 with next = 0 at every level
    nan = np.full((1, 128), np.nan, np.float32)
    assert oracle.voctree_leaves(tree, branch, depth, nan, True)[0] == -1
    assert oracle.voctree_leaves(tree, branch, depth, nan, False)[0] == -1 - starts[-1]
    # scores at or below the -1e30 floor never win either
    far = np.zeros((1, 128), np.float32)
    far[0, 0] = -1e35
    assert oracle.voctree_leaves(tree, branch, depth, far, True)[0] == -1


def test_voctree_oracle_sums_sequentially_in_fp32():
    # child 0 scores 2^24 + 1 - 2^24: 0 in sequential FP32 (2^24 + 1 rounds to 2^24), 1
    # exactly; child 1 scores 0.5 either way.  The reference's float loop picks child 1.
    branch, depth = 2, 1
    tree = np.zeros((2, 128), np.float32)
    tree[0, :3] = [2.0 ** 24, 1.0, -(2.0 ** 24)]
    tree[1, 0] = 0.5
    key = np.zeros((1, 128), np.float32)
    key[0, :3] = 1.0
    assert oracle.voctree_leaves(tree, branch, depth, key, True)[0] == 1
