"""GPU parity of the plain FP32 path against the exact CPU oracle.

The plain path must reproduce the oracle's match set exactly (identical (left, right)
pairs in identical order) and the per-feature decisions colMax / rowMax of cu:532-544.
"""
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def matcher():
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(16384, 16384)
    yield m
    m.close()


def _pair(nL, nR, frac=0.4, pair=0, config=1):
    qL, qR, truth = synth.make_pair(nL, nR, frac, config=config, pair=pair)
    return synth.dequantise(qL), synth.dequantise(qR), truth


def _check(m, L, R, nL, nR, min_score=0.9, min_ratio=0.8):
    m.setLeftFeatures(nL, L)
    m.setRightFeatures(nR, R)
    got = m.find_putative_matches_array(min_score, min_ratio)
    eL = oracle.effective_count(nL, 16384)
    eR = oracle.effective_count(nR, 16384)
    want = oracle.find_putative_matches(L[:eL], R[:eR], min_score, min_ratio) if eL and eR else \
        np.zeros((0, 2), np.int32)
    assert got.shape == want.shape, (got.shape, want.shape)
    assert np.array_equal(got, want)
    if eL and eR:
        cm, rm = m.get_maxima()
        ocm, orm = oracle.maxima(L[:eL], R[:eR], min_score, min_ratio)
        assert np.array_equal(cm, ocm)
        assert np.array_equal(rm, orm)
    return got


@pytest.mark.parametrize("nL,nR,seed", [
    (32, 32, 1), (64, 96, 2), (512, 512, 3), (512, 480, 4), (1000, 777, 5), (2048, 4096, 6),
    (4096, 4096, 0), (4096, 4096, 11), (4096, 4096, 12), (4096, 2080, 13), (160, 4096, 14),
])
def test_plain_matches_oracle(matcher, nL, nR, seed):
    say(f"case plain nL={nL} nR={nR} seed={seed}")
    L, R, _ = _pair(nL, nR, pair=seed)
    got = _check(matcher, L, R, nL, nR)
    st = matcher.stats()
    say(f"  matches={len(got)} rechecked={st['nRechecked']}")


@pytest.mark.parametrize("min_score,min_ratio", [(0.9, 0.8), (0.8, 0.9), (0.95, 0.6), (0.9, -1.0), (0.5, -1.0),
                                                  (0.0, 1.5)])
def test_thresholds(matcher, min_score, min_ratio):
    say(f"case thresholds {min_score} {min_ratio}")
    L, R, _ = _pair(1024, 1024, pair=21)
    _check(matcher, L, R, 1024, 1024, min_score, min_ratio)


def test_small_counts_round_to_empty(matcher):
    say("case small counts")
    L, R, _ = _pair(64, 64, pair=3)
    matcher.setLeftFeatures(31, L)
    matcher.setRightFeatures(64, R)
    assert matcher.findPutativeMatches() == []
    matcher.setLeftFeatures(64, L)
    matcher.setRightFeatures(0, R)
    assert matcher.findPutativeMatches() == []


def test_exact_duplicates_and_ties(matcher):
    say("case duplicates")
    L, R, _ = _pair(256, 256, pair=31)
    R = R.copy()
    R[:64] = L[:64]          # exact copies: scores ~1, clamp path of cu:154
    R[64:96] = L[100:132]
    R[96] = R[97] = L[200]   # duplicated right: exact tie in left 200's scan
    _check(matcher, L, R, 256, 256)
    _check(matcher, L, R, 256, 256, 0.9, -1.0)


def test_u8_upload_equals_float(matcher):
    say("case u8 upload")
    qL, qR, _ = synth.make_pair(2048, 1500, 0.4, config=1, pair=41)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.setLeftFeatures(2048, L)
    matcher.setRightFeatures(1500, R)
    a = matcher.find_putative_matches_array()
    matcher.set_left_u8(qL)
    matcher.set_right_u8(qR)
    b = matcher.find_putative_matches_array()
    assert np.array_equal(a, b)


def test_pair_16384_full_size(matcher):
    say("case pair 16384x16384")
    L, R, _ = _pair(16384, 16384, pair=0, config=2)
    got = _check(matcher, L, R, 16384, 16384)
    say(f"  16384x16384 matches={len(got)} rechecked={matcher.stats()['nRechecked']}")


@pytest.mark.parametrize("nL,nR,seed", [(512, 480, 51), (2048, 1024, 52), (4096, 4096, 53)])
def test_every_query_rechecked_exactly(matcher, nL, nR, seed):
    """Certification off: every decision comes from the exact FP64 re-check kernel."""
    say(f"case recheck-all nL={nL} nR={nR}")
    L, R, _ = _pair(nL, nR, pair=seed)
    matcher.set_certification(0.0)
    try:
        _check(matcher, L, R, nL, nR)
        _check(matcher, L, R, nL, nR, 0.9, -1.0)
        assert matcher.stats()["nRechecked"] == nL + nR
    finally:
        matcher.set_certification(2.0 ** -15)


def test_result_larger_than_previous(matcher):
    """The match list is copied in one piece sized from the previous result; a result much
    larger than the previous one (here: 32 features, then 4096 features matched against
    copies of themselves) is completed by a second copy."""
    m = matcher
    L, R, _ = _pair(32, 32, pair=21)
    _check(m, L, R, 32, 32)
    small = m.stats()["nMatches"]
    rng = np.random.default_rng(21)
    q = rng.integers(0, 60, (4096, 128)).astype(np.uint8)
    X = synth.dequantise(q)
    got = _check(m, X, X.copy(), 4096, 4096)
    say(f"  previous result {small} matches, this one {len(got)}")
    assert len(got) > 1024 + small + small // 4  # larger than the first copy
