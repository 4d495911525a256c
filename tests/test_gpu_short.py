"""GPU tests of the BF16 shortlist form: recall and precision against the exact match
set of the oracle (the reduced-precision forms are measured, not required to be
identical)."""
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import learned, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pairs_set(a):
    return set(map(tuple, a.tolist()))


@pytest.fixture(scope="module")
def matcher():
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(16384, 16384)
    m.set_precision(v3dmatch.PREC_BF16)
    yield m
    m.close()


@pytest.mark.parametrize("nL,nR,seed,config", [
    (512, 480, 1, 1), (1000, 777, 2, 1), (4096, 4096, 0, 1), (4096, 4096, 3, 1), (4096, 2080, 4, 1),
    (160, 4096, 5, 1), (16384, 16384, 0, 2),
])
def test_bf16_recall(matcher, nL, nR, seed, config):
    say(f"case bf16 nL={nL} nR={nR} seed={seed}")
    qL, qR, _ = synth.make_pair(nL, nR, 0.4, config=config, pair=seed)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.setLeftFeatures(nL, L)
    matcher.setRightFeatures(nR, R)
    got = _pairs_set(matcher.find_putative_matches_array(0.9, 0.8))
    eL, eR = oracle.effective_count(nL, 16384), oracle.effective_count(nR, 16384)
    want = _pairs_set(oracle.find_putative_matches(L[:eL], R[:eR], 0.9, 0.8))
    inter = len(got & want)
    recall = inter / max(len(want), 1)
    precision = inter / max(len(got), 1)
    st = matcher.stats()
    say(f"  exact={len(want)} bf16={len(got)} recall={recall:.5f} precision={precision:.5f} "
        f"rescored={st.get('nRescored', 0)}")
    assert recall >= 0.995
    assert precision >= 0.995


def test_bf16_no_ratio_and_edges(matcher):
    say("case bf16 no-ratio")
    qL, qR, _ = synth.make_pair(2048, 2048, 0.4, config=1, pair=9)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.setLeftFeatures(2048, L)
    matcher.setRightFeatures(2048, R)
    got = _pairs_set(matcher.find_putative_matches_array(0.9, -1.0))
    want = _pairs_set(oracle.find_putative_matches(L, R, 0.9, -1.0))
    assert len(got & want) / max(len(want), 1) >= 0.995
    matcher.setLeftFeatures(31, L)
    assert matcher.findPutativeMatches() == []


@pytest.mark.parametrize("prec", [v3dmatch.PREC_BF16, v3dmatch.PREC_FP8, v3dmatch.PREC_LEARNED])
def test_repeated_calls_are_identical(matcher, prec):
    """The shortlist passes stream candidate tiles through double-buffered LDS stages; a
    stage read before its loads have landed would change the lists from call to call.  Five
    calls on the 16384^2 pair give identical decisions in both directions."""
    say(f"case repeated calls mode {prec}")
    qL, qR, _ = synth.make_pair(16384, 16384, 0.4, config=2, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.setLeftFeatures(16384, L)
    matcher.setRightFeatures(16384, R)
    if prec == v3dmatch.PREC_LEARNED:
        learned.Network(matcher).load(learned.WEIGHTS_PATH)
    matcher.set_precision(prec)
    matcher.set_fallback(0, 0.995)
    first = None
    for _ in range(5):
        matcher.find_async(0.9, 0.8)
        cm, rm = matcher.get_maxima()
        if first is None:
            first = (cm.copy(), rm.copy())
        else:
            assert np.array_equal(cm, first[0]) and np.array_equal(rm, first[1])
    matcher.set_fallback(128, 0.995)
    matcher.set_precision(v3dmatch.PREC_BF16)
