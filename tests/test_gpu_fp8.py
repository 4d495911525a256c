"""GPU tests of the FP8 (e4m3, per-descriptor power-of-two scale) shortlist form: recall
and precision against the exact match set — the oracle's at sizes it finishes in seconds,
the exact FP32 GPU form's (itself identical to the oracle's) at 32768 x 32768."""
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


def _pairs_set(a):
    return set(map(tuple, a.tolist()))


@pytest.fixture(scope="module")
def matcher():
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(32768, 32768)
    yield m
    m.close()


def _check(got, want, tag, st):
    inter = len(got & want)
    recall = inter / max(len(want), 1)
    precision = inter / max(len(got), 1)
    say(f"  {tag}: exact={len(want)} fp8={len(got)} recall={recall:.5f} precision={precision:.5f} "
        f"rescored={st.get('nRescored', 0)}")
    assert recall >= 0.995
    assert precision >= 0.995


@pytest.mark.parametrize("nL,nR,seed", [(512, 480, 1), (1000, 777, 2), (4096, 4096, 0), (4096, 2080, 4),
                                        (160, 4096, 5)])
def test_fp8_recall_oracle(matcher, nL, nR, seed):
    say(f"case fp8 nL={nL} nR={nR} seed={seed}")
    qL, qR, _ = synth.make_pair(nL, nR, 0.4, config=4, pair=seed)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.set_precision(v3dmatch.PREC_FP8)
    matcher.setLeftFeatures(nL, L)
    matcher.setRightFeatures(nR, R)
    got = _pairs_set(matcher.find_putative_matches_array(0.9, 0.8))
    eL, eR = oracle.effective_count(nL, 32768), oracle.effective_count(nR, 32768)
    want = _pairs_set(oracle.find_putative_matches(L[:eL], R[:eR], 0.9, 0.8))
    _check(got, want, "oracle", matcher.stats())


def test_fp8_recall_full_size(matcher):
    n = 32768
    say(f"case fp8 full size {n}x{n}")
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=4, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.setLeftFeatures(n, L)
    matcher.setRightFeatures(n, R)
    matcher.set_precision(v3dmatch.PREC_FP32)
    want = _pairs_set(matcher.find_putative_matches_array(0.9, 0.8))
    matcher.set_precision(v3dmatch.PREC_FP8)
    got = _pairs_set(matcher.find_putative_matches_array(0.9, 0.8))
    _check(got, want, "exact FP32 form", matcher.stats())


def test_fp8_scale_invariance(matcher):
    """Descriptors scaled by powers of two far outside the e4m3 range give the same match
    set: the per-descriptor scale absorbs the exponent."""
    say("case fp8 scale invariance")
    qL, qR, _ = synth.make_pair(2048, 2048, 0.4, config=4, pair=7)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    matcher.set_precision(v3dmatch.PREC_FP8)
    matcher.setLeftFeatures(2048, L)
    matcher.setRightFeatures(2048, R)
    base = _pairs_set(matcher.find_putative_matches_array(0.9, 0.8))
    assert len(base) > 0
    # L * 2^a and R * 2^-a: every product, norm product and score is unchanged exactly
    for a in (20, -30, 12):
        matcher.setLeftFeatures(2048, np.ldexp(L, a).astype(np.float32))
        matcher.setRightFeatures(2048, np.ldexp(R, -a).astype(np.float32))
        got = _pairs_set(matcher.find_putative_matches_array(0.9, 0.8))
        assert got == base, a
