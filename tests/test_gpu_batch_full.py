"""GPU tests on the full uneven batch: 256 image pairs, 500-16000 features per image
(log-uniform), 30-60% copies, uint8 descriptors.

* plain FP32 form: every pair's match set identical to the oracle's
  (V3D/CUDA/v3d_cudamatching.cu:522-552 restated on the CPU);
* learned form (trained network) with its sampled check and per-pair fallback: recall and
  precision >= 99.5% against the oracle's set, and every pair the check sent to the plain
  form has exactly the oracle's matches;
* the sampled check's counts per pair equal those recomputed on the host from the exact
  and the learned match sets (samples on the larger image, left on a tie);
* certified integer rescoring of the shortlists = exact FP64 rescoring of the same
  candidates (identical match sets, learned and BF16).
"""
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import learned, synth

pytestmark = pytest.mark.gpu

P = 256


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch():
    counts, desc, left, right, _ = synth.make_batch(P)
    offs = np.concatenate([[0], np.cumsum(counts)])
    images = [desc[offs[m]:offs[m + 1]] for m in range(len(counts))]
    say(f"case full batch: {P} pairs, {int(counts.sum())} features; oracle on the host")
    want = []
    for a, b in zip(left, right):
        L, R = synth.dequantise(images[a]), synth.dequantise(images[b])
        eL, eR = oracle.round32(L.shape[0]), oracle.round32(R.shape[0])
        want.append(oracle.find_putative_matches(L[:eL], R[:eR], 0.9, 0.8))
    m = v3dmatch.SIFT_Matching(0)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    learned.Network(m).load(learned.WEIGHTS_PATH)
    yield m, counts, left, right, want
    m.close()


def _sets(res):
    return [set(map(tuple, r.tolist())) for r in res]


def _recall(want, got):
    ws, gs = _sets(want), _sets(got)
    inter = sum(len(ws[p] & gs[p]) for p in range(P))
    per = [len(ws[p] & gs[p]) / max(len(ws[p]), 1) for p in range(P)]
    return inter / sum(len(w) for w in ws), inter / max(sum(len(g) for g in gs), 1), per


def _expected_check(counts, left, right, exact, got, samples=128):
    """Per pair (exact matches among the samples, missed by `got`): the samples are the
    min(samples, nA) evenly spaced features k * nA // S of the pair's larger image A (left
    on a tie); sample i is an exact match (i, j) of the exact set, missed when `got` lacks
    the same pair."""
    rc = counts - (counts & 31)
    ex, gs = _sets(exact), _sets(got)
    out = np.zeros((P, 2), np.int64)
    for p in range(P):
        nL, nR = int(rc[left[p]]), int(rc[right[p]])
        if nL == 0 or nR == 0:
            continue
        side = 0 if nL >= nR else 1
        nA = nL if side == 0 else nR
        S = min(samples, nA)
        by = {(a if side == 0 else b): (a, b) for a, b in ex[p]}
        for k in range(S):
            i = (k * nA) // S
            if i in by:
                out[p, 0] += 1
                out[p, 1] += by[i] not in gs[p]
    return out


def test_plain_full_batch_equals_oracle(batch):
    m, counts, left, right, want = batch
    m.set_precision(v3dmatch.PREC_FP32)
    got = m.batch_match(0.9, 0.8)
    bad = [p for p in range(P) if not np.array_equal(got[p], want[p])]
    say(f"case full batch plain: {sum(len(w) for w in want)} matches, pairs differing from the oracle: {bad}")
    assert not bad


def test_learned_full_batch_check_counts(batch):
    """The check's counts, with the fallback off (minRecall 0), against the host's count
    from the oracle's set and the learned sets."""
    m, counts, left, right, want = batch
    m.set_precision(v3dmatch.PREC_LEARNED)
    m.set_fallback(128, 0.0)
    got = m.batch_match(0.9, 0.8)
    chk = m.batch_get_check()
    exp = _expected_check(counts, left, right, want, got)
    rec, prec, per = _recall(want, got)
    say(f"case full batch learned, no fallback: recall={rec:.5f} precision={prec:.5f} "
        f"min pair recall={min(per):.4f} pairs below 0.995={sum(r < 0.995 for r in per)}; "
        f"sampled exact matches={int(chk[:, 0].sum())} missed={int(chk[:, 1].sum())} "
        f"decided exactly={int(chk[:, 2].sum())}")
    m.set_fallback(128, 0.995)
    assert np.array_equal(chk[:, 0], exp[:, 0])
    assert np.array_equal(chk[:, 1], exp[:, 1])
    assert not m.batch_get_fallback().any()
    assert rec >= 0.995 and prec >= 0.995


def test_learned_full_batch_with_fallback(batch):
    m, counts, left, right, want = batch
    m.set_precision(v3dmatch.PREC_LEARNED)
    m.set_fallback(128, 0.995)
    got = m.batch_match(0.9, 0.8)
    flags = m.batch_get_fallback()
    chk = m.batch_get_check()
    rec, prec, per = _recall(want, got)
    say(f"case full batch learned with fallback: recall={rec:.5f} precision={prec:.5f} "
        f"min pair recall={min(per):.4f} fallback pairs={int(flags.sum())}")
    for p in range(P):
        a, mi = int(chk[p, 0]), int(chk[p, 1])
        assert bool(flags[p]) == (mi > v3dmatch.SIFT_Matching.fallback_allowed_misses(a, 0.995)), p
        if flags[p]:
            assert np.array_equal(got[p], want[p]), p
    assert rec >= 0.995 and prec >= 0.995


@pytest.mark.parametrize("prec", [v3dmatch.PREC_LEARNED, v3dmatch.PREC_BF16])
def test_certified_rescoring_equals_exact(batch, prec):
    m, counts, left, right, want = batch
    m.set_precision(prec)
    m.set_fallback(0, 0.995)
    cert = m.batch_match(0.9, 0.8)
    m.set_certification(0.0)
    exact = m.batch_match(0.9, 0.8)
    m.set_certification(2.0 ** -15)
    m.set_fallback(128, 0.995)
    bad = [p for p in range(P) if not np.array_equal(cert[p], exact[p])]
    say(f"case certified rescoring mode {prec}: {sum(len(c) for c in cert)} matches, pairs differing: {bad}")
    assert not bad
