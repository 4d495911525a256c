"""GPU tests of the learned-code shortlist form: with the trained network
(v3dmatch/weights/shortlist_mlp.npz) the match set keeps >= 99.5% recall and precision
against the exact set; the same pipeline with an untrained network loses recall (what the
training gains); and the per-query candidate count trades recall for work."""
import os
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
    assert os.path.exists(learned.WEIGHTS_PATH), "trained weights missing"
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(16384, 16384)
    net = learned.Network(m).load(learned.WEIGHTS_PATH)
    yield m, net
    m.close()


def _recall(m, L, R, want, samples=0):
    """Recall / precision of the learned form itself: the sampled check (and with it the
    fallback to the plain form, which would hide a defect of the learned path) is off
    unless samples > 0."""
    m.setLeftFeatures(L.shape[0], L)
    m.setRightFeatures(R.shape[0], R)
    m.set_precision(v3dmatch.PREC_LEARNED)
    m.set_fallback(samples, 0.995)
    got = _pairs_set(m.find_putative_matches_array(0.9, 0.8))
    st = m.stats()
    m.set_fallback(128, 0.995)
    assert samples > 0 or st["nFallback"] == 0
    inter = len(got & want)
    return inter / max(len(want), 1), inter / max(len(got), 1), st


@pytest.mark.parametrize("n,seed", [(1024, 3), (4096, 0), (4096, 7)])
def test_learned_recall_oracle(matcher, n, seed):
    m, net = matcher
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=seed)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = _pairs_set(oracle.find_putative_matches(L, R, 0.9, 0.8))
    rec, prec, st = _recall(m, L, R, want)
    say(f"case learned n={n} seed={seed}: exact={len(want)} recall={rec:.5f} precision={prec:.5f} "
        f"rescored={st['nRescored']}")
    assert rec >= 0.995 and prec >= 0.995


def test_learned_recall_full_size(matcher):
    m, net = matcher
    n = 16384
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=2, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m.setLeftFeatures(n, L)
    m.setRightFeatures(n, R)
    m.set_precision(v3dmatch.PREC_FP32)
    want = _pairs_set(m.find_putative_matches_array(0.9, 0.8))
    rec, prec, st = _recall(m, L, R, want)
    say(f"case learned full size: exact={len(want)} recall={rec:.5f} precision={prec:.5f}")
    assert rec >= 0.995 and prec >= 0.995


def test_learned_large_candidate_sets():
    """Many queries and candidates: the default plan would put more than 1024 candidate tiles
    in one split, past what a packed list index holds, so the plan takes more splits.  The
    learned form's result against the plain form's (the exact set) on a 262144 x 131072 pair."""
    nL, nR = 262144, 131072
    say(f"case learned {nL}x{nR}")
    qL, qR, _ = synth.make_pair(nL, nR, 0.4, config=1, pair=17)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(nL, nR)
    learned.Network(m).load(learned.WEIGHTS_PATH)
    m.setLeftFeatures(nL, L)
    m.setRightFeatures(nR, R)
    m.set_precision(v3dmatch.PREC_FP32)
    want = _pairs_set(m.find_putative_matches_array(0.9, 0.8))
    rec, prec, st = _recall(m, L, R, want)
    m.close()
    say(f"  exact={len(want)} recall={rec:.5f} precision={prec:.5f}")
    assert rec >= 0.99 and prec >= 0.995


def test_training_gain(matcher):
    """The trained network against the same pipeline with its initial (untrained) weights."""
    m, net = matcher
    n = 4096
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=11)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = _pairs_set(oracle.find_putative_matches(L, R, 0.9, 0.8))
    rec_t, _, _ = _recall(m, L, R, want)
    other = v3dmatch.SIFT_Matching(0)
    other.enableRoundingWarning(False)
    other.allocate(n, n)
    raw = learned.Network(other)
    raw.set_params(learned.init_params(1))
    raw.calibrate_code_scale(np.concatenate([L[:2048], R[:2048]]))
    rec_u, _, _ = _recall(other, L, R, want)
    other.close()
    say(f"case training gain: trained recall {rec_t:.5f}, untrained recall {rec_u:.5f}")
    assert rec_t >= 0.995
    assert rec_u < rec_t - 0.01


def test_topk_tradeoff(matcher):
    m, net = matcher
    n = 4096
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=12)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = _pairs_set(oracle.find_putative_matches(L, R, 0.9, 0.8))
    res = {}
    for k in (2, 4, 8, 16):
        m.set_learned_topk(k)
        res[k] = _recall(m, L, R, want)
    m.set_learned_topk(6)
    say("case top-k: " + " ".join(f"k={k}: recall {r[0]:.5f} rescored {r[2]['nRescored']}" for k, r in res.items()))
    assert res[16][0] >= res[2][0]
    assert res[16][2]["nRescored"] > res[2][2]["nRescored"]


def test_margin_tradeoff(matcher):
    """The code-score margin prunes the rescored candidates (fewer with a smaller margin)
    and, at its default (0.07 scale^2), keeps recall and precision >= 99.5%."""
    m, net = matcher
    n = 4096
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=14)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = _pairs_set(oracle.find_putative_matches(L, R, 0.9, 0.8))
    res = {}
    for margin in (-1.0, 0.07, 0.02):
        m.set_learned_margin(margin)
        res[margin] = _recall(m, L, R, want)
    m.set_learned_margin(0.07)
    say("case margin: " + " ".join(f"margin={k}: recall {r[0]:.5f} precision {r[1]:.5f} rescored {r[2]['nRescored']}"
                                   for k, r in res.items()))
    assert res[-1.0][2]["nRescored"] > res[0.07][2]["nRescored"] > res[0.02][2]["nRescored"]
    assert res[0.07][0] >= 0.995 and res[0.07][1] >= 0.995


def test_learned_batch(matcher):
    """The learned form through the batch API on an uneven batch: recall and precision
    against the plain batch (identical to the single-pair exact path)."""
    m, net = matcher
    counts, desc, left, right, _ = synth.make_batch(24, lo=300, hi=6000, config=41)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    m.set_precision(v3dmatch.PREC_FP32)
    want = m.batch_match(0.9, 0.8)
    m.set_precision(v3dmatch.PREC_LEARNED)
    m.set_fallback(0, 0.995)  # the learned batch itself, without plain reruns
    got = m.batch_match(0.9, 0.8)
    m.set_fallback(128, 0.995)
    ws = set((p, a, b) for p in range(len(left)) for a, b in want[p].tolist())
    gs = set((p, a, b) for p in range(len(left)) for a, b in got[p].tolist())
    inter = len(ws & gs)
    say(f"case learned batch: exact={len(ws)} learned={len(gs)} recall={inter / len(ws):.5f} "
        f"precision={inter / max(len(gs), 1):.5f} tuning {m.batch_get_tuning(v3dmatch.PREC_LEARNED)}")
    assert inter >= 0.995 * len(ws) and inter >= 0.995 * len(gs)
    m.set_precision(v3dmatch.PREC_FP32)


def _oracle_pairs(images, left, right):
    out = []
    for a, b in zip(left, right):
        L, R = synth.dequantise(images[a]), synth.dequantise(images[b])
        eL, eR = oracle.round32(L.shape[0]), oracle.round32(R.shape[0])
        out.append(np.zeros((0, 2), np.int32) if eL == 0 or eR == 0
                   else np.asarray(oracle.find_putative_matches(L[:eL], R[:eR], 0.9, 0.8), np.int32).reshape(-1, 2))
    return out


def test_learned_batch_ragged_images(matcher):
    """Every batch form on ragged images: an image without features, images below the
    32-feature granularity (rounded to none) or just above it, an image paired with itself,
    one used by several pairs and an empty image paired with itself.  Pairs without features
    on a side return no matches and have zero check counts; the plain form returns the
    oracle's set pair by pair; the BF16 form and the learned form (check off and on) keep
    >= 99.5% recall and precision against it."""
    m, net = matcher
    counts0, desc0, left0, right0, _ = synth.make_batch(10, lo=40, hi=1500, config=51)
    x1 = synth.make_pair(20, 700, 0.4, config=52, pair=0)
    x2 = synth.make_pair(33, 64, 0.5, config=53, pair=1)
    n = len(counts0)
    E, T, B, C, D = n, n + 1, n + 2, n + 3, n + 4
    counts = np.concatenate([counts0, [0, 20, 700, 33, 64]]).astype(np.int32)
    desc = np.concatenate([desc0, x1[0], x1[1], x2[0], x2[1]])
    left = np.concatenate([left0, [E, 0, T, B, C, D, B, 0, E]]).astype(np.int32)
    right = np.concatenate([right0, [0, E, B, T, D, C, B, 2, E]]).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(counts)])
    images = [desc[offs[k]:offs[k + 1]] for k in range(len(counts))]
    want = _oracle_pairs(images, left, right)
    empty = [p for p in range(len(left)) if oracle.round32(int(counts[left[p]])) == 0
             or oracle.round32(int(counts[right[p]])) == 0]
    say(f"case ragged batch: {len(left)} pairs, {len(empty)} without features on a side")
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    m.set_precision(v3dmatch.PREC_FP32)
    got = m.batch_match(0.9, 0.8)
    for p in range(len(left)):
        assert np.array_equal(got[p], want[p]), p
    ws = set((p, a, b) for p in range(len(left)) for a, b in want[p].tolist())
    for mode, samples in ((v3dmatch.PREC_BF16, 0), (v3dmatch.PREC_LEARNED, 0), (v3dmatch.PREC_LEARNED, 128)):
        m.set_precision(mode)
        m.set_fallback(samples, 0.995)
        got = m.batch_match(0.9, 0.8)
        for p in empty:
            assert len(got[p]) == 0, (mode, p)
        gs = set((p, a, b) for p in range(len(left)) for a, b in got[p].tolist())
        inter = len(ws & gs)
        say(f"  mode {mode} samples {samples}: recall {inter / len(ws):.5f} precision {inter / max(len(gs), 1):.5f}")
        assert inter >= 0.995 * len(ws) and inter >= 0.995 * len(gs)
        if mode == v3dmatch.PREC_LEARNED and samples:
            chk = m.batch_get_check()
            for p in empty:
                assert not chk[p].any(), (p, chk[p])
            assert chk[[p for p in range(len(left)) if p not in empty], 3].min() > 0
    m.set_fallback(128, 0.995)
    m.set_precision(v3dmatch.PREC_FP32)


def test_batch_graph_replays_equal_eager(matcher):
    """Every form's batch run replayed from its captured hipGraph several times, with eager
    work of the same handle (a single-pair learned call) and synchronisations between the
    replays, equals the run launched kernel by kernel: match sets, and for the learned form
    (sampled check and plain fallback on) the per-pair check counts and fallback flags."""
    m, net = matcher
    counts, desc, left, right, _ = synth.make_batch(24, lo=300, hi=6000, config=47)
    qL, qR, _ = synth.make_pair(2048, 2048, 0.4, config=1, pair=21)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    m.set_fallback(128, 0.995)
    for mode in (v3dmatch.PREC_FP32, v3dmatch.PREC_BF16, v3dmatch.PREC_LEARNED):
        m.set_precision(mode)
        m.batch_use_graph(False)
        eager = m.batch_match(0.9, 0.8)
        chk = m.batch_get_check() if mode == v3dmatch.PREC_LEARNED else None
        flags = m.batch_get_fallback() if mode == v3dmatch.PREC_LEARNED else None
        m.batch_use_graph(True)
        for rep in range(3):
            got = m.batch_match(0.9, 0.8)
            say(f"case graph replay: mode {mode} replay {rep}: {sum(len(g) for g in got)} matches")
            assert all(np.array_equal(a, b) for a, b in zip(got, eager)), (mode, rep)
            if chk is not None:
                assert np.array_equal(m.batch_get_check(), chk) and np.array_equal(m.batch_get_fallback(), flags)
            # eager work of the same handle between replays
            m.setLeftFeatures(2048, L)
            m.setRightFeatures(2048, R)
            m.find_putative_matches_array(0.9, 0.8)
            m.synchronize()
    m.set_precision(v3dmatch.PREC_FP32)


def test_learned_batch_tile_shapes(matcher):
    """Every work-group shape of the learned code pass (128 or 256 queries per 64-candidate
    tile) lists the same candidates: identical match sets, with the check on; the shape the
    tuner chose is reported."""
    m, net = matcher
    counts, desc, left, right, _ = synth.make_batch(24, lo=300, hi=6000, config=43)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    m.set_precision(v3dmatch.PREC_LEARNED)
    res = {}
    for q in (256, 128):
        m.batch_set_tile_shape(v3dmatch.PREC_LEARNED, q)
        res[q] = m.batch_match(0.9, 0.8)
        shape = m.batch_get_tile_shape(v3dmatch.PREC_LEARNED)
        say(f"case learned tile shape {q}: {shape}, {sum(len(r) for r in res[q])} matches")
        assert shape == {"query_rows": q, "candidate_rows": 64}
    assert all(np.array_equal(a, b) for a, b in zip(res[256], res[128]))
    m.batch_set_tile_shape(v3dmatch.PREC_LEARNED, 0)
    m.batch_set_tuning(autotune=True)
    tuned = m.batch_match(0.9, 0.8)
    say(f"case learned tile shape tuned: {m.batch_get_tile_shape(v3dmatch.PREC_LEARNED)} "
        f"{m.batch_get_tuning(v3dmatch.PREC_LEARNED)}")
    assert m.batch_get_tile_shape(v3dmatch.PREC_LEARNED)["query_rows"] in (128, 256)
    assert all(np.array_equal(a, b) for a, b in zip(res[256], tuned))
    with pytest.raises(v3dmatch.V3DMError):
        m.batch_set_tile_shape(v3dmatch.PREC_LEARNED, 512)
    m.set_precision(v3dmatch.PREC_FP32)


def test_sampled_check_and_fallback(matcher):
    """Every learned call decides 128 sampled left features exactly; with the trained
    network the check passes (no fallback), with a poor shortlist (untrained network, one
    candidate per query) the sampled recall drops and the result is the plain form's."""
    m, net = matcher
    n = 4096
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=13)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = oracle.find_putative_matches(L, R, 0.9, 0.8)
    m.setLeftFeatures(n, L)
    m.setRightFeatures(n, R)
    m.set_precision(v3dmatch.PREC_LEARNED)
    got = m.find_putative_matches_array(0.9, 0.8)
    st = m.stats()
    say(f"case sampled check (trained): sampled={st['nSampled']} missed={st['nSampleMissed']} "
        f"fallback={st['nFallback']}")
    assert st["nSampled"] == 128 and st["nFallback"] == 0
    other = v3dmatch.SIFT_Matching(0)
    other.enableRoundingWarning(False)
    other.allocate(n, n)
    raw = learned.Network(other)
    raw.set_params(learned.init_params(2))
    raw.calibrate_code_scale(np.concatenate([L[:2048], R[:2048]]))
    other.setLeftFeatures(n, L)
    other.setRightFeatures(n, R)
    other.set_precision(v3dmatch.PREC_LEARNED)
    other.set_learned_topk(1)
    other.set_fallback(128, 1.0)
    got = other.find_putative_matches_array(0.9, 0.8)
    st = other.stats()
    say(f"case sampled check (untrained, top-1): sampled={st['nSampled']} missed={st['nSampleMissed']} "
        f"fallback={st['nFallback']}")
    assert st["nSampleMissed"] > 0 and st["nFallback"] == 1
    assert np.array_equal(got, want)  # the plain form's (exact) result
    other.set_fallback(0, 0.995)
    got = other.find_putative_matches_array(0.9, 0.8)
    st = other.stats()
    assert st["nSampled"] == 0 and st["nFallback"] == 0
    other.close()


def test_sampled_check_follows_features(matcher):
    """The sampled check reads the features on its own stream: after every change of the
    features its counts are those of the new features.  Expected counts: the oracle's exact
    decisions of the 128 evenly spaced left features against the learned decisions (read
    back with the fallback threshold at 0, so that no fallback replaces them)."""
    m, net = matcher
    n = 2048
    sets = []
    for seed in (31, 32, 33):
        qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=seed)
        L, R = synth.dequantise(qL), synth.dequantise(qR)
        sets.append((L, R, oracle.maxima(L, R, 0.9, 0.8)[0]))
    other = v3dmatch.SIFT_Matching(0)
    other.enableRoundingWarning(False)
    other.allocate(n, n)
    raw = learned.Network(other)
    raw.set_params(learned.init_params(5))
    raw.calibrate_code_scale(np.concatenate([sets[0][0][:1024], sets[0][1][:1024]]))
    other.set_precision(v3dmatch.PREC_LEARNED)
    other.set_learned_topk(1)  # a weak shortlist: misses to count
    other.set_fallback(128, 0.0)
    rows = (np.arange(128, dtype=np.int64) * n) // 128
    total_missed = 0
    for step, k in enumerate((0, 1, 0, 2, 2, 1)):
        L, R, exact = sets[k]
        if step == 0 or k != (0, 1, 0, 2, 2, 1)[step - 1]:
            other.setLeftFeatures(n, L)
            other.setRightFeatures(n, R)
        other.find_async(0.9, 0.8)
        learned_cm, _ = other.get_maxima()
        st = other.stats()
        x, d = exact[rows], learned_cm[rows]
        want_missed = int(np.sum((x >= 0) & (d != x)))
        say(f"case check after feature change: step {step} set {k}: sampled={st['nSampled']} "
            f"missed={st['nSampleMissed']} (expected {want_missed})")
        assert st["nSampled"] == 128 and st["nFallback"] == 0
        assert st["nSampleMissed"] == want_missed
        total_missed += want_missed
    assert total_missed > 0
    other.close()


def test_learned_batch_fallback(matcher):
    """Batch: with the trained network no pair falls back; with an untrained network and a
    single candidate per query, the pairs whose sampled recall falls short get exactly the
    plain form's matches."""
    m, net = matcher
    counts, desc, left, right, _ = synth.make_batch(12, lo=600, hi=5000, config=43)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    m.set_precision(v3dmatch.PREC_FP32)
    plain = m.batch_match(0.9, 0.8)
    m.set_precision(v3dmatch.PREC_LEARNED)
    m.batch_match(0.9, 0.8)
    st = m.stats()
    say(f"case batch check (trained): sampled={st['nSampled']} missed={st['nSampleMissed']} "
        f"fallback pairs={st['nFallback']}")
    assert st["nSampled"] > 0 and st["nFallback"] == 0
    other = v3dmatch.SIFT_Matching(0)
    raw = learned.Network(other)
    raw.set_params(learned.init_params(3))
    raw.calibrate_code_scale(synth.dequantise(desc[:4096]))
    other.batch_set_images(counts, desc)
    other.batch_set_pairs(left, right)
    other.set_precision(v3dmatch.PREC_LEARNED)
    other.set_learned_topk(1)
    other.set_fallback(128, 1.0)
    got = other.batch_match(0.9, 0.8)
    flags = other.batch_get_fallback()
    st = other.stats()
    say(f"case batch check (untrained, top-1): missed={st['nSampleMissed']} fallback pairs={int(flags.sum())}")
    assert flags.sum() > 0 and st["nFallback"] == int(flags.sum())
    for p in range(len(left)):
        if flags[p]:
            assert np.array_equal(got[p], plain[p]), p
    other.close()


def test_time_dominant_kernel(matcher):
    """The measurement helper times each mode's dominant kernel and leaves the last result
    readable."""
    m, net = matcher
    n = 4096
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=1, pair=13)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m.setLeftFeatures(n, L)
    m.setRightFeatures(n, R)
    for prec in (v3dmatch.PREC_FP32, v3dmatch.PREC_BF16, v3dmatch.PREC_FP8, v3dmatch.PREC_LEARNED):
        m.set_precision(prec)
        want = m.find_putative_matches_array(0.9, 0.8)
        ms = m.time_dominant_kernel(5)
        say(f"case time dominant kernel: mode {prec}: {ms * 1e3:.1f} us per launch")
        assert 0.0 < ms < 50.0
        assert np.array_equal(m.get_matches(), want)
    m.set_precision(v3dmatch.PREC_FP32)
