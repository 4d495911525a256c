"""GPU tests of caller batching: an image sequence in .features archives matched by the
batched sequence driver (C++ program v3dm_match_sequence and v3dmatch.sequence) gives,
pair by pair, what the reference's caller loop gets (V3D/Apps/SfM/match_sequence.cpp:
165-240: setLeftFeatures / setRightFeatures / findPutativeMatches per pair, counts
min(maxFeatures, size & 0xffe0)).  The per-pair reference here is the single-pair
findPutativeMatches on the same truncated descriptors (itself identical to the oracle),
and for the small views the oracle directly."""
import os
import subprocess
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import features, sequence, synth

pytestmark = pytest.mark.gpu

# sizes: a multiple of 32, ragged ones, one below 32, one above the maximum
SIZES = [1500, 1237, 2048, 31, 1900, 2500, 777, 1333]
MAXF = 2200
RADIUS = 2


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _views(seed=5):
    """uint8 views; view v reuses (noisy, requantised) rows of view v - 1 so that nearby
    views share matches."""
    rng = np.random.default_rng(seed)
    out = []
    for v, n in enumerate(SIZES):
        q = synth.fresh(synth.seed(777, 9, v, 0), n)
        if v > 0 and out[-1].shape[0] > 0:
            prev = synth.dequantise(out[-1])
            k = min(n, prev.shape[0]) * 2 // 5
            src = rng.choice(prev.shape[0], k, replace=False)
            dst = rng.choice(n, k, replace=False)
            x = np.maximum(prev[src] + rng.normal(0, 0.04, (k, 128)).astype(np.float32), 0)
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            q[dst] = np.minimum(255, np.floor(512 * x + 0.5)).astype(np.uint8)
        out.append(q)
    return out


@pytest.fixture(scope="module")
def sequence_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("seq")
    names = []
    for v, q in enumerate(_views()):
        name = str(d / f"view{v:03d}")
        desc = synth.dequantise(q)
        xy = np.stack([np.arange(len(q)), np.arange(len(q))[::-1]], 1).astype(np.float32)
        features.write(name + ".features", desc, xy=xy, binary=(v % 3 != 2))  # every third in text
        names.append(name)
    lst = d / "list.txt"
    lst.write_text("\n".join(names) + "\n")
    return d, names, lst


def _single_pair_reference(names, min_score=0.9, min_ratio=0.8, exact_count=False):
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(MAXF, MAXF)
    ref = {}
    try:
        descs = []
        for n in names:
            size, _ = features.count(n + ".features")
            descs.append(features.read(n + ".features", cap=sequence.caller_count(size, MAXF, exact_count)).desc)
        for a, b in sequence.sequence_pairs(len(names), RADIUS):
            m.setLeftFeatures(descs[a].shape[0], descs[a])
            m.setRightFeatures(descs[b].shape[0], descs[b])
            ref[(a, b)] = m.find_putative_matches_array(min_score, min_ratio)
    finally:
        m.close()
    return ref, descs


def _parse(path):
    out = {}
    tok = open(path).read().split()
    k = 0
    while k < len(tok):
        a, b, n = int(tok[k]), int(tok[k + 1]), int(tok[k + 2])
        k += 3
        out[(a, b)] = np.array(tok[k:k + 2 * n], np.int32).reshape(n, 2)
        k += 2 * n
    return out


def test_sequence_program_matches_the_caller_loop(sequence_dir):
    d, names, lst = sequence_dir
    say(f"case sequence program: {len(names)} views, radius {RADIUS}, max features {MAXF}")
    ref, descs = _single_pair_reference(names)
    out = d / "matches.txt"
    exe = v3dmatch.lib_path("v3dm_match_sequence")
    r = subprocess.run([exe, str(lst), str(RADIUS), "--max-features", str(MAXF), "--out", str(out), "--window", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    say(r.stdout.strip().splitlines()[-1])
    got = _parse(out)
    assert list(got) == sequence.sequence_pairs(len(names), RADIUS)
    for key, want in ref.items():
        assert np.array_equal(got[key], want), key
    # the caller's log lines (match_sequence.cpp:201-202, :224)
    assert "matching view 0 with view 1" in r.stdout and "nFeatures1 = 1472, nFeatures2 = 1216" in r.stdout
    # the oracle directly on the small views
    for a, b in [(0, 1), (2, 3), (6, 7)]:
        if descs[a].shape[0] and descs[b].shape[0]:
            want = oracle.find_putative_matches(descs[a], descs[b], 0.9, 0.8)
        else:
            want = np.zeros((0, 2), np.int32)
        assert np.array_equal(got[(a, b)], want), (a, b)
    say(f"  pairs={len(got)} matches={sum(len(v) for v in got.values())}")


def test_sequence_python_driver(sequence_dir):
    d, names, lst = sequence_dir
    say("case sequence python driver (paths and arrays, window 2)")
    ref, descs = _single_pair_reference(names)
    m = v3dmatch.SIFT_Matching(0)
    try:
        got = sequence.match_sequence(m, [n + ".features" for n in names], RADIUS, max_features=MAXF, window=2)
        assert list(got) == sequence.sequence_pairs(len(names), RADIUS)
        for key, want in ref.items():
            assert np.array_equal(got[key], want), key
        full = [features.read(n + ".features").desc for n in names]
        got2 = sequence.match_sequence(m, full, RADIUS, max_features=MAXF, window=100)
        for key, want in ref.items():
            assert np.array_equal(got2[key], want), key
    finally:
        m.close()


def test_sequence_program_bf16_equals_batch_bf16(sequence_dir):
    d, names, lst = sequence_dir
    say("case sequence program bf16")
    out = d / "matches_bf16.txt"
    exe = v3dmatch.lib_path("v3dm_match_sequence")
    r = subprocess.run([exe, str(lst), str(RADIUS), "--max-features", str(MAXF), "--out", str(out), "--precision",
                        "bf16", "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _parse(out)
    m = v3dmatch.SIFT_Matching(0)
    try:
        m.set_precision(v3dmatch.PREC_BF16)
        want = sequence.match_sequence(m, [n + ".features" for n in names], RADIUS, max_features=MAXF)
    finally:
        m.close()
    for key in want:
        assert np.array_equal(got[key], want[key]), key
    ref, _ = _single_pair_reference(names)
    exact = sum(len(v) for v in ref.values())
    inter = sum(len(set(map(tuple, got[k].tolist())) & set(map(tuple, ref[k].tolist()))) for k in ref)
    say(f"  bf16 recall {inter / max(exact, 1):.5f} over {exact} exact matches")
    assert inter >= 0.995 * exact


def test_sequence_program_errors(sequence_dir, tmp_path):
    exe = v3dmatch.lib_path("v3dm_match_sequence")
    lst = tmp_path / "missing.txt"
    lst.write_text(str(tmp_path / "nothere") + "\n" + str(tmp_path / "nothere2") + "\n")
    r = subprocess.run([exe, str(lst), "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
