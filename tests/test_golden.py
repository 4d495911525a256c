"""Committed golden fixtures (tests/golden, written by tests/golden/make_golden.py):
the oracle, the generator and the dequantisation must reproduce them bit for bit."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle
from v3dmatch import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("n,seed", [(32, 11), (96, 12), (512, 13)])
def test_fixture_pairs(n, seed):
    z = np.load(os.path.join(GOLD, f"pairs_n{n}.npz"))
    qL, qR, truth = synth.make_pair(n, n, 0.4, config=0, pair=seed)
    assert np.array_equal(qL, z["qL"]) and np.array_equal(qR, z["qR"]) and np.array_equal(truth, z["truth"])
    L, R = synth.dequantise(z["qL"]), synth.dequantise(z["qR"])
    assert np.array_equal(oracle.find_putative_matches(L, R, 0.9, 0.8), z["matches"])
    assert np.array_equal(oracle.find_putative_matches(L, R, 0.9, -1.0), z["matches_no_ratio"])
    cm, rm = oracle.maxima(L, R, 0.9, 0.8)
    assert np.array_equal(cm, z["colMax"]) and np.array_equal(rm, z["rowMax"])


def test_single_pair_4096_hashes():
    with open(os.path.join(GOLD, "hashes.json")) as f:
        h = json.load(f)
    qL, qR, _ = synth.make_pair(4096, 4096, 0.4, config=1, pair=0)
    assert hashlib.sha256(qL.tobytes()).hexdigest() == h["sha256_qL"]
    assert hashlib.sha256(qR.tobytes()).hexdigest() == h["sha256_qR"]
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    assert hashlib.sha256(L.tobytes()).hexdigest() == h["sha256_L_float"]
    m = oracle.find_putative_matches(L, R, 0.9, 0.8)
    assert len(m) == h["n_matches"]
    assert hashlib.sha256(np.ascontiguousarray(m, np.int32).tobytes()).hexdigest() == h["sha256_matches_int32"]
