"""The C++ drop-in classes, driven like the reference's callers, return the oracle's
results on the GPU: SIFT_Matching (match_sequence.cpp) and VocabularyTreeTraversal."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle
from v3dmatch import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_dropin_cpp_caller(tmp_path):
    sys.path.insert(0, os.path.join(ROOT, "tests", "cpp"))
    from build_dropin import build
    exe = build(str(tmp_path))
    print("case dropin 3000x2500", flush=True)
    r = subprocess.run([exe, "3000", "2500", "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    n = int(lines[0])
    got = np.array([list(map(int, l.split())) for l in lines[1:]], np.int32).reshape(-1, 2)
    assert len(got) == n
    qL, qR, _ = synth.make_pair(3000, 2500, 0.4, config=0, pair=3)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = oracle.find_putative_matches(L[:2976], R[:2496], 0.9, 0.8)
    assert np.array_equal(got, want)


def test_dropin_cpp_voctree(tmp_path):
    sys.path.insert(0, os.path.join(ROOT, "tests", "cpp"))
    from build_dropin import build
    exe = build(str(tmp_path), "dropin_voctree")
    print("case dropin voctree <10,3>", flush=True)
    rng = np.random.default_rng(5)
    n_nodes = 10 + 100 + 1000
    tree = rng.standard_normal((n_nodes, 128)).astype(np.float32)
    keys = rng.standard_normal((3000, 128)).astype(np.float32)
    keys[7] = tree[3]  # a key equal to a node
    tree.tofile(tmp_path / "tree.bin")
    keys.tofile(tmp_path / "keys.bin")
    r = subprocess.run([exe, str(tmp_path / "tree.bin"), str(tmp_path / "keys.bin"), "3000"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.array(r.stdout.split(), np.int32)
    # as the reference: allocate(3000) rounds the capacity down to 2976 with a warning and
    # traverseTree reduces nFeatures to it (cu:806-811, :843-848); the rest stays untouched
    assert "should be a multiple of 32" in r.stderr and "Reducing nFeatures" in r.stderr
    assert np.array_equal(got[:2976], oracle.voctree_leaves(tree, 10, 3, keys[:2976], global_ids=True))
    assert np.all(got[2976:] == 0)
