"""The C++ drop-in class, driven like the reference's match_sequence caller, returns the
oracle's match set on the GPU."""
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
