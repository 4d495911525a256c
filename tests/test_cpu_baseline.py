"""The timed CPU baseline (FP32 port of the reference algorithm) agrees with the exact
oracle on small pairs (CPU only)."""
import numpy as np
import pytest

from oracle import oracle
from v3dmatch import synth


@pytest.mark.parametrize("nL,nR,seed", [(32, 32, 1), (300, 257, 2), (1024, 1024, 3), (2048, 1000, 4)])
def test_cpu_baseline_matches_oracle(nL, nR, seed):
    qL, qR, _ = synth.make_pair(nL, nR, 0.4, config=0, pair=seed)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    for ms, mr in [(0.9, 0.8), (0.9, -1.0)]:
        a = oracle.cpu_baseline_find(L, R, ms, mr, 4)
        b = oracle.find_putative_matches(L, R, ms, mr)
        assert np.array_equal(a, b)
