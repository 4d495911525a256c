"""The sampled check's fallback rule (v3dm_fallback_allowed_misses, host code of the C ABI):
a learned result falls back to the plain form when its sampled misses exceed the number a
recall of minRecall explains at the 5% level.  Pinned against scipy's binomial
distribution, which shares no code with the library's recurrence."""
import numpy as np
import pytest
from scipy.stats import binom

from v3dmatch import SIFT_Matching

ALPHA = 0.05


@pytest.mark.parametrize("recall", [0.5, 0.9, 0.98, 0.995, 0.999, 0.9999])
def test_allowed_misses_is_the_binomial_quantile(recall):
    p_miss = 1.0 - float(np.float32(recall))  # the library receives a float
    for a in range(0, 129):
        k = SIFT_Matching.fallback_allowed_misses(a, recall)
        assert 0 <= k <= a
        if a == 0:
            assert k == 0
            continue
        # k is the smallest count whose upper tail P[X > k] is at most alpha
        assert binom.sf(k, a, p_miss) <= ALPHA + 1e-12, (a, k)
        if k > 0:
            assert binom.sf(k - 1, a, p_miss) > ALPHA - 1e-12, (a, k)


def test_allowed_misses_edges():
    for a in (1, 7, 60, 128):
        assert SIFT_Matching.fallback_allowed_misses(a, 1.0) == 0  # any miss falls back
        assert SIFT_Matching.fallback_allowed_misses(a, 0.0) == a  # never
    # examples: at 0.995 one miss in 60 accepted samples is within the target, two are not;
    # 128 accepted samples tolerate two
    assert SIFT_Matching.fallback_allowed_misses(60, 0.995) == 1
    assert SIFT_Matching.fallback_allowed_misses(128, 0.995) == 2
    assert SIFT_Matching.fallback_allowed_misses(10, 0.995) == 0
    # large counts (r^a underflows in plain arithmetic) and invalid arguments
    for a in (10 ** 5, 10 ** 6):
        k = SIFT_Matching.fallback_allowed_misses(a, 0.995)
        p_miss = 1.0 - float(np.float32(0.995))
        assert binom.sf(k, a, p_miss) <= ALPHA + 1e-9 and binom.sf(k - 1, a, p_miss) > ALPHA - 1e-9, (a, k)
    assert SIFT_Matching.fallback_allowed_misses(-1, 0.995) == -1
    assert SIFT_Matching.fallback_allowed_misses(1 << 25, 0.995) == -1
    assert SIFT_Matching.fallback_allowed_misses(10, 1.5) == -1
