"""The sampled check's fallback rules (host code of the C ABI).

Threshold rule (the default, v3dm_fallback_allowed_misses): a learned result falls back to
the plain form when its sampled recall (a - m) / a is below minRecall; pinned by exact
rational arithmetic on the float the library receives.  Binomial rule
(v3dm_fallback_allowed_misses_binomial): it falls back when its sampled misses exceed the
number a recall of minRecall explains at the 5% level; pinned against scipy's binomial
distribution, which shares no code with the library's recurrence."""
from fractions import Fraction

import numpy as np
import pytest
from scipy.stats import binom

from v3dmatch import SIFT_Matching

ALPHA = 0.05
THRESHOLD, BINOMIAL = 0, 1


@pytest.mark.parametrize("recall", [0.5, 0.9, 0.98, 0.995, 0.999, 0.9999])
def test_threshold_rule_is_the_recall_bound(recall):
    r = Fraction(float(np.float32(recall)))  # the library receives a float
    for a in range(0, 300):
        k = SIFT_Matching.fallback_allowed_misses(a, recall, THRESHOLD)
        assert 0 <= k <= a
        if a == 0:
            assert k == 0
            continue
        # k is the largest miss count whose recall is not below the float minRecall (the
        # float quotient of two small integers agrees with the exact one on this grid)
        assert Fraction(a - k, a) >= r or Fraction(float(np.float32(a - k) / np.float32(a))) >= r, (a, k)
        if k < a:
            assert Fraction(float(np.float32(a - k - 1) / np.float32(a))) < r, (a, k)


def test_threshold_rule_examples():
    # 0.995: a miss among fewer than 200 exact matches falls back; 200 tolerate one
    assert SIFT_Matching.fallback_allowed_misses(128, 0.995) == 0
    assert SIFT_Matching.fallback_allowed_misses(199, 0.995) == 0
    assert SIFT_Matching.fallback_allowed_misses(200, 0.995) == 1
    assert SIFT_Matching.fallback_allowed_misses(399, 0.995) == 1
    assert SIFT_Matching.fallback_allowed_misses(400, 0.995) == 2
    assert SIFT_Matching.fallback_allowed_misses(10, 0.9) == 1
    for a in (1, 7, 60, 128):
        assert SIFT_Matching.fallback_allowed_misses(a, 1.0) == 0  # any miss falls back
        assert SIFT_Matching.fallback_allowed_misses(a, 0.0) == a  # never
    assert SIFT_Matching.fallback_allowed_misses(-1, 0.995) == -1
    assert SIFT_Matching.fallback_allowed_misses(1 << 25, 0.995) == -1
    assert SIFT_Matching.fallback_allowed_misses(10, 1.5) == -1


@pytest.mark.parametrize("recall", [0.5, 0.9, 0.98, 0.995, 0.999, 0.9999])
def test_binomial_rule_is_the_binomial_quantile(recall):
    p_miss = 1.0 - float(np.float32(recall))  # the library receives a float
    for a in range(0, 129):
        k = SIFT_Matching.fallback_allowed_misses(a, recall, BINOMIAL)
        assert 0 <= k <= a
        if a == 0:
            assert k == 0
            continue
        # k is the smallest count whose upper tail P[X > k] is at most alpha
        assert binom.sf(k, a, p_miss) <= ALPHA + 1e-12, (a, k)
        if k > 0:
            assert binom.sf(k - 1, a, p_miss) > ALPHA - 1e-12, (a, k)


def test_binomial_rule_edges():
    for a in (1, 7, 60, 128):
        assert SIFT_Matching.fallback_allowed_misses(a, 1.0, BINOMIAL) == 0
        assert SIFT_Matching.fallback_allowed_misses(a, 0.0, BINOMIAL) == a
    # at 0.995 one miss in 60 samples is within the test, two are not; 128 tolerate two
    assert SIFT_Matching.fallback_allowed_misses(60, 0.995, BINOMIAL) == 1
    assert SIFT_Matching.fallback_allowed_misses(128, 0.995, BINOMIAL) == 2
    assert SIFT_Matching.fallback_allowed_misses(10, 0.995, BINOMIAL) == 0
    # large counts (r^a underflows in plain arithmetic) and invalid arguments
    for a in (10 ** 5, 10 ** 6):
        k = SIFT_Matching.fallback_allowed_misses(a, 0.995, BINOMIAL)
        p_miss = 1.0 - float(np.float32(0.995))
        assert binom.sf(k, a, p_miss) <= ALPHA + 1e-9 and binom.sf(k - 1, a, p_miss) > ALPHA - 1e-9, (a, k)
    assert SIFT_Matching.fallback_allowed_misses(-1, 0.995, BINOMIAL) == -1
    assert SIFT_Matching.fallback_allowed_misses(1 << 25, 0.995, BINOMIAL) == -1
    assert SIFT_Matching.fallback_allowed_misses(10, 1.5, BINOMIAL) == -1
