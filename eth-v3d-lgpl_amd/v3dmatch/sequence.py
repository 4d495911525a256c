"""Batched putative matching of an image sequence (the caller loop of
V3D/Apps/SfM/match_sequence.cpp:165-240) over the batch API.

The reference matches each view against the views that follow it within 2 * radius - 1
(match_sequence.cpp:165, :187-189), one setLeftFeatures / setRightFeatures /
findPutativeMatches per pair, with the feature count min(maxFeatures, size & 0xffe0)
(match_sequence.cpp:183, :219).  Here the pairs of a window of views are matched in one
batch call; per pair the result is what the reference's loop gets for that pair.  The
same driver exists as a C++ program (v3dmatch/_lib/v3dm_match_sequence, source
eth-v3d-lgpl_amd/apps/match_sequence.cpp).
"""
import numpy as np

from . import features as feat


def caller_count(size, max_features, exact=False):
    """Features the reference's caller passes for an image of `size` features:
    min(maxFeatures, size & 0xffe0) -- the mask wraps at 65536 and above; exact=True gives
    size rounded down to a multiple of 32 instead."""
    rule = (int(size) & ~31) if exact else (int(size) & 0xFFE0)
    return min(int(max_features), rule)


def sequence_pairs(n_views, radius):
    """(cur, other) in the reference's loop order."""
    out = []
    for cur in range(n_views):
        for other in range(cur + 1, cur + 2 * radius):
            if other >= n_views:
                continue
            out.append((cur, other))
    return out


def match_sequence(matcher, views, radius, max_features=7168, min_score=0.9, min_ratio=0.8, window=256,
                   exact_count=False):
    """Match an image sequence with one batch call per window of `window` views.

    views: list of descriptor arrays (n_v, 128) float32 or paths of .features archives.
    Returns {(cur, other): (m, 2) int32 matches}, in the reference's pair order."""
    n_views = len(views)
    cache = {}

    def desc(v):
        if v not in cache:
            x = views[v]
            if isinstance(x, (str, bytes)) or hasattr(x, "__fspath__"):
                size, _ = feat.count(x)
                x = feat.read(x, cap=caller_count(size, max_features, exact_count)).desc
            else:
                x = np.asarray(x, np.float32).reshape(-1, 128)
                x = x[:caller_count(x.shape[0], max_features, exact_count)]
            cache[v] = np.ascontiguousarray(x, np.float32)
        return cache[v]

    result = {}
    for w0 in range(0, n_views, window):
        w1 = min(n_views, w0 + window)
        last = min(n_views, w1 + 2 * radius - 1)
        pairs = [(a, b) for (a, b) in sequence_pairs(n_views, radius) if w0 <= a < w1]
        if not pairs:
            continue
        imgs = [desc(v) for v in range(w0, last)]
        matcher.batch_set_images([d.shape[0] for d in imgs], np.concatenate(imgs) if imgs else
                                 np.zeros((0, 128), np.float32))
        matcher.batch_set_pairs([a - w0 for a, _ in pairs], [b - w0 for _, b in pairs])
        per_pair = matcher.batch_match(min_score, min_ratio)
        for (a, b), m in zip(pairs, per_pair):
            result[(a, b)] = m
        for v in range(w0, w1):
            cache.pop(v, None)
    return result
