"""GPU parity of vocabulary-tree quantisation (VocabularyTree::findBestLeafIndex,
V3D/Base/v3d_voctree.h:63-99; VocabularyTreeTraversal::traverseTree,
V3D/CUDA/v3d_cudamatching.cu:705-869) against the CPU oracle: identical leaves for every
key, visual words and node ids, on random and hierarchical trees, planted exact and
near ties, NaN / zero / far keys, ragged key counts, the tree file loader and the
device-pointer entry point."""
import sys

import numpy as np
import pytest

from oracle import oracle
from v3dmatch import voctree

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _unit(x):
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def _random_tree(branch, depth, seed):
    n_nodes, _ = voctree.sizes(branch, depth)
    return _unit(np.random.default_rng(seed).standard_normal((n_nodes, 128)))


def _hier_tree(branch, depth, seed, spread=0.35):
    """k-means-like tree: children are perturbations of their parent's direction."""
    rng = np.random.default_rng(seed)
    n_nodes, _ = voctree.sizes(branch, depth)
    tree = np.zeros((n_nodes, 128), np.float32)
    tree[:branch] = _unit(np.abs(rng.standard_normal((branch, 128))))
    start, ln = 0, branch
    for d in range(1, depth):
        nxt = start + ln
        for p in range(start, start + ln):
            c0 = (p + 1) * branch
            tree[c0:c0 + branch] = _unit(tree[p] + spread * np.abs(rng.standard_normal((branch, 128))) / np.sqrt(128))
        start, ln = nxt, ln * branch
    return tree


def _keys_near_leaves(tree, branch, depth, n, seed, noise=0.05):
    rng = np.random.default_rng(seed)
    n_nodes, n_leaves = voctree.sizes(branch, depth)
    leaves = tree[n_nodes - n_leaves:]
    pick = rng.integers(0, n_leaves, n)
    return _unit(np.abs(leaves[pick] + noise * rng.standard_normal((n, 128))))


def _check(t, tree, branch, depth, keys, label):
    got_w = t.visual_words(keys)
    got_g = t.traverseTree(keys.shape[0], keys)
    want_w = oracle.voctree_leaves(tree, branch, depth, keys, global_ids=False)
    want_g = oracle.voctree_leaves(tree, branch, depth, keys, global_ids=True)
    bad = np.flatnonzero(got_w != want_w)
    assert bad.size == 0, (label, bad[:10], got_w[bad[:10]], want_w[bad[:10]])
    assert np.array_equal(got_g, want_g), label
    return t.synchronize()


@pytest.mark.parametrize("branch,depth,kind,n", [
    (10, 3, "random", 20000), (50, 3, "hier", 60000), (8, 4, "hier", 30001), (300, 2, "random", 9999),
    (10, 5, "hier", 40000), (1, 3, "random", 100), (64, 2, "hier", 4096),
])
def test_voctree_matches_oracle(branch, depth, kind, n):
    say(f"case voctree <{branch},{depth}> {kind} n={n}")
    tree = _random_tree(branch, depth, 3) if kind == "random" else _hier_tree(branch, depth, 3)
    keys = _keys_near_leaves(tree, branch, depth, n, 4) if kind == "hier" else \
        _unit(np.random.default_rng(4).standard_normal((n, 128)))
    with voctree.VocabularyTree(branch, depth) as t:
        t.setNodes(tree)
        exact = _check(t, tree, branch, depth, keys, kind)
        say(f"  keys decided exactly: {exact}")


def test_voctree_ties_and_special_keys():
    branch, depth = 16, 3
    say("case voctree ties / NaN / zero / far keys")
    tree = _hier_tree(branch, depth, 7)
    # exact ties: children 5 and 9 of the root equal child 2; near ties: child 11 one ulp
    # away from child 3 in one component
    tree[5] = tree[2]
    tree[9] = tree[2]
    tree[11] = tree[3]
    tree[11, 17] = np.nextafter(tree[3, 17], np.float32(2))
    n_nodes, _ = voctree.sizes(branch, depth)
    keys = _keys_near_leaves(tree, branch, depth, 3000, 8)
    keys[:64] = np.repeat(tree[2:3], 64, 0)        # ties at the first level
    keys[64:128] = np.repeat(tree[3:4], 64, 0)     # one-ulp near ties
    keys[128] = np.nan                             # NaN: no child beats -1e30
    keys[129] = 0.0                                # all scores 0: first child each level
    keys[130] = -1e35 * tree[0]                    # far below the -1e30 floor
    keys[131, :] = np.inf
    with voctree.VocabularyTree(branch, depth) as t:
        t.setNodes(tree)
        exact = _check(t, tree, branch, depth, keys, "special")
        say(f"  keys decided exactly: {exact}")
        assert exact >= 128
        for n in (0, 1, 31, 33):
            _check(t, tree, branch, depth, keys[:n], f"n={n}")


class _Hip:
    """hipMalloc / hipMemcpy / hipFree from the HIP runtime the library is linked with."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]

    def malloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.rt.hipMalloc(self.C.byref(p), nbytes) == 0
        return p.value

    def h2d(self, dst, a):
        assert self.rt.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.rt.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def free(self, p):
        self.rt.hipFree(p)


def test_voctree_tree_file_and_device_entry(tmp_path):
    branch, depth = 20, 3
    say("case voctree tree file + device pointers")
    tree = _hier_tree(branch, depth, 11)
    path = tmp_path / "tree.bin"
    tree.tofile(path)
    keys = _keys_near_leaves(tree, branch, depth, 50000, 12)
    want = oracle.voctree_leaves(tree, branch, depth, keys)
    with voctree.VocabularyTree(branch, depth) as t:
        assert t.loadTreeFromFile(path)
        assert np.array_equal(t.visual_words(keys), want)
        hip = _Hip()
        dk, dl = hip.malloc(keys.nbytes), hip.malloc(4 * keys.shape[0])
        try:
            hip.h2d(dk, keys)
            t.traverse_device(keys.shape[0], dk, dl)
            t.synchronize()
            got = np.empty(keys.shape[0], np.int32)
            hip.d2h(got, dl)
        finally:
            hip.free(dk)
            hip.free(dl)
        assert np.array_equal(got, want)
    # a short file leaves the remaining nodes zero, as the reference's loader does
    short = tmp_path / "short.bin"
    tree[:100].tofile(short)
    padded = np.zeros_like(tree)
    padded[:100] = tree[:100]
    with voctree.VocabularyTree(branch, depth) as t:
        t.loadTreeFromFile(short)
        assert np.array_equal(t.visual_words(keys[:2000]), oracle.voctree_leaves(padded, branch, depth, keys[:2000]))
        with pytest.raises(Exception):
            t.loadTreeFromFile(tmp_path / "missing.bin")


def test_voctree_large_batch():
    branch, depth = 50, 3
    n = 200000
    say(f"case voctree <{branch},{depth}> n={n}")
    tree = _hier_tree(branch, depth, 21)
    keys = _keys_near_leaves(tree, branch, depth, n, 22, noise=0.08)
    with voctree.VocabularyTree(branch, depth) as t:
        t.setNodes(tree)
        exact = _check(t, tree, branch, depth, keys, "large")
        say(f"  keys decided exactly: {exact} of {n}")
