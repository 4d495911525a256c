"""Vocabulary-tree quantisation on the GPU (VocabularyTreeTraversal<BranchFactor, Depth>,
V3D/CUDA/v3d_cudamatching.h:69-105, and VocabularyTree<Depth, BranchFactor>::getVisualWord,
V3D/Base/v3d_voctree.h:19-99) over the C ABI of libv3dmatch.so.

Every key descends the tree level by level to the first child of maximal score, the score
being the sequential FP32 sum of the 128 products (the CPU tree's rule).  `traverseTree`
returns node ids as the CUDA traversal does; `visual_words` / `getVisualWord` return leaf
indices in [0, leafCount) as the CPU tree does.
"""
import ctypes as C

import numpy as np

from ._native import V3DMError, load_library

NONE = np.iinfo(np.int32).min  # a descent that left the node array


def sizes(branch, depth):
    """(getNodeCount, getLeaveCount)."""
    lib = load_library()
    n, l = C.c_int64(0), C.c_int64(0)
    rc = lib.v3dm_voctree_sizes(int(branch), int(depth), C.byref(n), C.byref(l))
    if rc:
        raise V3DMError(rc, lib.v3dm_last_error(None).decode())
    return n.value, l.value


class VocabularyTree:
    """A tree of `depth` levels with `branch` children per node on one GPU."""

    def __init__(self, branch, depth, device=0):
        self._lib = load_library()
        self.branch, self.depth = int(branch), int(depth)
        self.node_count, self.leaf_count = sizes(branch, depth)
        self._h = C.c_void_p()
        self._check(self._lib.v3dm_voctree_create(C.byref(self._h), int(device), self.branch, self.depth))

    def _check(self, rc):
        if rc:
            raise V3DMError(rc, self._lib.v3dm_last_error(None).decode())

    def close(self):
        if self._h:
            self._lib.v3dm_voctree_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # VocabularyTreeTraversal names (cu:795-860)
    def setNodes(self, nodes):
        a = np.ascontiguousarray(nodes, np.float32).reshape(-1, 128)
        if a.shape[0] != self.node_count:
            raise ValueError(f"expected {self.node_count} nodes, got {a.shape[0]}")
        self._check(self._lib.v3dm_voctree_set_nodes(self._h, a.ctypes.data))

    def loadTreeFromFile(self, path):
        self._check(self._lib.v3dm_voctree_load(self._h, str(path).encode()))
        return True

    def _traverse(self, keys, global_ids):
        k = np.ascontiguousarray(keys, np.float32).reshape(-1, 128)
        out = np.empty(k.shape[0], np.int32)
        self._check(self._lib.v3dm_voctree_traverse(self._h, k.shape[0], k.ctypes.data, out.ctypes.data,
                                                     1 if global_ids else 0))
        return out

    def traverseTree(self, nFeatures, features):
        """Node ids of the first nFeatures keys (the CUDA traversal's output)."""
        return self._traverse(np.asarray(features, np.float32).reshape(-1, 128)[:nFeatures], True)

    def visual_words(self, keys):
        """Visual word (leaf index) of every key, as getVisualWord."""
        return self._traverse(keys, False)

    def getVisualWord(self, key):
        return int(self._traverse(np.asarray(key, np.float32).reshape(1, 128), False)[0])

    def traverse_device(self, n, d_keys, d_leaves, global_ids=False):
        """Enqueue the descent of n device-resident keys (device pointers)."""
        self._check(self._lib.v3dm_voctree_traverse_device(self._h, int(n), C.c_void_p(d_keys),
                                                            C.c_void_p(d_leaves), 1 if global_ids else 0))

    def synchronize(self):
        """Wait for the tree's stream; returns the keys the last call decided exactly."""
        n = C.c_int(0)
        self._check(self._lib.v3dm_voctree_synchronize(self._h, C.byref(n)))
        return n.value
