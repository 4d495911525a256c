// Vocabulary-tree quantisation on gfx950 (interface between v3dm_voctree.cpp and
// v3dm_voctree.hip).
//
// Reference: VocabularyTree<Depth, BranchFactor>::findBestLeafIndex
// (V3D/Base/v3d_voctree.h:63-99) and VocabularyTreeTraversal<BranchFactor, Depth>
// (V3D/CUDA/v3d_cudamatching.cu:705-869).  Node array: the tree file's layout, level after
// level, the children of node x at (x + 1) * branch, the root's children at 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace v3dm {

constexpr int kVtCntSlots = 0;   // counters[0]: 32-key slots of the running level
constexpr int kVtCntExact = 1;   // counters[1]: keys decided exactly at the running level
constexpr int kVtCntFlagged = 2; // counters[2]: keys decided exactly over the whole call
constexpr int kVtCounters = 4;
constexpr int kVtNone = INT32_MIN;  // child range outside the node array
constexpr int kVtBinStride = 32;    // ints between bucket counters (one 128-byte line each)
constexpr int kVtLdsBins = 16384;   // parent levels up to this many nodes are counted in LDS

struct VtLevel {
    int branch;
    int level;              // 0 .. depth - 1
    long long parentStart;  // first node id of level - 1 (unused at level 0)
    int nParents;           // nodes of level - 1 (1 at level 0: the root)
    long long nNodes;
};

// reset the per-level counters (a kernel, so that no memset node is needed)
void launch_vt_reset(int* counters, bool all, hipStream_t st);
// bucket keys by their parent node (level >= 1); keys whose parent is not a node of level
// - 1 go straight to the exact list
// cnt, cursor: nParents * kVtBinStride ints
void launch_vt_bucket(int const* node, int n, VtLevel const& lv, int* cnt, int* cursor, int* order, int orderCap,
                      int* exactList, int* counters, hipStream_t st);
// certified MFMA scores of every key against the children of its parent; settled keys get
// their child, the others go to the exact list
void launch_vt_score(float const* keys, float const* keyNorm, int n, float const* nodes, float maxNodeNorm,
                     VtLevel const& lv, int const* order, bool identity, int maxSlots, int* node,
                     int* exactList, int* counters, hipStream_t st);
// exact decision (sequential FP32 sums, the reference's scan) of the listed keys
void launch_vt_exact(float const* keys, float const* nodes, VtLevel const& lv, int const* exactList,
                     int* node, int* counters, hipStream_t st);
void launch_vt_finish(int const* node, int n, long long leafStart, bool globalIds, int32_t* out, hipStream_t st);

}  // namespace v3dm
