// Vocabulary-tree quantisation: host side and C ABI.
//
// Replaces VocabularyTreeTraversal<BranchFactor, Depth> (V3D/CUDA/v3d_cudamatching.h:69-105,
// cu:795-869) and the per-key VocabularyTree::getVisualWord loop of the callers
// (V3D/Base/v3d_voctree.h:57-99, used at V3D/Apps/SfM/match_using_voctree.cpp:189-190).
// The result of every key is the reference's CPU descent (first child of maximal sequential
// FP32 score per level); globalIds selects the node id the CUDA traversal writes instead of
// the visual word (node id - first leaf id).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/v3dm.h"
#include "v3dm_device.h"
#include "v3dm_ext.h"
#include "v3dm_kernels.h"
#include "v3dm_voctree.h"

using namespace v3dm;

struct v3dm_voctree_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int branch = 0, depth = 0;
    long long nNodes = 0;
    std::vector<long long> levelStarts;  // depth + 1 entries
    long long maxParents = 1;            // nodes of the widest parent level (branch^(depth-1))
    float* dNodes = nullptr;             // (nNodes + 32) x 128, the last 32 rows zero
    float maxNodeNorm = 0.0f;
    bool haveNodes = false;

    int cap = 0;
    float* dKeys = nullptr;
    float* keyNorm = nullptr;
    unsigned* maxBits = nullptr;
    int* node = nullptr;
    int* order = nullptr;
    int* exactList = nullptr;
    int32_t* dOut = nullptr;
    int* cnt = nullptr;
    int* cursor = nullptr;
    int* counters = nullptr;
    int lastExact = 0;
};

namespace {

int vfail(int code, const std::string& msg) { return ctx_fail(nullptr, code, msg.c_str()); }

#define VTCHECK(call)                                                                                       \
    do {                                                                                                    \
        hipError_t const e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                               \
            return vfail(e_ == hipErrorOutOfMemory ? V3DM_ERR_NOMEM : V3DM_ERR_HIP,                         \
                         std::string(#call) + " failed: " + hipGetErrorString(e_));                          \
    } while (0)

template <typename T>
void vfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void releaseKeys(v3dm_voctree_s* t)
{
    vfree(t->dKeys);
    vfree(t->keyNorm);
    vfree(t->maxBits);
    vfree(t->node);
    vfree(t->order);
    vfree(t->exactList);
    vfree(t->dOut);
    t->cap = 0;
}

long long orderCapacity(long long n, long long parents) { return n + 31 * std::min(n, parents) + 32; }

int ensureCapacity(v3dm_voctree_s* t, int n)
{
    if (n <= t->cap) return V3DM_OK;
    releaseKeys(t);
    int const c = std::max(n, 1024);
    VTCHECK(hipMalloc(&t->dKeys, (size_t)c * kDim * sizeof(float)));
    VTCHECK(hipMalloc(&t->keyNorm, (size_t)c * sizeof(float)));
    VTCHECK(hipMalloc(&t->maxBits, 4 * sizeof(unsigned)));
    VTCHECK(hipMalloc(&t->node, (size_t)c * sizeof(int)));
    VTCHECK(hipMalloc(&t->order, (size_t)orderCapacity(c, t->maxParents) * sizeof(int)));
    VTCHECK(hipMalloc(&t->exactList, (size_t)c * sizeof(int)));
    VTCHECK(hipMalloc(&t->dOut, (size_t)c * sizeof(int32_t)));
    t->cap = c;
    return V3DM_OK;
}

// The descent for n device-resident keys; leaves (device) get the result.
int enqueueTraverse(v3dm_voctree_s* t, int n, float const* dKeys, int32_t* dLeaves, bool globalIds)
{
    hipStream_t const st = t->stream;
    launch_vt_reset(t->counters, true, st);
    if (n <= 0) return V3DM_OK;
    VTCHECK(hipMemsetAsync(t->maxBits, 0, sizeof(unsigned), st));
    launch_norms(dKeys, n, t->keyNorm, t->maxBits, st);
    VtLevel lv;
    lv.branch = t->branch;
    lv.nNodes = t->nNodes;
    for (int d = 0; d < t->depth; ++d) {
        lv.level = d;
        if (d == 0) {
            lv.parentStart = 0;
            lv.nParents = 1;
            launch_vt_score(dKeys, t->keyNorm, n, t->dNodes, t->maxNodeNorm, lv, nullptr, true, (n + 31) / 32,
                            t->node, t->exactList, t->counters, st);
        } else {
            launch_vt_reset(t->counters, false, st);
            lv.parentStart = t->levelStarts[d - 1];
            lv.nParents = (int)(t->levelStarts[d] - t->levelStarts[d - 1]);
            long long const oc = orderCapacity(n, lv.nParents) - 32;
            launch_vt_bucket(t->node, n, lv, t->cnt, t->cursor, t->order, (int)oc, t->exactList, t->counters, st);
            launch_vt_score(dKeys, t->keyNorm, n, t->dNodes, t->maxNodeNorm, lv, t->order, false,
                            (int)((oc + 31) / 32), t->node, t->exactList, t->counters, st);
        }
        launch_vt_exact(dKeys, t->dNodes, lv, t->exactList, t->node, t->counters, st);
    }
    launch_vt_finish(t->node, n, t->levelStarts[t->depth - 1], globalIds, dLeaves, st);
    VTCHECK(hipGetLastError());
    return V3DM_OK;
}

}  // namespace

extern "C" {

int v3dm_voctree_sizes(int branchFactor, int depth, int64_t* nodeCount, int64_t* leafCount)
{
    if (branchFactor < 1 || depth < 1) return vfail(V3DM_ERR_INVALID, "branchFactor and depth must be >= 1");
    long long nodes = 0, len = 1;
    for (int d = 0; d < depth; ++d) {
        len *= branchFactor;
        nodes += len;
        if (nodes > (1LL << 31) - 64) return vfail(V3DM_ERR_INVALID, "tree above 2^31 nodes");
    }
    if (nodeCount) *nodeCount = nodes;
    if (leafCount) *leafCount = len;
    return V3DM_OK;
}

int v3dm_voctree_create(v3dm_voctree* out, int device, int branchFactor, int depth)
{
    if (!out) return vfail(V3DM_ERR_INVALID, "null tree handle pointer");
    *out = nullptr;
    int64_t nodes = 0;
    int rc = v3dm_voctree_sizes(branchFactor, depth, &nodes, nullptr);
    if (rc) return rc;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return vfail(V3DM_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= count) return vfail(V3DM_ERR_INVALID, "device out of range");
    VTCHECK(hipSetDevice(device));
    auto* t = new v3dm_voctree_s();
    t->device = device;
    t->branch = branchFactor;
    t->depth = depth;
    t->nNodes = nodes;
    t->levelStarts.resize(depth + 1);
    long long s = 0, len = 1;
    for (int d = 0; d < depth; ++d) {  // v3d_voctree.h:21-35
        t->levelStarts[d] = s;
        len *= branchFactor;
        s += len;
    }
    t->levelStarts[depth] = s;
    t->maxParents = depth > 1 ? t->levelStarts[depth - 1] - t->levelStarts[depth - 2] : 1;
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&t->dNodes, (size_t)(nodes + 32) * kDim * sizeof(float));
    if (e == hipSuccess) e = hipMemset(t->dNodes, 0, (size_t)(nodes + 32) * kDim * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&t->cnt, (size_t)t->maxParents * kVtBinStride * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&t->cursor, (size_t)t->maxParents * kVtBinStride * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&t->counters, kVtCounters * sizeof(int));
    if (e != hipSuccess) {
        v3dm_voctree_destroy(t);
        return vfail(e == hipErrorOutOfMemory ? V3DM_ERR_NOMEM : V3DM_ERR_HIP,
                     std::string("tree allocation failed: ") + hipGetErrorString(e));
    }
    *out = t;
    return V3DM_OK;
}

void v3dm_voctree_destroy(v3dm_voctree t)
{
    if (!t) return;
    (void)hipSetDevice(t->device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    releaseKeys(t);
    vfree(t->dNodes);
    vfree(t->cnt);
    vfree(t->cursor);
    vfree(t->counters);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

int v3dm_voctree_set_nodes(v3dm_voctree t, const float* nodes)
{
    if (!t || !nodes) return vfail(V3DM_ERR_INVALID, "null argument");
    VTCHECK(hipSetDevice(t->device));
    // error bound: the largest node norm, rounded up (any NaN / Inf makes it non-finite,
    // which sends every key to the exact path)
    double mx = 0.0;
    bool finite = true;
    for (long long r = 0; r < t->nNodes; ++r) {
        double s = 0.0;
        for (int k = 0; k < kDim; ++k) {
            double const v = nodes[r * kDim + k];
            s += v * v;
        }
        if (!std::isfinite(s)) finite = false;
        mx = std::max(mx, s);
    }
    t->maxNodeNorm = finite ? (float)(std::sqrt(mx) * (1.0 + 1e-6)) : INFINITY;
    VTCHECK(hipMemcpyAsync(t->dNodes, nodes, (size_t)t->nNodes * kDim * sizeof(float), hipMemcpyHostToDevice,
                           t->stream));
    VTCHECK(hipStreamSynchronize(t->stream));
    t->haveNodes = true;
    return V3DM_OK;
}

int v3dm_voctree_load(v3dm_voctree t, const char* path)
{
    if (!t || !path) return vfail(V3DM_ERR_INVALID, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return vfail(V3DM_ERR_IO, std::string("Could not open ") + path);  // v3d_voctree.h:44-48
    // raw floats; a short file leaves the remaining nodes zero, as the reference's read into
    // its zero-initialised buffer does (v3d_voctree.h:49)
    std::vector<float> nodes((size_t)t->nNodes * kDim, 0.0f);
    size_t const got = fread(nodes.data(), sizeof(float), nodes.size(), f);
    fclose(f);
    (void)got;
    return v3dm_voctree_set_nodes(t, nodes.data());
}

int v3dm_voctree_traverse_device(v3dm_voctree t, int n, const float* d_keys, int32_t* d_leaves, int globalIds)
{
    if (!t) return vfail(V3DM_ERR_INVALID, "null tree handle");
    if (!t->haveNodes) return vfail(V3DM_ERR_STATE, "traverse before setNodes / loadTreeFromFile");
    if (n < 0 || (n > 0 && (!d_keys || !d_leaves))) return vfail(V3DM_ERR_INVALID, "bad key arguments");
    VTCHECK(hipSetDevice(t->device));
    int rc = ensureCapacity(t, n);
    if (rc) return rc;
    return enqueueTraverse(t, n, d_keys, d_leaves, globalIds != 0);
}

int v3dm_voctree_traverse(v3dm_voctree t, int n, const float* keys, int32_t* leaves, int globalIds)
{
    if (!t) return vfail(V3DM_ERR_INVALID, "null tree handle");
    if (!t->haveNodes) return vfail(V3DM_ERR_STATE, "traverse before setNodes / loadTreeFromFile");
    if (n < 0 || (n > 0 && (!keys || !leaves))) return vfail(V3DM_ERR_INVALID, "bad key arguments");
    VTCHECK(hipSetDevice(t->device));
    int rc = ensureCapacity(t, n);
    if (rc) return rc;
    if (n > 0) VTCHECK(hipMemcpyAsync(t->dKeys, keys, (size_t)n * kDim * sizeof(float), hipMemcpyHostToDevice,
                                      t->stream));
    rc = enqueueTraverse(t, n, t->dKeys, t->dOut, globalIds != 0);
    if (rc) return rc;
    if (n > 0)
        VTCHECK(hipMemcpyAsync(leaves, t->dOut, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, t->stream));
    int c[kVtCounters];
    VTCHECK(hipMemcpyAsync(c, t->counters, sizeof(c), hipMemcpyDeviceToHost, t->stream));
    VTCHECK(hipStreamSynchronize(t->stream));
    t->lastExact = c[kVtCntFlagged];
    return V3DM_OK;
}

int v3dm_voctree_synchronize(v3dm_voctree t, int* nExact)
{
    if (!t) return vfail(V3DM_ERR_INVALID, "null tree handle");
    VTCHECK(hipSetDevice(t->device));
    int c[kVtCounters];
    VTCHECK(hipMemcpyAsync(c, t->counters, sizeof(c), hipMemcpyDeviceToHost, t->stream));
    VTCHECK(hipStreamSynchronize(t->stream));
    t->lastExact = c[kVtCntFlagged];
    if (nExact) *nExact = t->lastExact;
    return V3DM_OK;
}

}  // extern "C"
