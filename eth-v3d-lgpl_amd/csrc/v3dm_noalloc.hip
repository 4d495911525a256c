// findPutativeMatches (plain FP32 form) on device-resident descriptors without allocating:
// v3dm_match_device_noalloc.  For callers that cannot allocate device memory (code under
// HIP graph capture, serving loops, harnesses that forbid it).
//
// Every buffer beyond the caller's inputs and outputs lives in one file-scope device array,
// `scratch`, of kNoAllocScratchBytes: padded copies of both descriptor sets (zero rows up to
// a multiple of 128, which the pass kernels read whole), row norms, the partial top-2 of
// both directions, the decisions, the open-query list and re-check partials, the mutual
// check's status words, counters and striped counts.  The host never uses the array's name
// as an address: the first call on a device (with that device current) asks
// hipGetSymbolAddress for that device's copy and keeps the pointer per device.  The
// kernels are the handle path's (v3dm_fp32.hip, v3dm_select.hip); the result equals
// v3dm_find_putative_matches on the same descriptors.
//
// All calls on one GPU share its array: calls that overlap in time on one GPU (other
// streams, other threads) race.  Counters and norm maxima are cleared by a kernel at the
// start of every call; the status words carry a per-device epoch.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>

#include "../../include/v3dm.h"
#include "v3dm_device.h"
#include "v3dm_kernels.h"

namespace v3dm {

namespace {

constexpr int kNoAllocCap = V3DM_NOALLOC_MAX_FEATURES;  // rows per side (a multiple of 128)
constexpr size_t kAlign = 256;

constexpr size_t alignUp(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// byte offsets inside the scratch array
constexpr size_t kOffL = 0;
constexpr size_t kOffR = kOffL + alignUp((size_t)kNoAllocCap * kDim * sizeof(float));
constexpr size_t kOffNormL = kOffR + alignUp((size_t)kNoAllocCap * kDim * sizeof(float));
constexpr size_t kOffNormR = kOffNormL + alignUp((size_t)kNoAllocCap * sizeof(float));
constexpr size_t kOffMaxNorm = kOffNormR + alignUp((size_t)kNoAllocCap * sizeof(float));
constexpr size_t kOffColPart = kOffMaxNorm + kAlign;
constexpr size_t kOffRowPart = kOffColPart + alignUp((size_t)kMaxSplits * kNoAllocCap * sizeof(Top2));
constexpr size_t kOffColMax = kOffRowPart + alignUp((size_t)(kNoAllocCap / 128) * kNoAllocCap * sizeof(Top2));
constexpr size_t kOffRowMax = kOffColMax + alignUp((size_t)kNoAllocCap * sizeof(int));
constexpr size_t kOffFlag = kOffRowMax + alignUp((size_t)kNoAllocCap * sizeof(int));
constexpr size_t kOffRpart = kOffFlag + alignUp((size_t)2 * kNoAllocCap * sizeof(int));
constexpr size_t kOffStatus = kOffRpart + alignUp((size_t)2 * kNoAllocCap * kRecheckMaxChunks * sizeof(Top2));
constexpr size_t kOffStripes =
    kOffStatus + alignUp((size_t)(kNoAllocCap / kMutualSlice + 1) * sizeof(unsigned long long));
constexpr size_t kOffCounters = kOffStripes + alignUp((size_t)kStripes * kStripeStride * sizeof(int));
constexpr size_t kNoAllocScratchBytes = kOffCounters + kAlign;

static_assert(kNoAllocCap % 128 == 0, "the cap is a whole number of 128-row blocks");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// src rows [0, n) -> dst, rows [n, nPad) zero; one float4 per thread
__global__ __launch_bounds__(256) void k_noalloc_copy_pad(float const* __restrict__ src, int n, int nPad,
                                                          float* __restrict__ dst)
{
    size_t const i = (size_t)blockIdx.x * 256 + threadIdx.x;  // float4 index
    size_t const row = i / (kDim / 4);
    if (row >= (size_t)nPad) return;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (row < (size_t)n) v = reinterpret_cast<float4 const*>(src)[i];
    reinterpret_cast<float4*>(dst)[i] = v;
}

__global__ void k_noalloc_clear(int* __restrict__ counters, unsigned int* __restrict__ maxNorm)
{
    int const t = threadIdx.x;
    if (t < kNumCounters) counters[t] = 0;
    if (t < 2) maxNorm[t] = 0u;
}

// the match count, or -1 when the mutual check's ordering wait ran out (its epoch in
// counters[kCntMutualError])
__global__ void k_noalloc_count(int const* __restrict__ counters, unsigned epoch, int* __restrict__ count)
{
    if (threadIdx.x == 0) *count = (counters[kCntMutualError] == (int)epoch) ? -1 : counters[kCntPairs];
}

struct DeviceScratch {
    unsigned char* base = nullptr;
    std::atomic<unsigned> epoch{0};
};

DeviceScratch gScratch[64];
std::mutex gScratchMutex;

// The current device's copy of `scratch` (looked up once per device), or an error.
hipError_t scratch_of_current_device(DeviceScratch*& out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    DeviceScratch& s = gScratch[dev];
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!s.base) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        s.base = static_cast<unsigned char*>(ptr);
    }
    out = &s;
    return hipSuccess;
}

thread_local char gNoAllocErr[256] = "";

int noalloc_fail(int code, char const* msg, hipError_t e = hipSuccess)
{
    if (e != hipSuccess) snprintf(gNoAllocErr, sizeof gNoAllocErr, "%s: %s", msg, hipGetErrorString(e));
    else snprintf(gNoAllocErr, sizeof gNoAllocErr, "%s", msg);
    return code;
}

}  // namespace

}  // namespace v3dm

using namespace v3dm;

extern "C" {

int v3dm_match_device_noalloc(const float* leftDesc, int nLeft, const float* rightDesc, int nRight, float minScore,
                              float minRatio, int32_t* pairs, int32_t* count, void* stream)
{
    if (nLeft < 0 || nRight < 0 || nLeft > kNoAllocCap || nRight > kNoAllocCap)
        return noalloc_fail(V3DM_ERR_INVALID, "feature counts must be in [0, V3DM_NOALLOC_MAX_FEATURES]");
    if ((nLeft > 0 && !leftDesc) || (nRight > 0 && !rightDesc) || !pairs || !count)
        return noalloc_fail(V3DM_ERR_INVALID, "null pointer");
    int const nL = nLeft / 32 * 32, nR = nRight / 32 * 32;  // cu:493-518
    DeviceScratch* ds = nullptr;
    hipError_t e = scratch_of_current_device(ds);
    if (e != hipSuccess) return noalloc_fail(V3DM_ERR_HIP, "hipGetSymbolAddress(scratch)", e);
    hipStream_t const st = static_cast<hipStream_t>(stream);
    unsigned char* const b = ds->base;
    float* const L = reinterpret_cast<float*>(b + kOffL);
    float* const R = reinterpret_cast<float*>(b + kOffR);
    float* const normL = reinterpret_cast<float*>(b + kOffNormL);
    float* const normR = reinterpret_cast<float*>(b + kOffNormR);
    unsigned int* const maxNorm = reinterpret_cast<unsigned int*>(b + kOffMaxNorm);
    Top2* const colPart = reinterpret_cast<Top2*>(b + kOffColPart);
    Top2* const rowPart = reinterpret_cast<Top2*>(b + kOffRowPart);
    int* const colMax = reinterpret_cast<int*>(b + kOffColMax);
    int* const rowMax = reinterpret_cast<int*>(b + kOffRowMax);
    int* const flagList = reinterpret_cast<int*>(b + kOffFlag);
    Top2* const rpart = reinterpret_cast<Top2*>(b + kOffRpart);
    unsigned long long* const status = reinterpret_cast<unsigned long long*>(b + kOffStatus);
    int* const stripes = reinterpret_cast<int*>(b + kOffStripes);
    int* const counters = reinterpret_cast<int*>(b + kOffCounters);

    unsigned ep = ds->epoch.fetch_add(1) + 1;
    if (ep == 0) ep = ds->epoch.fetch_add(1) + 1;  // epoch 0 marks never-written status words
    int const capL = std::max(128, (nL + 127) / 128 * 128), capR = std::max(128, (nR + 127) / 128 * 128);
    bool const work = nL > 0 && nR > 0;
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, st, counters, maxNorm);
    if (work) {
        hipLaunchKernelGGL(k_noalloc_copy_pad, dim3((capL * (kDim / 4) + 255) / 256), dim3(256), 0, st, leftDesc, nL,
                           capL, L);
        hipLaunchKernelGGL(k_noalloc_copy_pad, dim3((capR * (kDim / 4) + 255) / 256), dim3(256), 0, st, rightDesc, nR,
                           capR, R);
        launch_norms(L, nL, normL, maxNorm + 0, st);
        launch_norms(R, nR, normR, maxNorm + 1, st);
        FusedPlan const p = plan_fused_fp32(nL, nR, 512);
        launch_fused_fp32(L, R, nL, nR, p, colPart, capL, rowPart, capR, st);
        launch_merge_decide(colPart, p.splits, capL, rowPart, p.leftBlocks, capR, nL, nR, normL, normR, maxNorm,
                            kEpsRelFp32, minScore, minRatio, colMax, rowMax, flagList, capL, counters, st);
        launch_recheck_exact(L, R, nL, nR, flagList, capL, counters, minScore, minRatio, colMax, rowMax, rpart,
                             kRecheckGrid, st);
    }
    launch_mutual_compact(colMax, rowMax, work ? nL : 0, nR, status, ep, stripes, reinterpret_cast<int2*>(pairs),
                          counters, st);
    hipLaunchKernelGGL(k_noalloc_count, dim3(1), dim3(64), 0, st, counters, ep, count);
    e = hipGetLastError();
    if (e != hipSuccess) return noalloc_fail(V3DM_ERR_HIP, "launch", e);
    return V3DM_OK;
}

const char* v3dm_noalloc_last_error(void) { return gNoAllocErr; }

}  // extern "C"
