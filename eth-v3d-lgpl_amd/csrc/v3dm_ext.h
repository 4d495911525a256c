// Internal interface between the C ABI and the shortlist / guided / batch extensions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/v3dm.h"

namespace v3dm {

// A guided-matching mask applied since the last findPutativeMatches (cu:554-694): kind 0 =
// epipolar (Sampson distance with F), 1 = homography transfer distance with H.
struct GuidedMask {
    int kind;
    float m[9];
    float sqrThreshold;
};

struct ExtState;
struct BatchState;
struct LearnedState;
struct MlpWeights;

// accessors into the handle (v3dm_api.cpp)
hipStream_t ctx_stream(v3dm_context* h);
ExtState*& ctx_ext(v3dm_context* h);
BatchState*& ctx_batch(v3dm_context* h);
LearnedState*& ctx_learned(v3dm_context* h);
int ctx_fail(v3dm_context* h, int code, const char* msg);
int enqueue_plain(v3dm_context* h, float minScore, float minRatio);
// HIP events around the dominant kernel(s) of a call, when profiling is enabled
int prof_begin(v3dm_context* h);
int prof_end(v3dm_context* h);

// extension hooks (v3dm_ext.cpp)
bool ext_supports(v3dm_context* h, int mode);
int ext_enqueue(v3dm_context* h, int mode, float minScore, float minRatio);
// before a result is read: sampled check of a learned result, plain rerun if it fails
int ext_finish(v3dm_context* h);
// the dominant pass of shortlist mode `mode`, reps times between the two events
int ext_time_pass(v3dm_context* h, int mode, int reps, hipEvent_t start, hipEvent_t stop, int* launches);
// exact second best candidate of every query of the last shortlist call (host arrays)
int ext_second_best(v3dm_context* h, int* colSecond, int* rowSecond);
void ext_features_changed(v3dm_context* h, bool left);
// before the features (or the buffers holding them) are written: wait for the sampled
// check's side stream, which reads them, and make its next use wait for the handle's stream
void ext_quiesce(v3dm_context* h);
// wait for the side stream (the handle's synchronise covers the whole call)
void ext_synchronize(v3dm_context* h);
// a guided call replaces the last result: drop its pending sampled check and second bests
void ext_discard_result(v3dm_context* h);
void ext_release(ExtState* e);
void ext_destroy(ExtState*& e);
// batched matching (v3dm_batch.cpp)
void batch_destroy(BatchState*& b);
// learned network (v3dm_learned.cpp)
void learned_destroy(LearnedState*& L);
MlpWeights const* learned_weights(LearnedState* L);
float learned_code_scale(LearnedState* L);

}  // namespace v3dm
