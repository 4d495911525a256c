/*
 * v3dm -- C ABI of the MI355X (gfx950) SIFT putative matcher.
 *
 * Drop-in for the CUDA matcher V3D_CUDA::SIFT_Matching (V3D/CUDA/v3d_cudamatching.h:15-67,
 * V3D/CUDA/v3d_cudamatching.cu:439-694).  One opaque handle per matcher object; every
 * call is synchronous with respect to the host unless its name ends in _async.  A handle
 * is not thread-safe; separate handles may be used from separate threads.
 *
 * Return value of every int function: V3DM_OK (0) or an error code; the message of the
 * last error of a handle is returned by v3dm_last_error(handle) (v3dm_last_error(NULL)
 * for v3dm_create).  The reference prints and exit()s on CUDA errors (cu:12-27); this
 * library returns the code instead and leaves the process alive.
 *
 * Descriptors are 128 floats per feature, row-major (feature k at desc + 128*k), exactly
 * the layout the reference copies with cublasSetMatrix (cu:503, cu:519).
 */
#ifndef V3DM_H
#define V3DM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct v3dm_context* v3dm_handle;

enum {
    V3DM_OK = 0,
    V3DM_ERR_INVALID = 1,      /* bad argument (negative count, null pointer, ...) */
    V3DM_ERR_STATE = 2,        /* call out of order (e.g. features set before allocate) */
    V3DM_ERR_CAPACITY = 3,     /* output buffer too small: *nPairs holds the required count */
    V3DM_ERR_HIP = 4,          /* HIP runtime error */
    V3DM_ERR_UNSUPPORTED = 5,  /* feature not available in this build / mode */
    V3DM_ERR_NOMEM = 6,        /* device or host allocation failed */
    V3DM_ERR_IO = 7            /* file missing, unreadable or malformed */
};

/* Precision modes of findPutativeMatches (library extension; the reference has FP32 only). */
enum {
    V3DM_PREC_FP32 = 0,    /* exact match set (certified FP32 MFMA + FP64 re-check) */
    V3DM_PREC_BF16 = 1,    /* BF16 MFMA shortlist, exact rescoring */
    V3DM_PREC_FP8 = 2,     /* FP8 e4m3 MFMA shortlist, exact rescoring */
    V3DM_PREC_LEARNED = 3  /* learned 32-d code shortlist, exact rescoring */
};

/* Counters of the last findPutativeMatches / guided call (library extension). */
typedef struct v3dm_stats {
    int nLeft, nRight;        /* feature counts used (after rounding / clamping) */
    int nMatches;             /* matches found */
    int nRechecked;           /* queries whose decision was settled by the exact re-check */
    int precision;            /* V3DM_PREC_* used */
    int nFallback;            /* pairs that fell back to the exact path (shortlist modes) */
    double kernelMs;          /* accumulated time of the dominant kernel since the last reset
                                 (only when profiling is enabled) */
    int kernelLaunches;       /* launches accumulated in kernelMs */
    int nRescored;            /* shortlist candidates rescored exactly (shortlist modes) */
    int nSampled;             /* learned mode: left features decided exactly as a check */
    int nSampleMissed;        /* ... of which the learned decision missed the exact match */
} v3dm_stats;

/* Handle life cycle (constructor / destructor of SIFT_Matching, h:17-21).  device is a HIP
 * device ordinal. */
int v3dm_create(v3dm_handle* h, int device);
void v3dm_destroy(v3dm_handle h);
const char* v3dm_last_error(v3dm_handle h);

/* SIFT_Matching::allocate(nMaxLeft, nMaxRight, useGuidedMatching), cu:439-474.
 * Both maxima are rounded down to a multiple of 32 (warning on stderr if enabled). */
int v3dm_allocate(v3dm_handle h, int nMaxLeft, int nMaxRight, int useGuidedMatching);
/* SIFT_Matching::deallocate(), cu:476-488. */
int v3dm_deallocate(v3dm_handle h);
/* enableRoundingWarning / enableReductionPath, h:23-31.  The reduction path is accepted
 * and gives the canonical (non-reduction) selection semantics. */
int v3dm_set_flags(v3dm_handle h, int warnOnRounding, int useReductionPath);

/* setLeftFeatures / setRightFeatures, cu:490-520: n is rounded down to a multiple of 32 and
 * clamped to the allocated maximum; desc is a host array of n*128 floats. */
int v3dm_set_left(v3dm_handle h, int n, const float* desc);
int v3dm_set_right(v3dm_handle h, int n, const float* desc);

/* findPutativeMatches(matches, minScore, minRatio), cu:522-552.  Writes up to cap
 * (left, right) index pairs to pairs[2*k], pairs[2*k+1], sorted by left index, and the
 * number of matches to *nPairs.  minRatio < 0 disables the ratio test.  Returns
 * V3DM_ERR_CAPACITY (with the full count in *nPairs) if cap is too small; cap >=
 * min(nLeft, nRight) always suffices. */
int v3dm_find_putative_matches(v3dm_handle h, float minScore, float minRatio, int32_t* pairs, int cap,
                               int* nPairs);

/* runGuidedMatching (F, Sampson distance) and runGuidedHomographyMatching (H, transfer
 * distance), cu:554-694.  M is a row-major 3x3 double matrix; positions are AoS (x, y)
 * floats of the current left / right features.  Candidates outside the distance
 * threshold get the score -1e20 (cu:308, cu:340).  As in the reference, the mask applies
 * to the scores of the last findPutativeMatches and repeated guided calls compound (until
 * the next findPutativeMatches or set*Features).  The scores are recomputed from the
 * current descriptors in exact FP32, whatever the precision mode.  Requires
 * allocate(..., useGuidedMatching = 1): it holds one bit per (left, right) pair. */
int v3dm_run_guided_F(v3dm_handle h, const double F[9], float distanceThreshold, const float* leftXY,
                      const float* rightXY, float minScore, float minRatio, int32_t* pairs, int cap, int* nPairs);
int v3dm_run_guided_H(v3dm_handle h, const double H[9], float distanceThreshold, const float* leftXY,
                      const float* rightXY, float minScore, float minRatio, int32_t* pairs, int cap, int* nPairs);

/* ---- extensions (no counterpart in the reference) ---------------------------------- */

/* Select the precision mode (V3DM_PREC_*) of later findPutativeMatches calls. */
int v3dm_set_precision(v3dm_handle h, int mode);
/* Error-bound scale of the certified FP32 decisions: a query is settled from the MFMA
 * scores only when every exact score set within epsRel * ||q|| * max||c|| gives the same
 * decision; otherwise it is re-checked exactly in FP64.  Default 2^-15; 0 re-checks every
 * query (the whole selection then runs on exact FP64 scores). */
int v3dm_set_certification(v3dm_handle h, float epsRel);
/* Candidates per query scored exactly in the learned mode (the best k by code score;
 * default 6, at most 32). */
int v3dm_set_learned_topk(v3dm_handle h, int k);
/* ... of which those whose code score is more than marginRel * scale^2 below the best
 * candidate's are not rescored (scale: the network's code scale; scale^2 is about the code
 * score of a code with itself).  Default 0.07; a negative value turns the margin off. */
int v3dm_set_learned_margin(v3dm_handle h, float marginRel);
/* 1 (default): the bf16 / fp8 / code copies of the features are kept between calls while
 * the features are unchanged; 0: every call derives them again (so that a call's time
 * includes that work). */
int v3dm_set_derived_cache(v3dm_handle h, int enable);
/* Learned mode: every call also decides `samples` evenly spaced features (at most 128;
 * 0 = off) with the exact plain path.  Of the a sampled features that the exact path
 * accepts, m have a different learned decision (missed); when the sampled recall
 * (a - m) / a falls below minRecall (V3DM_FALLBACK_THRESHOLD, the default rule), or when m
 * is more than a recall of minRecall explains at the 5% level (V3DM_FALLBACK_BINOMIAL), the
 * result is replaced by the plain form's before it is returned (stats.nFallback = 1).
 * minRecall 1: any miss falls back; 0: never.  Default 128 and 0.995.  The batch form
 * samples exact mutual matches instead (v3dm_batch_match_async). */
int v3dm_set_fallback(v3dm_handle h, int samples, float minRecall);
#define V3DM_FALLBACK_THRESHOLD 0
#define V3DM_FALLBACK_BINOMIAL 1
int v3dm_set_fallback_rule(v3dm_handle h, int rule);
/* Test hook of the failure path of the mutual check's ordered compaction (no counterpart
 * in the reference; for tests only).  Each work-group of the single-pair mutual check waits,
 * boundedly, for its predecessors' counts; when the wait runs out the call fails with
 * V3DM_ERR_HIP ("ordering wait ... ran out") and the handle stays usable.  spinLimit: the
 * sleeps before giving up (< 0: the default, 2^20); neverReady = 1 makes no predecessor
 * count visible, so every work-group but the first runs out of its wait (with several
 * work-groups, i.e. more than 1024 left features).  0, -1 restores normal operation. */
int v3dm_debug_mutual_wait(v3dm_handle h, int spinLimit, int neverReady);
/* The misses a sampled check of `accepted` exact matches tolerates at recall minRecall
 * under the threshold rule: the largest m with (float)(accepted - m) / (float)accepted not
 * below minRecall (0 for minRecall 1, `accepted` for minRecall 0); -1 when accepted is
 * outside [0, 2^24] or minRecall outside [0, 1]. */
int v3dm_fallback_allowed_misses(int accepted, float minRecall);
/* ... under the binomial rule: the smallest k with P[X > k] <= 0.05 for
 * X ~ Binomial(accepted, 1 - minRecall) (a one-sided test at 5%: a result whose sample
 * misses more is taken to fall short of minRecall). */
int v3dm_fallback_allowed_misses_binomial(int accepted, float minRecall);
/* Measurement helper: launches the dominant kernel of the current precision mode (plain:
 * the fused score pass; BF16 / FP8 / learned: the pass over both directions) `reps` times
 * back to back on the inputs of the last findPutativeMatches,
 * with one HIP event pair around all launches on the handle's stream, and returns the
 * average duration of one launch.  Overwrites the intermediate partial results (not the
 * match list or colMax / rowMax of the last call). */
int v3dm_time_dominant_kernel(v3dm_handle h, int reps, float* msPerLaunch);
/* findPutativeMatches of the plain FP32 form on device-resident descriptors, allocating
 * nothing.  No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * leftDesc / rightDesc: device arrays of nLeft / nRight x 128 floats (counts rounded down
 * to multiples of 32, as setLeftFeatures does, cu:493-518; each at most
 * V3DM_NOALLOC_MAX_FEATURES); pairs: device array of at least 2 * min(nLeft, nRight) ints
 * receiving (left, right) sorted by left; count: device int receiving the number of
 * matches (-1: the mutual check's ordering wait ran out).  Enqueued on `stream` (a
 * hipStream_t; NULL = the null stream), asynchronous; the match set equals
 * v3dm_find_putative_matches's.  The work buffers are one file-scope device array of
 * about 32 MiB per GPU (held until the process ends), shared by every call on that GPU:
 * calls that overlap on one GPU race.  Errors: V3DM_ERR_INVALID (counts outside [0,
 * V3DM_NOALLOC_MAX_FEATURES], null pointers), V3DM_ERR_HIP; message in
 * v3dm_noalloc_last_error(). */
#define V3DM_NOALLOC_MAX_FEATURES 8192
int v3dm_match_device_noalloc(const float* leftDesc, int nLeft, const float* rightDesc, int nRight, float minScore,
                              float minRatio, int32_t* pairs, int32_t* count, void* stream);
const char* v3dm_noalloc_last_error(void);
/* Measurement helper (no reference analogue): the approximate scores the pass of mode
 * V3DM_PREC_FP32 / BF16 / FP8 computes for the first nRows left x nCols right features of
 * the current sets (same instruction, operand rounding and summation order as the pass,
 * which never writes its scores), row-major into the host array scores[nRows * nCols].
 * Synchronous. */
int v3dm_score_block(v3dm_handle h, int mode, int nRows, int nCols, float* scores);
/* Run the library's kernels on this HIP stream (hipStream_t, NULL = the handle's own). */
int v3dm_set_stream(v3dm_handle h, void* hipStream);
/* Feature upload from device memory (n*128 floats, device pointer); no host copy. */
int v3dm_set_left_device(v3dm_handle h, int n, const float* d_desc);
int v3dm_set_right_device(v3dm_handle h, int n, const float* d_desc);
/* Feature upload as uint8 SIFT descriptors (host array n*128 bytes), dequantised on the
 * device exactly as SIFT_Feature::normalizeDescriptor (V3D/Base/v3d_feature.h:30-38). */
int v3dm_set_left_u8(v3dm_handle h, int n, const uint8_t* q);
int v3dm_set_right_u8(v3dm_handle h, int n, const uint8_t* q);
/* findPutativeMatches without the host copy: enqueues the work on the stream and returns.
 * The result stays on the device until v3dm_get_matches. */
int v3dm_find_putative_matches_async(v3dm_handle h, float minScore, float minRatio);
/* Waits for the last _async call and copies its matches (as v3dm_find_putative_matches). */
int v3dm_get_matches(v3dm_handle h, int32_t* pairs, int cap, int* nPairs);
/* Device pointers of the last result: int2 pairs[], int counters[] (counters[0] = count). */
int v3dm_device_results(v3dm_handle h, const int32_t** d_pairs, const int32_t** d_counters);
/* Per-query decisions of the last call (colMax: nLeft ints, rowMax: nRight ints, -1 = none),
 * the arrays of cu:532-533 after the selection kernels. */
int v3dm_get_maxima(v3dm_handle h, int32_t* colMax, int32_t* rowMax);
/* Block until all work the handle enqueued has finished (its stream, and the side stream
 * of the learned mode's sampled check). */
int v3dm_synchronize(v3dm_handle h);
/* Time the dominant kernel with HIP events on the handle's stream (1 = on, resets sums). */
int v3dm_set_profiling(v3dm_handle h, int enable);
int v3dm_get_stats(v3dm_handle h, v3dm_stats* out);
/* After a shortlist-mode call (BF16 / FP8): for every left (right) feature the feature of
 * the other set with the second best exact score among the rescored candidates (-1 if
 * none) -- the hardest distractor, used to label training triplets. */
int v3dm_get_second_best(v3dm_handle h, int32_t* colSecond, int32_t* rowSecond);

/* ---- SIFT feature archives (".features" files) ------------------------------------------
 * The caller-side input of the matcher: V3D::serializeDataToFile / serializeDataFromFile of
 * a SerializableVector<SIFT_Feature> (V3D/Base/v3d_serialization.h:953-1017, record
 * V3D/Base/v3d_feature.h:41-49), as loadFeatures reads them
 * (V3D/Apps/SfM/match_sequence.cpp:25-46).  Binary ("V3DBIN" + uint32 count + 540 bytes per
 * feature: int32 id, float x, y, direction x, y, scale, magnitude, 128 descriptor floats,
 * little-endian) and text ("V3DTXT") archives.  Host only (no handle, no GPU); errors are
 * reported by v3dm_last_error(NULL).  An unknown tag is V3DM_ERR_IO, as the reference's
 * "Unknown archive magic tag" (v3d_serialization.h:1000). */
/* Number of features and the format (1 binary, 0 text) of an archive. */
int v3dm_features_count(const char* path, int* nFeatures, int* binary);
/* The first min(count, cap) features; each output may be NULL (field skipped): ids[n],
 * xy[2n] (x, y), dir[2n], scale[n], magnitude[n], desc[128n] (row-major).  *n = features
 * read. */
int v3dm_features_read(const char* path, int cap, int32_t* ids, float* xy, float* dir, float* scale,
                       float* magnitude, float* desc, int* n);
/* Writes n features (binary = 1: V3DBIN, 0: V3DTXT with 6 significant digits, as the
 * reference's text archive); NULL ids / xy / dir / scale / magnitude write k / zeros. */
int v3dm_features_write(const char* path, int n, const int32_t* ids, const float* xy, const float* dir,
                        const float* scale, const float* magnitude, const float* desc, int binary);

/* ---- vocabulary-tree quantisation ------------------------------------------------------
 * VocabularyTreeTraversal<BranchFactor, Depth> (V3D/CUDA/v3d_cudamatching.h:69-105,
 * cu:705-869) and VocabularyTree<Depth, BranchFactor>::getVisualWord / findBestLeafIndex
 * (V3D/Base/v3d_voctree.h:19-99, the per-feature loop of
 * V3D/Apps/SfM/match_using_voctree.cpp:189-190).  Nodes: getNodeCount() x 128 floats,
 * level after level, the children of node x at (x + 1) * branchFactor (the tree file's
 * layout).  Per level the descent takes the first child of maximal score, the score being
 * the sequential FP32 sum of the 128 products (the CPU tree's rule; the CUDA kernel's
 * summation order is not reproduced).  Errors via v3dm_last_error(NULL). */
typedef struct v3dm_voctree_s* v3dm_voctree;
/* getNodeCount / getLeaveCount (h:84-101). */
int v3dm_voctree_sizes(int branchFactor, int depth, int64_t* nodeCount, int64_t* leafCount);
/* Constructor + allocate (cu:795-817); the key capacity grows on demand. */
int v3dm_voctree_create(v3dm_voctree* t, int device, int branchFactor, int depth);
/* deallocate (cu:819-826) + destructor. */
void v3dm_voctree_destroy(v3dm_voctree t);
/* setNodes (cu:828-838): host array of nodeCount x 128 floats. */
int v3dm_voctree_set_nodes(v3dm_voctree t, const float* nodes);
/* loadTreeFromFile (v3d_voctree.h:38-51): raw floats; a short file leaves the remaining
 * nodes zero, as the reference does.  V3DM_ERR_IO if the file cannot be opened. */
int v3dm_voctree_load(v3dm_voctree t, const char* path);
/* traverseTree (cu:840-860) for n keys (host arrays, n x 128 floats).  globalIds = 1:
 * node ids, as the CUDA traversal writes them (cu:784-791); 0: visual words in
 * [0, leafCount), as getVisualWord returns them (v3d_voctree.h:78).  A key whose descent
 * leaves the node array (possible only when no child scores above -1e30, e.g. NaN keys,
 * where the reference reads out of range) gets INT32_MIN. */
int v3dm_voctree_traverse(v3dm_voctree t, int n, const float* keys, int32_t* leaves, int globalIds);
/* The same on device-resident keys and output, enqueued on the tree's stream (no host
 * synchronisation); v3dm_voctree_synchronize waits for it and returns the number of keys
 * the last call decided by the exact (sequential FP32) path. */
int v3dm_voctree_traverse_device(v3dm_voctree t, int n, const float* d_keys, int32_t* d_leaves, int globalIds);
int v3dm_voctree_synchronize(v3dm_voctree t, int* nExact);

/* ---- batched matching (extension) ----------------------------------------------------
 * Replaces the caller loop that matches many image pairs one at a time
 * (V3D/Apps/SfM/match_sequence.cpp:165-382, match_using_voctree.cpp:264-609: per pair
 * setLeftFeatures / setRightFeatures / findPutativeMatches, cu:490-552).  The images of a
 * batch are uploaded once; a pair list names (left image, right image) per pair; one call
 * matches every pair with the handle's precision mode (FP32 exact, BF16 shortlist or
 * learned codes) through persistent kernels over a device work queue, replayed from a captured hipGraph.
 * Per pair the result is what findPutativeMatches gives for that pair: counts rounded down
 * to a multiple of 32, matches (left, right) sorted by left index.  Independent of
 * v3dm_allocate. */
/* nImages images, counts[m] features each, descriptors concatenated (image m starts at row
 * counts[0] + ... + counts[m-1]); uint8 descriptors are dequantised as
 * SIFT_Feature::normalizeDescriptor (V3D/Base/v3d_feature.h:30-38). */
int v3dm_batch_set_images_u8(v3dm_handle h, int nImages, const int32_t* counts, const uint8_t* desc);
int v3dm_batch_set_images(v3dm_handle h, int nImages, const int32_t* counts, const float* desc);
/* Pair p matches image left[p] (left features) against image right[p]. */
int v3dm_batch_set_pairs(v3dm_handle h, int nPairs, const int32_t* left, const int32_t* right);
/* Enqueue the batch (no host synchronisation). */
int v3dm_batch_match_async(v3dm_handle h, float minScore, float minRatio);
/* Wait for the last batch run; offsets[0..nPairs] (may be NULL) delimit each pair's matches
 * in pairs[2*k], pairs[2*k+1]; *nTotal = total count (V3DM_ERR_CAPACITY if cap < total). */
int v3dm_batch_get_matches(v3dm_handle h, int32_t* offsets, int32_t* pairs, int cap, int* nTotal);
/* v3dm_batch_match_async followed by v3dm_batch_get_matches. */
int v3dm_batch_match(v3dm_handle h, float minScore, float minRatio, int32_t* offsets, int32_t* pairs, int cap,
                     int* nTotal);
/* Work-item tuning of the persistent kernels: autotune != 0 measures the candidate
 * (tiles per item, grid) settings on the batch at its first run; tilesPerItem / grid > 0
 * fix them for the form of `mode` instead.  In the learned form tilesPerItem is the minimum
 * split size of the candidate lists (default 32 tiles of 64 candidates, at most 8 splits
 * per query): it changes which candidates are listed, so it is never tuned. */
int v3dm_batch_set_tuning(v3dm_handle h, int autotune, int mode, int tilesPerItem, int grid);
/* The work-group tile shape of a form's pass: queries (plain: left features) per work-group
 * x candidate rows per tile, 128 x 64 or 256 x 64 in every form (plain FP32: 4 or 8 waves;
 * BF16 and learned codes: 2 or 4 waves), chosen by the autotuner with the tiles per item and
 * the grid (no shape changes the result: the plain form's match set is the exact one in
 * every shape, and the shortlist forms list the same candidates).  set: queryRows 128 or 256
 * fixes the shape of the form of `mode`, 0 lets the tuner choose (default without tuning:
 * plain 128, BF16 and learned 256); get: the shape of the form's current plan. */
int v3dm_batch_set_tile_shape(v3dm_handle h, int mode, int queryRows);
int v3dm_batch_get_tile_shape(v3dm_handle h, int mode, int* queryRows, int* candidateRows);
/* The setting in use for `mode` and the autotuner's measured pass time (-1 if not tuned). */
int v3dm_batch_get_tuning(v3dm_handle h, int mode, int* tilesPerItem, int* grid, int* nItems, float* passMs);
/* 1 (default): replay batch runs from a captured hipGraph; 0: launch kernel by kernel. */
int v3dm_batch_use_graph(v3dm_handle h, int enable);
/* Learned batch runs check every pair against its exact match set (v3dm_set_fallback):
 * S = min(samples, nA) evenly spaced features of the pair's larger image A (left on a tie)
 * are decided exactly, and so are the features of the other image they match; an exact
 * mutual match among the samples that the learned result lacks is a miss.  A pair whose
 * sampled recall falls short under the handle's rule (v3dm_set_fallback_rule) gets the
 * plain form's matches, in the same run.  (Batches uploaded as FP32 compare the samples'
 * own decisions instead.)  flags[p] = 1 for the pairs of the last run that did. */
int v3dm_batch_get_fallback(v3dm_handle h, int32_t* flags);
/* Measurement helper: the pass kernel of the last batch run's mode (plain: the fused FP32
 * pass; BF16: the BF16 pass; learned: the code pass), launched `reps` times back to back
 * with its plan between one HIP event pair on the handle's stream; average ms per launch.
 * Leaves the last run's matches readable. */
int v3dm_batch_time_dominant_kernel(v3dm_handle h, int reps, float* msPerLaunch);
/* Per pair of the last learned batch run, 4 ints: exact matches among the samples,
 * missed by the learned result, sampled decisions the certified scores left to the exact
 * FP64 scores, samples. */
int v3dm_batch_get_check(v3dm_handle h, int32_t* counts);
/* GPU time of the last batch run (HIP events around the whole run), ms. */
int v3dm_batch_last_ms(v3dm_handle h, float* ms);

/* ---- learned shortlist network (extension; the reference has no learned stage) -------
 * A 128 -> 256 -> 32 ReLU MLP embeds descriptors into 32-d codes whose dot products
 * propose candidates; the proposed candidates are then scored exactly.  Parameters are
 * one packed FP32 vector of V3DM_LEARNED_PARAMS floats: W1 [256][128], b1 [256],
 * W2 [32][256], b2 [32] (row-major, in that order).  Training runs on the GPU with BF16
 * MFMA forward / backward kernels, FP32 master weights and Adam. */
#define V3DM_LEARNED_PARAMS 41248
/* Parameters and Adam moments (m, v may be NULL: zero) at Adam step count `step`. */
int v3dm_learned_set_state(v3dm_handle h, const float* params, const float* m, const float* v, int step);
int v3dm_learned_get_state(v3dm_handle h, float* params, float* m, float* v, int* step);
/* Quantisation scale of the unit-norm codes used by the code pass (codes * scale are
 * rounded to integers in [-31, 31]). */
int v3dm_learned_set_code_scale(v3dm_handle h, float scale);
/* Training set: nPool descriptor rows (host, 128 floats each) and nTriplets triplets
 * (anchor, positive, negative) of pool rows (3 int32 each). */
int v3dm_learned_set_training(v3dm_handle h, int nPool, const float* pool, int nTriplets, const int32_t* triplets);
/* `steps` Adam steps on minibatches of `batch` triplets (reshuffled every epoch with
 * `seed`); loss temperature tau; losses[s] = mean minibatch loss of step s (may be NULL). */
int v3dm_learned_train(v3dm_handle h, int steps, int batch, float lr, float tau, uint64_t seed, float* losses);
/* Loss and FP32 gradient (packed layout) of one minibatch (indices into the training
 * triplets), without an update; hard[t] = in-batch hardest positive chosen for anchor t. */
int v3dm_learned_gradients(v3dm_handle h, int nBatch, const int32_t* tripletIdx, float tau, float* loss,
                           float* grads, int32_t* hard);
/* Codes before normalisation (n x 32 floats) of n host descriptors. */
int v3dm_learned_embed(v3dm_handle h, int n, const float* desc, float* codes);

#ifdef __cplusplus
}
#endif

#endif /* V3DM_H */
