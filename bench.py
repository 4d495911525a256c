"""Benchmark of the gfx950 SIFT putative matcher (one GPU).

Headline workload (the metric's configuration): the uneven batch of 256 image pairs, each
image with 500-16000 features (log-uniform), 30-60% noisy copies per pair, 128-d uint8
descriptors, minScore 0.9, minRatio 0.8.  A step is one batch run (v3dm_batch_match_async)
of the learned shortlist form: embedding of every image (derived-copy cache off), code pass,
exact rescoring, the sampled exact check of every pair and the plain FP32 rerun of the
pairs that fail it, mutual check -- with the descriptors resident in HBM and the matches
left in HBM.  The plain FP32 form (identical to the oracle's match set) and the BF16 form
run the same batch; the learned form's recall and precision are measured against the
plain set, per pair and in aggregate.  The headline `value` is the learned form's image
pairs/s, `roofline` its dominant kernel (the code pass) at the algorithmic work of one
2*N*M*32 int8 product per pair, and `cpu_baseline` the CPU port (oracle/cpu_baseline.c,
FP32 blocked GEMM + top-2 + mutual, OpenMP) over the same 256 pairs.

Also reported: the single 16384 x 16384 pair in all four forms, the FP8 form on a 32768 x
32768 pair, the plain form on a 4096 x 4096 pair with its CPU baseline, the FP64 score
error of the FP32 / BF16 / FP8 passes per pair size, guided matching, vocabulary-tree
quantisation and the robust-orientation consumer.  HBM traffic per launch comes from the
rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/traffic.json (null when
that file has no entry for the kernel).

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "eth-v3d-lgpl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

PEAK_FP32_MFMA = 157.3e12
PEAK_BF16_MFMA = 2.5e15
PEAK_FP8_MFMA = 5.0e15
PEAK_INT8_MFMA = 5.0e15
PEAK_HBM = 8.0e12
N_PAIR = 16384
N_FP8 = 32768
N_SMALL = 4096
BATCH_PAIRS = 256
MIN_SCORE, MIN_RATIO = 0.9, 0.8
METRIC = "image pairs matched/s @\u226599.5% match recall, learned/BF16 vs plain FP32"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration(nbytes):
    """One hipMemcpy device-to-device of nbytes (the runtime's copy kernel reads and writes
    exactly nbytes), used to scale the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"bench: copied {nbytes} bytes device to device")


def _sets(res):
    return [set(map(tuple, r.tolist())) for r in res]


def time_batch(m, prec, steps, warmup, read_back_steps=3):
    """One mode on the batch: the first run (autotuning, graph capture) returns the match
    lists; then `warmup` and `steps` back-to-back runs with the matches left in HBM (ms per
    batch); GPU time of the last run (events around the run); and runs whose matches are
    copied to host memory before the next run (the reference caller's semantics)."""
    m.set_precision(prec)
    res = m.batch_match(MIN_SCORE, MIN_RATIO)
    for _ in range(warmup):
        m.batch_match_async(MIN_SCORE, MIN_RATIO)
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.batch_match_async(MIN_SCORE, MIN_RATIO)
    m.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    m.batch_get_matches()
    gpu_ms = m.batch_last_ms()
    t0 = time.perf_counter()
    for _ in range(read_back_steps):
        res = m.batch_match(MIN_SCORE, MIN_RATIO)
    ms_rb = (time.perf_counter() - t0) * 1e3 / read_back_steps
    tuning = m.batch_get_tuning(prec)
    tuning["shape"] = m.batch_get_tile_shape(prec)  # query (plain: left) rows x candidate rows per work-group
    return res, {"ms_per_batch": ms, "pairs_per_s": BATCH_PAIRS / (ms * 1e-3), "gpu_ms_last_run": gpu_ms,
                 "ms_per_batch_read_back": ms_rb, "tuning": tuning}


def cpu_baseline_batch(images, left, right):
    """The CPU port on every pair of the batch once (FP32 blocked GEMM + top-2 + mutual,
    OpenMP over the host threads)."""
    from oracle import oracle
    threads = oracle.cpu_max_threads()
    fl = [None] * len(images)
    for i in set(left.tolist()) | set(right.tolist()):
        n = images[i].shape[0] // 32 * 32
        fl[i] = np.ascontiguousarray(synth_dequant(images[i][:n]))
    oracle.cpu_baseline_find(fl[left[0]][:256], fl[right[0]][:256], MIN_SCORE, MIN_RATIO, threads)  # warm-up
    out = []
    t0 = time.perf_counter()
    for a, b in zip(left, right):
        L, R = fl[a], fl[b]
        out.append(oracle.cpu_baseline_find(L, R, MIN_SCORE, MIN_RATIO, threads)
                   if L.shape[0] and R.shape[0] else np.zeros((0, 2), np.int32))
    dt = time.perf_counter() - t0
    return {"value": len(left) / dt, "unit": "image pairs/s", "cores": threads, "kind": "port",
            "sample": f"the whole {len(left)}-pair batch once ({dt:.1f} s; FP32 blocked GEMM + top-2 + mutual, "
                      f"OpenMP {threads} threads)"}, out


def synth_dequant(q):
    from v3dmatch import synth
    return synth.dequantise(q)


def bench_batch(args):
    """The headline: the 256-pair uneven batch in the plain, BF16 and learned forms."""
    import v3dmatch
    from v3dmatch import learned, synth
    log(f"bench: generating the {BATCH_PAIRS}-pair uneven batch")
    counts, desc, left, right, _ = synth.make_batch(BATCH_PAIRS)
    offs = np.concatenate([[0], np.cumsum(counts)])
    images = [desc[offs[i]:offs[i + 1]] for i in range(len(counts))]
    rc = counts - (counts & 31)
    score_pairs = int(sum(int(rc[a]) * int(rc[b]) for a, b in zip(left, right)))
    m = v3dmatch.SIFT_Matching(0)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    m.set_derived_cache(False)  # every run embeds / converts the images again
    learned.Network(m).load(learned.WEIGHTS_PATH)
    out = {"workload": f"{BATCH_PAIRS} image pairs, 500-16000 features per image (log-uniform), 30-60% noisy "
                       f"copies per pair, 128-d uint8 descriptors, minScore {MIN_SCORE}, minRatio {MIN_RATIO}",
           "pairs": BATCH_PAIRS, "features": int(counts.sum()), "score_pairs": score_pairs}
    log("bench: batch plain_fp32")
    plain, out["plain_fp32"] = time_batch(m, v3dmatch.PREC_FP32, args.steps, args.warmup)
    out["plain_fp32"]["kernel"] = "k_batch_fused_fp32"
    out["plain_fp32"]["kernel_ms"] = m.batch_time_dominant_kernel(3)
    out["plain_fp32"]["flop_per_launch"] = 2.0 * score_pairs * 128
    ex = _sets(plain)
    log("bench: batch bf16_shortlist")
    bf, out["bf16_shortlist"] = time_batch(m, v3dmatch.PREC_BF16, args.steps, args.warmup)
    out["bf16_shortlist"]["kernel"] = "k_batch_pass_bf16"
    out["bf16_shortlist"]["kernel_ms"] = m.batch_time_dominant_kernel(5)
    log("bench: batch learned_shortlist")
    ln, out["learned_shortlist"] = time_batch(m, v3dmatch.PREC_LEARNED, args.steps, args.warmup)
    lf = out["learned_shortlist"]
    flags = m.batch_get_fallback()
    chk = m.batch_get_check()
    lf["kernel"] = "k_batch_pass_code"
    lf["kernel_ms"] = m.batch_time_dominant_kernel(10)
    lf["ops_per_launch"] = 2.0 * score_pairs * 32  # one 32-d int8 product per score pair
    lf["sampled_check"] = {"rule": "fallback to the plain form when the sampled recall of the pair's exact "
                                   "mutual matches is below 0.995",
                           "samples": int(chk[:, 3].sum()), "exact_matches_sampled": int(chk[:, 0].sum()),
                           "missed": int(chk[:, 1].sum()), "decided_exactly": int(chk[:, 2].sum()),
                           "fallback_pairs": int(flags.sum())}
    for name, res in (("bf16_shortlist", bf), ("learned_shortlist", ln)):
        gs = _sets(res)
        inter = [len(ex[p] & gs[p]) for p in range(BATCH_PAIRS)]
        per = [inter[p] / max(len(ex[p]), 1) for p in range(BATCH_PAIRS)]
        o = out[name]
        o["matches"] = sum(len(g) for g in gs)
        o["recall"] = sum(inter) / max(sum(len(e) for e in ex), 1)
        o["precision"] = sum(inter) / max(o["matches"], 1)
        o["min_pair_recall"] = min(per)
        o["pairs_below_0995"] = sum(1 for r in per if r < 0.995)
        o["speedup_vs_plain"] = out["plain_fp32"]["ms_per_batch"] / o["ms_per_batch"]
        o["plain_equivalent_tflops"] = 2.0 * score_pairs * 128 / (o["ms_per_batch"] * 1e-3) / 1e12
    out["plain_fp32"]["matches"] = sum(len(e) for e in ex)
    out["plain_fp32"]["recall_reference"] = "itself (identical to the oracle: tests/test_gpu_batch_full.py)"
    m.close()
    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline over the batch")
        cpu, cpu_sets = cpu_baseline_batch(images, left, right)
        cpu["match_sets_identical_to_gpu_plain"] = all(
            np.array_equal(cpu_sets[p], plain[p]) for p in range(BATCH_PAIRS))
    if not args.no_oracle:
        from oracle import oracle
        log("bench: oracle on the batch")
        same = True
        for p, (a, b) in enumerate(zip(left, right)):
            L = synth.dequantise(images[a][:int(rc[a])])
            R = synth.dequantise(images[b][:int(rc[b])])
            want = oracle.find_putative_matches(L, R, MIN_SCORE, MIN_RATIO) if len(L) and len(R) else \
                np.zeros((0, 2), np.int32)
            same = same and np.array_equal(want, plain[p])
        out["plain_fp32"]["match_sets_identical_to_oracle"] = bool(same)
    return out, cpu


def time_form(m, steps, warmup):
    """ms per step over `steps` back-to-back calls (no event recording in this loop); the
    same number of calls again with HIP events around the dominant kernel inside every call
    (kernel_ms_in_step: includes the events' own dispatch gap); and the dominant kernel
    launched `steps` times back to back between one event pair on the handle's stream
    (kernel_ms: the figure the roofline uses, comparable with a kernel trace)."""
    for _ in range(warmup):
        m.find_async(MIN_SCORE, MIN_RATIO)
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.find_async(MIN_SCORE, MIN_RATIO)
    m.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / steps
    m.set_profiling(True)
    for _ in range(steps):
        m.find_async(MIN_SCORE, MIN_RATIO)
    m.synchronize()
    st = m.stats()
    m.set_profiling(False)
    st["kernel_ms_in_step"] = st["kernelMs"] / max(st["kernelLaunches"], 1)
    kernel_ms = m.time_dominant_kernel(max(steps, 10))
    return ms_step, kernel_ms, st


def time_read_back(m, calls):
    """ms per call for a caller that reads every result before the next call (the
    reference's synchronous findPutativeMatches + match copy): findPutativeMatches with the
    match list copied to host memory, `calls` times."""
    m.find_putative_matches_array(MIN_SCORE, MIN_RATIO)
    t0 = time.perf_counter()
    for _ in range(calls):
        m.find_putative_matches_array(MIN_SCORE, MIN_RATIO)
    return (time.perf_counter() - t0) * 1e3 / calls


def cpu_baseline_pair(L, R, budget_s=10.0, max_reps=20):
    from oracle import oracle
    threads = oracle.cpu_max_threads()
    oracle.cpu_baseline_find(L[:1024], R[:1024], MIN_SCORE, MIN_RATIO, threads)  # warm-up
    reps, t0 = 0, time.perf_counter()
    out = None
    while reps < max_reps and (time.perf_counter() - t0) < budget_s:
        out = oracle.cpu_baseline_find(L, R, MIN_SCORE, MIN_RATIO, threads)
        reps += 1
    dt = time.perf_counter() - t0
    return {"value": reps / dt, "unit": "image pairs/s", "cores": threads, "kind": "port",
            "sample": f"{reps} x the full {L.shape[0]}x{R.shape[0]} pair (FP32 blocked GEMM + top-2 + mutual, "
                      f"OpenMP {threads} threads)"}, out


def bench_pair(args):
    """The single 16384 x 16384 pair in all four forms (device-resident inputs and matches)."""
    import v3dmatch
    from v3dmatch import learned, synth
    log(f"bench: generating the {N_PAIR}x{N_PAIR} pair")
    qL, qR, _ = synth.make_pair(N_PAIR, N_PAIR, 0.4, config=2, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(N_PAIR, N_PAIR)
    m.setLeftFeatures(N_PAIR, L)
    m.setRightFeatures(N_PAIR, R)
    m.set_derived_cache(False)
    flop = 2.0 * N_PAIR * N_PAIR * 128
    log("bench: pair plain FP32")
    ms, kms, st = time_form(m, args.steps, args.warmup)
    exact = m.get_matches()
    forms = {"plain_fp32": {"pairs_per_s": 1e3 / ms, "ms_per_step": ms, "kernel": "k_fused_fp32",
                            "kernel_ms": kms, "kernel_tflops": flop / (kms * 1e-3) / 1e12,
                            "frac": flop / (kms * 1e-3) / PEAK_FP32_MFMA, "matches": int(len(exact)),
                            "rechecked_queries": int(st["nRechecked"]),
                            "ms_per_call_read_back": time_read_back(m, args.steps)}}
    ex = set(map(tuple, exact.tolist()))
    learned.Network(m).load(learned.WEIGHTS_PATH)
    for prec, name, kernel, peak, work in (
            (v3dmatch.PREC_BF16, "bf16_shortlist", "k_pass_bf16", PEAK_BF16_MFMA, flop),
            (v3dmatch.PREC_FP8, "fp8_shortlist", "k_pass_fp8", PEAK_FP8_MFMA, flop),
            (v3dmatch.PREC_LEARNED, "learned_shortlist", "k_pass_code2", PEAK_INT8_MFMA,
             2.0 * N_PAIR * N_PAIR * 32)):
        log(f"bench: pair {name}")
        m.set_precision(prec)
        ms, kms, st = time_form(m, args.steps, args.warmup)
        gs = set(map(tuple, m.get_matches().tolist()))
        inter = len(ex & gs)
        forms[name] = {"pairs_per_s": 1e3 / ms, "ms_per_step": ms, "kernel": kernel, "kernel_ms": kms,
                       "work_per_pair": work, "rate_t": work / (kms * 1e-3) / 1e12,
                       "frac": work / (kms * 1e-3) / peak, "matches": len(gs),
                       "recall": inter / max(len(ex), 1), "precision": inter / max(len(gs), 1),
                       "rescored_candidates": int(st["nRescored"]), "speedup_vs_plain":
                           forms["plain_fp32"]["ms_per_step"] / ms,
                       "ms_per_call_read_back": time_read_back(m, args.steps)}
        if prec == v3dmatch.PREC_LEARNED:
            st = m.stats()
            forms[name]["sampled_check_last_call"] = {"samples": int(st["nSampled"]),
                                                      "missed": int(st["nSampleMissed"]),
                                                      "fallback_to_plain": int(st["nFallback"])}
        m.set_precision(v3dmatch.PREC_FP32)
    out = {"workload": f"one pair of {N_PAIR}x{N_PAIR} descriptors, 40% noisy copies", "forms": forms,
           "note": "a step is one findPutativeMatches with the matches left in HBM; the learned form evaluates "
                   "each call's sampled check (and any fallback) before the next call"}
    if not args.no_oracle:
        from oracle import oracle
        log("bench: pair oracle check")
        out["plain_match_set_identical_to_oracle"] = bool(
            np.array_equal(exact, oracle.find_putative_matches(L, R, MIN_SCORE, MIN_RATIO)))
    m.close()
    return out


def bench_fp8(args):
    """The FP8 form on a 32768 x 32768 pair: time, recall and precision against the plain
    form's (exact) set, and the FP8 pass's fraction of the FP8 MFMA peak."""
    import v3dmatch
    from v3dmatch import synth
    log(f"bench: generating the {N_FP8}x{N_FP8} pair")
    qL, qR, _ = synth.make_pair(N_FP8, N_FP8, 0.4, config=4, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(N_FP8, N_FP8)
    m.setLeftFeatures(N_FP8, L)
    m.setRightFeatures(N_FP8, R)
    m.set_derived_cache(False)
    ms_plain, kms_plain, _ = time_form(m, max(3, args.steps // 4), 2)
    exact = set(map(tuple, m.get_matches().tolist()))
    m.set_precision(v3dmatch.PREC_FP8)
    ms, kms, st = time_form(m, args.steps, args.warmup)
    gs = set(map(tuple, m.get_matches().tolist()))
    m.close()
    inter = len(exact & gs)
    work = 2.0 * N_FP8 * N_FP8 * 128
    return {"workload": f"one pair of {N_FP8}x{N_FP8} descriptors, per-descriptor power-of-two scales, "
                        "40% noisy copies",
            "plain_fp32_ms": ms_plain, "plain_fp32_kernel_ms": kms_plain,
            "fp8_ms_per_step": ms, "pairs_per_s": 1e3 / ms, "speedup_vs_plain": ms_plain / ms,
            "kernel": "k_pass_fp8", "kernel_ms": kms, "work_per_pair": work,
            "tflops": work / (kms * 1e-3) / 1e12, "frac_fp8_peak": work / (kms * 1e-3) / PEAK_FP8_MFMA,
            "matches": len(gs), "exact_matches": len(exact), "recall": inter / max(len(exact), 1),
            "precision": inter / max(len(gs), 1), "rescored_candidates": int(st["nRescored"]),
            "recall_reference": "the plain FP32 form's set (identical to the oracle's)"}


def bench_small_pair(args):
    """The CPU-runnable configuration: one 4096 x 4096 pair, plain form, GPU and CPU."""
    import v3dmatch
    from v3dmatch import synth
    qL, qR, _ = synth.make_pair(N_SMALL, N_SMALL, 0.4, config=0, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(N_SMALL, N_SMALL)
    m.setLeftFeatures(N_SMALL, L)
    m.setRightFeatures(N_SMALL, R)
    ms, kms, _ = time_form(m, args.steps, args.warmup)
    got = m.get_matches()
    rb = time_read_back(m, args.steps)
    m.close()
    flop = 2.0 * N_SMALL * N_SMALL * 128
    out = {"workload": f"one pair of {N_SMALL}x{N_SMALL} descriptors, 40% noisy copies, plain FP32 form",
           "ms_per_step": ms, "pairs_per_s": 1e3 / ms, "ms_per_call_read_back": rb, "kernel": "k_fused_fp32",
           "kernel_ms": kms, "frac_fp32_peak": flop / (kms * 1e-3) / PEAK_FP32_MFMA, "matches": int(len(got))}
    if not args.no_cpu:
        cb, cpu_pairs = cpu_baseline_pair(L, R, budget_s=5.0)
        cb["match_set_identical_to_gpu"] = bool(np.array_equal(cpu_pairs, got))
        out["cpu_baseline"] = cb
    if not args.no_oracle:
        from oracle import oracle
        out["match_set_identical_to_oracle"] = bool(np.array_equal(got, oracle.find_putative_matches(L, R)))
    return out


def bench_score_error():
    """Largest FP64 score error of the FP32, BF16 and FP8 passes on a 1024 x 1024 block of
    the generator's pairs at the three pair sizes: max |s - s64| / (|q| |c|) and, for the
    FP32 pass, max |s - fl32(s64)| / (|q| max|c|) (the certified decisions assume < 2^-15)."""
    import v3dmatch
    from v3dmatch import synth
    out = {}
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(1024, 1024)
    for config, n in ((0, N_SMALL), (2, N_PAIR), (4, N_FP8)):
        qL, qR, _ = synth.make_pair(n, n, 0.4, config=config, pair=3)
        L, R = synth.dequantise(qL[:1024]), synth.dequantise(qR[:1024])
        m.setLeftFeatures(1024, L)
        m.setRightFeatures(1024, R)
        s64 = L.astype(np.float64) @ R.astype(np.float64).T
        nq = np.linalg.norm(L.astype(np.float64), axis=1)[:, None]
        nc = np.linalg.norm(R.astype(np.float64), axis=1)
        row = {}
        for prec, name in ((v3dmatch.PREC_FP32, "fp32"), (v3dmatch.PREC_BF16, "bf16"), (v3dmatch.PREC_FP8, "fp8")):
            s = m.score_block(prec, 1024, 1024).astype(np.float64)
            row[name] = {"max_abs_error": float(np.abs(s - s64).max()),
                         "max_rel_error_q_c": float((np.abs(s - s64) / (nq * nc[None, :])).max())}
            if prec == v3dmatch.PREC_FP32:
                row[name]["max_certified_error"] = float(
                    (np.abs(s - s64.astype(np.float32)) / (nq * nc.max())).max())
        out[f"{n}x{n}"] = row
    m.close()
    return out


class _Hip:
    """Device buffers from the HIP runtime the library is linked with."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]

    def upload(self, a):
        p = self.C.c_void_p()
        assert self.rt.hipMalloc(self.C.byref(p), a.nbytes) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p.value

    def alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.rt.hipMalloc(self.C.byref(p), nbytes) == 0
        return p.value

    def download(self, a, p):
        assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0

    def free(self, p):
        self.rt.hipFree(p)


def synthetic_tree(branch, depth, seed=21, spread=0.35):
    """k-means-like vocabulary tree: every node's children perturb its direction."""
    rng = np.random.default_rng(seed)
    n_nodes = sum(branch ** (d + 1) for d in range(depth))
    tree = np.zeros((n_nodes, 128), np.float32)
    unit = lambda x: (x / np.linalg.norm(x, axis=-1, keepdims=True)).astype(np.float32)  # noqa: E731
    tree[:branch] = unit(np.abs(rng.standard_normal((branch, 128))))
    start, ln = 0, branch
    for _ in range(1, depth):
        par = tree[start:start + ln]
        kids = par[:, None, :] + spread * np.abs(rng.standard_normal((ln, branch, 128))) / np.sqrt(128)
        c0 = (start + 1) * branch
        tree[c0:c0 + ln * branch] = unit(kids).reshape(-1, 128)
        start, ln = start + ln, ln * branch
    leaves = tree[n_nodes - branch ** depth:]
    return tree, leaves, unit


def bench_voctree(steps, warmup, cpu, branch=50, depth=3, n=1_000_000):
    """Vocabulary-tree quantisation (VocabularyTree::getVisualWord for every key): n keys
    near the leaves of a synthetic <branch, depth> tree, keys and words in HBM."""
    from v3dmatch import voctree
    tree, leaves, unit = synthetic_tree(branch, depth)
    rng = np.random.default_rng(22)
    keys = unit(np.abs(leaves[rng.integers(0, leaves.shape[0], n)] + 0.08 * rng.standard_normal((n, 128))))
    hip = _Hip()
    dk, dw = hip.upload(keys), hip.alloc(4 * n)
    res = {"workload": f"{n} keys, <{branch},{depth}> tree ({tree.shape[0]} nodes), keys near leaves"}
    with voctree.VocabularyTree(branch, depth) as t:
        t.setNodes(tree)
        for _ in range(warmup):
            t.traverse_device(n, dk, dw)
        t.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            t.traverse_device(n, dk, dw)
        exact = t.synchronize()
        dt = (time.perf_counter() - t0) / steps
        words = np.empty(n, np.int32)
        hip.download(words, dw)
    hip.free(dk)
    hip.free(dw)
    flop = 2.0 * n * branch * depth * 128
    res.update({"ms_per_run": dt * 1e3, "keys_per_s": n / dt, "exact_keys": int(exact),
                "tflops": flop / dt / 1e12, "frac_fp32_mfma": flop / dt / PEAK_FP32_MFMA})
    if cpu:
        from oracle import oracle
        s = 20000
        t0 = time.perf_counter()
        want = oracle.voctree_leaves(tree, branch, depth, keys[:s])
        cdt = time.perf_counter() - t0
        res["cpu_baseline"] = {"keys_per_s": s / cdt, "cores": oracle.cpu_max_threads(), "kind": "port",
                               "sample": f"first {s} keys, oracle descent (sequential FP32 scores), OpenMP"}
        res["words_identical_to_oracle_on_sample"] = bool(np.array_equal(words[:s], want))
    return res


def bench_ransac(steps, n_in=3300, n_out=1400):
    """The match lists' consumer (computeRobustOrientationMLE, host code, one thread): a
    synthetic two-view scene with about as many correspondences as the 16384-feature pair's
    match list has (3300 inliers with 0.5 px noise, 1400 outliers), 1000 samples, 2 px
    threshold, iterative refinement, seeds 1..steps."""
    from v3dmatch import ransac, synth
    corrs, K, R, T, is_in = synth.two_view_scene(n_in, n_out, seed=11)
    res = None
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        res = ransac.robust_orientation(corrs, K, K, 2.0, 1000, True, True, seed=s)
    dt = (time.perf_counter() - t0) / steps
    got = np.zeros(len(corrs), bool)
    got[res["inliers"]] = True
    Rr = res["rotation"]
    rot = float(np.degrees(np.arccos(np.clip((np.trace(Rr.T @ R) - 1.0) / 2.0, -1.0, 1.0))))
    tdir = float(np.degrees(np.arccos(np.clip(res["translation"] @ T, -1.0, 1.0))))
    return {"workload": f"{n_in + n_out} correspondences ({n_in} inliers, 0.5 px noise, {n_out} outliers), "
                        "1000 samples, 2 px, refinement on; host code, one thread",
            "ms_per_call": dt * 1e3, "samples_drawn": res["samples"], "hypotheses": res["hypotheses"],
            "rotation_error_deg": rot, "direction_error_deg": tdir,
            "inlier_recall": float((got & is_in).sum() / is_in.sum()),
            "inlier_precision": float((got & is_in).sum() / max(got.sum(), 1))}


def bench_guided(L, R, steps, warmup):
    """Guided matching (runGuidedHomographyMatching) on the single pair: copies sit at H p +
    noise, distractors anywhere; each call uploads the positions, masks, rescores the pair
    exactly and returns the match list to the host."""
    import v3dmatch
    from v3dmatch import synth
    _, _, truth = synth.make_pair(N_PAIR, N_PAIR, 0.4, config=2, pair=0)
    rng = np.random.default_rng(5)
    H = np.array([[1.02, 0.03, 12.0], [-0.02, 0.99, -7.0], [1e-5, -2e-5, 1.0]])
    lxy = np.stack([rng.uniform(0, 1024, N_PAIR), rng.uniform(0, 768, N_PAIR)], 1)
    rxy = np.stack([rng.uniform(0, 1024, N_PAIR), rng.uniform(0, 768, N_PAIR)], 1)
    cp = truth >= 0
    hp = (H @ np.c_[lxy[truth[cp]], np.ones(cp.sum())].T).T
    rxy[cp] = hp[:, :2] / hp[:, 2:3] + rng.standard_normal((cp.sum(), 2))
    lxy, rxy = lxy.astype(np.float32), rxy.astype(np.float32)
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(N_PAIR, N_PAIR, True)
    m.setLeftFeatures(N_PAIR, L)
    m.setRightFeatures(N_PAIR, R)
    put = m.find_putative_matches_array(MIN_SCORE, MIN_RATIO)
    for _ in range(warmup):
        m.run_guided_array(True, H, 3.0, lxy, rxy, MIN_SCORE, MIN_RATIO)
    t0 = time.perf_counter()
    for _ in range(steps):
        g = m.run_guided_array(True, H, 3.0, lxy, rxy, MIN_SCORE, MIN_RATIO)
    dt = (time.perf_counter() - t0) / steps
    m.close()
    return {"workload": f"{N_PAIR}x{N_PAIR} pair, homography mask, 3 px", "ms_per_call_host_inclusive": dt * 1e3,
            "calls_per_s": 1.0 / dt, "putative_matches": int(len(put)), "guided_matches": len(g)}


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic.json (rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes scaled by the copy calibration), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-oracle", action="store_true", help="skip the oracle match-set checks")
    ap.add_argument("--only-batch", action="store_true", help="the headline batch only")
    ap.add_argument("--no-extras", action="store_true", help="skip guided matching, tree and robust orientation")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration << 20)

    batch, cpu = bench_batch(args)
    lf = batch["learned_shortlist"]
    log(f"bench: batch ms per batch plain {batch['plain_fp32']['ms_per_batch']:.3f} bf16 "
        f"{batch['bf16_shortlist']['ms_per_batch']:.3f} learned {lf['ms_per_batch']:.3f} "
        f"(x{lf['speedup_vs_plain']:.2f}, recall {lf['recall']:.5f}, fallback pairs "
        f"{lf['sampled_check']['fallback_pairs']})")
    ops = lf["ops_per_launch"]
    achieved = ops / (lf["kernel_ms"] * 1e-3)
    out = {
        "metric": METRIC,
        "value": lf["pairs_per_s"],
        "unit": "image pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": lf["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic (seeded SIFT-like generator, uint8-quantised descriptors)",
        "config": {"workload": batch["workload"] + "; learned shortlist form (128-256-32 network, top-6 codes with "
                                                   "the code-score margin, exact rescoring, sampled check and "
                                                   "per-pair plain fallback)",
                   "pairs": BATCH_PAIRS, "features": batch["features"], "score_pairs": batch["score_pairs"],
                   "form": "learned_shortlist"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_INT8_MFMA / 1e12, "unit": "TOP/s",
                     "frac": achieved / PEAK_INT8_MFMA, "traffic": traffic_for("k_batch_pass_code"),
                     "kernel": "k_batch_pass_code", "kernel_ms": lf["kernel_ms"], "ops_per_launch": ops},
        "speedup_vs_plain": lf["speedup_vs_plain"],
        "recall": lf["recall"],
        "precision": lf["precision"],
        "min_pair_recall": lf["min_pair_recall"],
        "batch": batch,
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if not args.only_batch:
        out["pair_16384"] = bench_pair(args)
        out["fp8_32768"] = bench_fp8(args)
        log("bench: 4096 pair")
        out["pair_4096"] = bench_small_pair(args)
        log("bench: score errors")
        out["score_error"] = bench_score_error()
        if not args.no_extras:
            qL, qR, _ = __import__("v3dmatch").synth.make_pair(N_PAIR, N_PAIR, 0.4, config=2, pair=0)
            from v3dmatch import synth
            log("bench: guided matching")
            out["guided"] = bench_guided(synth.dequantise(qL), synth.dequantise(qR), 10, 2)
            log("bench: vocabulary tree")
            out["voctree"] = bench_voctree(10, 2, not args.no_cpu)
            log("bench: robust orientation (host)")
            out["ransac"] = bench_ransac(5)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
