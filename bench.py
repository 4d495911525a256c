"""Benchmark of the gfx950 SIFT putative matcher (one GPU).

Workload: one image pair of 16384 left x 16384 right SIFT-like descriptors (128-d FP32,
synthetic: Gamma(0.5,1) components, 0.2 clip, uint8 quantisation, 40% noisy copies),
minScore 0.9, minRatio 0.8.  A step is one findPutativeMatches on the pair with the
descriptors resident in HBM and the match list left in HBM.  The plain FP32 form and the
BF16, FP8 and learned-code shortlist forms are timed; the plain form's match set is
checked against the exact CPU oracle once per run, and the shortlist forms' recall and
precision are measured against it.  The headline is the fastest of the learned and BF16
forms whose recall and precision are at least 99.5%.

Also reported (not the headline): the 256-pair uneven batch, guided matching on the same
pair, and vocabulary-tree quantisation of 10^6 keys on a <50, 3> tree.  The HBM traffic of
the headline kernel comes from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in
profiles/traffic.json (null when that file has no entry for the kernel).

Prints one JSON line (see the keys below).  --gpus above 1 is refused: this build runs
on one GPU.
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
N_FEAT = 16384
MIN_SCORE, MIN_RATIO = 0.9, 0.8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def cpu_baseline(L, R, budget_s=10.0, max_reps=20):
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


def bench_batch(steps, warmup, n_pairs=256):
    """The uneven batch: n_pairs image pairs, 500-16000 features per image (log-uniform),
    30-60% copies, uint8 descriptors; plain FP32 and BF16 shortlist forms through the batch
    API (persistent kernels, graph replay).  Recall of the BF16 form against the plain
    form's match set (the plain batch equals the single-pair path pair by pair)."""
    import v3dmatch
    from v3dmatch import synth

    log(f"bench: generating the {n_pairs}-pair uneven batch")
    counts, desc, left, right, _ = synth.make_batch(n_pairs)
    m = v3dmatch.SIFT_Matching(0)
    m.batch_set_images(counts, desc)
    m.batch_set_pairs(left, right)
    res = {"workload": f"{n_pairs} pairs, 500-16000 features per image (log-uniform), 30-60% copies, uint8",
           "features": int(counts.sum()),
           "score_pairs": int(sum(int(counts[a] // 32 * 32) * int(counts[b] // 32 * 32)
                                  for a, b in zip(left, right)))}
    m.set_derived_cache(False)  # each run derives its bf16 / code copies of the images
    forms = [("plain_fp32", v3dmatch.PREC_FP32), ("bf16_shortlist", v3dmatch.PREC_BF16)]
    from v3dmatch import learned
    if os.path.exists(learned.WEIGHTS_PATH):
        learned.Network(m).load(learned.WEIGHTS_PATH)
        forms.append(("learned_shortlist", v3dmatch.PREC_LEARNED))
    sets = {}
    for name, prec in forms:
        log(f"bench: batch {name}")
        m.set_precision(prec)
        m.batch_match(MIN_SCORE, MIN_RATIO)  # first run: autotuning and graph capture
        for _ in range(warmup):
            m.batch_match_async(MIN_SCORE, MIN_RATIO)
        m.synchronize()
        t0 = time.perf_counter()
        gpu_ms = 0.0
        for _ in range(steps):
            m.batch_match_async(MIN_SCORE, MIN_RATIO)
            offs, pairs = m.batch_get_matches()
            gpu_ms += m.batch_last_ms()
        dt = (time.perf_counter() - t0) / steps
        sets[name] = set((p, a, b) for p in range(n_pairs) for a, b in pairs[offs[p]:offs[p + 1]].tolist())
        res[name] = {"ms_per_batch": dt * 1e3, "gpu_ms_per_batch": gpu_ms / steps,
                     "pairs_per_s": n_pairs / dt, "matches": len(sets[name]),
                     "tuning": m.batch_get_tuning(prec)}
    ex = sets["plain_fp32"]
    for name, _ in forms[1:]:
        gs = sets[name]
        inter = len(ex & gs)
        res[name]["recall"] = inter / max(len(ex), 1)
        res[name]["precision"] = inter / max(len(gs), 1)
        res[name]["speedup_vs_plain"] = res["plain_fp32"]["ms_per_batch"] / res[name]["ms_per_batch"]
    m.close()
    return res


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
    _, _, truth = synth.make_pair(N_FEAT, N_FEAT, 0.4, config=2, pair=0)
    rng = np.random.default_rng(5)
    H = np.array([[1.02, 0.03, 12.0], [-0.02, 0.99, -7.0], [1e-5, -2e-5, 1.0]])
    lxy = np.stack([rng.uniform(0, 1024, N_FEAT), rng.uniform(0, 768, N_FEAT)], 1)
    rxy = np.stack([rng.uniform(0, 1024, N_FEAT), rng.uniform(0, 768, N_FEAT)], 1)
    cp = truth >= 0
    hp = (H @ np.c_[lxy[truth[cp]], np.ones(cp.sum())].T).T
    rxy[cp] = hp[:, :2] / hp[:, 2:3] + rng.standard_normal((cp.sum(), 2))
    lxy, rxy = lxy.astype(np.float32), rxy.astype(np.float32)
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(N_FEAT, N_FEAT, True)
    m.setLeftFeatures(N_FEAT, L)
    m.setRightFeatures(N_FEAT, R)
    put = m.find_putative_matches_array(MIN_SCORE, MIN_RATIO)
    for _ in range(warmup):
        m.run_guided_array(True, H, 3.0, lxy, rxy, MIN_SCORE, MIN_RATIO)
    t0 = time.perf_counter()
    for _ in range(steps):
        g = m.run_guided_array(True, H, 3.0, lxy, rxy, MIN_SCORE, MIN_RATIO)
    dt = (time.perf_counter() - t0) / steps
    m.close()
    return {"workload": f"{N_FEAT}x{N_FEAT} pair, homography mask, 3 px", "ms_per_call_host_inclusive": dt * 1e3,
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
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-oracle", action="store_true", help="skip the oracle match-set check")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch measurement")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-extras", action="store_true", help="skip the guided and vocabulary-tree measurements")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import v3dmatch
    from v3dmatch import synth

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration << 20)

    log(f"bench: generating {N_FEAT}x{N_FEAT} pair")
    qL, qR, _ = synth.make_pair(N_FEAT, N_FEAT, 0.4, config=2, pair=0)
    L, R = synth.dequantise(qL), synth.dequantise(qR)

    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(N_FEAT, N_FEAT)
    m.setLeftFeatures(N_FEAT, L)
    m.setRightFeatures(N_FEAT, R)
    # every timed call derives its bf16 / fp8 / learned-code copies again
    m.set_derived_cache(False)

    log("bench: timing plain FP32 form")
    ms_plain, kms_plain, st_plain = time_form(m, args.steps, args.warmup)
    exact = m.get_matches()
    flop = 2.0 * N_FEAT * N_FEAT * 128
    forms = {"plain_fp32": {"pairs_per_s": 1e3 / ms_plain, "ms_per_step": ms_plain,
                            "kernel": "k_fused_fp32", "kernel_ms": kms_plain,
                            "kernel_ms_in_step": st_plain["kernel_ms_in_step"],
                            "kernel_tflops": flop / (kms_plain * 1e-3) / 1e12,
                            "matches": int(len(exact)), "rechecked_queries": int(st_plain["nRechecked"]),
                            "ms_per_call_read_back": time_read_back(m, args.steps)}}

    ex = set(map(tuple, exact.tolist()))

    def time_shortlist(prec, name, kernel, form_flop, dtype, peak):
        log(f"bench: timing {name} form")
        m.set_precision(prec)
        ms, kms, st = time_form(m, args.steps, args.warmup)
        gs = set(map(tuple, m.get_matches().tolist()))
        m.set_precision(v3dmatch.PREC_FP32)
        inter = len(ex & gs)
        forms[name] = {"pairs_per_s": 1e3 / ms, "ms_per_step": ms, "kernel": kernel, "kernel_ms": kms,
                       "kernel_ms_in_step": st["kernel_ms_in_step"],
                       "kernel_tflops": form_flop / (kms * 1e-3) / 1e12, "flop_per_launch": form_flop,
                       "dtype": dtype, "peak": peak, "matches": len(gs),
                       "recall": inter / max(len(ex), 1), "precision": inter / max(len(gs), 1),
                       "rescored_candidates": int(st["nRescored"]), "speedup_vs_plain": ms_plain / ms}
        m.set_precision(prec)
        forms[name]["ms_per_call_read_back"] = time_read_back(m, args.steps)
        m.set_precision(v3dmatch.PREC_FP32)
        if prec == v3dmatch.PREC_LEARNED:
            forms[name]["sampled_check"] = {"samples": int(st["nSampled"]), "missed": int(st["nSampleMissed"]),
                                            "fallback_to_plain": int(st["nFallback"])}

    # one k_pass_bf16 / k_pass_fp8 launch covers both directions
    time_shortlist(v3dmatch.PREC_BF16, "bf16_shortlist", "k_pass_bf16", 2 * flop, "bf16", PEAK_BF16_MFMA)
    time_shortlist(v3dmatch.PREC_FP8, "fp8_shortlist", "k_pass_fp8", 2 * flop, "fp8", PEAK_FP8_MFMA)
    from v3dmatch import learned
    if os.path.exists(learned.WEIGHTS_PATH):
        learned.Network(m).load(learned.WEIGHTS_PATH)
        # one k_pass_code2 launch covers both directions
        time_shortlist(v3dmatch.PREC_LEARNED, "learned_shortlist", "k_pass_code2",
                       2 * 2.0 * N_FEAT * N_FEAT * 32, "int8", PEAK_INT8_MFMA)
    forms["plain_fp32"].update({"flop_per_launch": flop, "dtype": "fp32", "peak": PEAK_FP32_MFMA})

    # headline: the fastest of the learned and BF16 shortlist forms that keeps >= 99.5%
    # recall and precision of the exact match set; else the plain form
    ok = [k for k in ("learned_shortlist", "bf16_shortlist")
          if k in forms and forms[k]["recall"] >= 0.995 and forms[k]["precision"] >= 0.995]
    head = min(ok, key=lambda k: forms[k]["ms_per_step"]) if ok else "plain_fp32"
    hf = forms[head]
    peak = hf["peak"]
    achieved = hf["flop_per_launch"] / (hf["kernel_ms"] * 1e-3)
    out = {
        "metric": "image pairs matched/s @>=99.5% match recall, learned/BF16 vs plain FP32",
        "value": hf["pairs_per_s"],
        "unit": "image pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hf["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": hf["dtype"],
        "data": "synthetic (seeded SIFT-like generator, uint8-quantised descriptors)",
        "config": {"workload": f"single pair {N_FEAT}x{N_FEAT} descriptors, 128-d, 40% noisy copies, "
                               f"minScore {MIN_SCORE}, minRatio {MIN_RATIO}; headline form {head}",
                   "n_left": N_FEAT, "n_right": N_FEAT, "form": head},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TOP/s" if hf["dtype"] == "int8" else "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic_for(hf["kernel"]), "kernel": hf["kernel"], "kernel_ms": hf["kernel_ms"],
                     "kernel_ms_in_step": hf["kernel_ms_in_step"],
                     "flop_per_launch": hf["flop_per_launch"]},
        "forms": forms,
    }
    gpu_pairs = exact
    if not args.no_oracle:
        from oracle import oracle
        log("bench: oracle match-set check")
        want = oracle.find_putative_matches(L, R, MIN_SCORE, MIN_RATIO)
        forms["plain_fp32"]["match_set_identical_to_oracle"] = bool(np.array_equal(gpu_pairs, want))
        for name in [k for k in forms if k != "plain_fp32"]:
            forms[name]["recall_reference"] = "plain_fp32 match set (identical to the oracle)" \
                if forms["plain_fp32"]["match_set_identical_to_oracle"] else "plain_fp32 match set"
    if not args.no_cpu:
        log("bench: CPU baseline leg")
        cb, cpu_pairs = cpu_baseline(L, R)
        out["cpu_baseline"] = cb
        if not args.no_oracle:
            out["cpu_baseline"]["match_set_identical_to_oracle"] = bool(np.array_equal(cpu_pairs, want))
    m.close()
    if not args.no_batch:
        out["batch"] = bench_batch(args.batch_steps, 2)
    if not args.no_extras:
        log("bench: guided matching")
        out["guided"] = bench_guided(L, R, 10, 2)
        log("bench: vocabulary tree")
        out["voctree"] = bench_voctree(10, 2, not args.no_cpu)
        log("bench: robust orientation (host)")
        out["ransac"] = bench_ransac(5)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
