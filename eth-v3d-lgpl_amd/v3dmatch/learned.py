"""Learned shortlist network: host side of training, checkpoints and code calibration.

The network (128 -> 256 -> 32 ReLU MLP, packed FP32 parameter vector of
V3DM_LEARNED_PARAMS floats) is trained on the GPU by the library's own BF16 forward /
backward kernels and Adam (v3dm_learned_train).  Training triplets (anchor, positive,
negative) come from synthetic image pairs labelled by the exact matcher: the positive of a
left feature is its exact mutual match, the negative the right feature with the second
best exact score (its hardest distractor).  Checkpoints are .npz files (parameters, Adam
moments, step, code scale) read without pickle.
"""
import ctypes as C
import os

import numpy as np

from ._native import load_library
from . import synth

N_PARAMS = 41248
OFF_W1, OFF_B1, OFF_W2, OFF_B2 = 0, 32768, 33024, 41216
DEFAULT_TAU = 0.05
CODE_LEVELS = 31
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "shortlist_mlp.npz")


def pack(params):
    """dict(W1, b1, W2, b2) -> packed float32 vector."""
    v = np.empty(N_PARAMS, dtype=np.float32)
    v[OFF_W1:OFF_B1] = np.asarray(params["W1"], np.float32).reshape(-1)
    v[OFF_B1:OFF_W2] = np.asarray(params["b1"], np.float32).reshape(-1)
    v[OFF_W2:OFF_B2] = np.asarray(params["W2"], np.float32).reshape(-1)
    v[OFF_B2:] = np.asarray(params["b2"], np.float32).reshape(-1)
    return v


def unpack(v):
    v = np.asarray(v, np.float32)
    return {"W1": v[OFF_W1:OFF_B1].reshape(256, 128).copy(), "b1": v[OFF_B1:OFF_W2].copy(),
            "W2": v[OFF_W2:OFF_B2].reshape(32, 256).copy(), "b2": v[OFF_B2:].copy()}


def init_params(seed=0):
    """He-style initial parameters (float32 dict)."""
    rng = np.random.default_rng(seed)
    return {"W1": (rng.standard_normal((256, 128)) * np.sqrt(2.0 / 128)).astype(np.float32),
            "b1": np.zeros(256, np.float32),
            "W2": (rng.standard_normal((32, 256)) * np.sqrt(1.0 / 256)).astype(np.float32),
            "b2": np.zeros(32, np.float32)}


class Network:
    """The learned network of one matcher handle (v3dmatch.SIFT_Matching)."""

    def __init__(self, matcher):
        self.m = matcher
        self._lib = load_library()
        self.code_scale = 0.0

    def _check(self, rc):
        self.m._check(rc)

    # -- state -----------------------------------------------------------------------
    def set_params(self, params, m=None, v=None, step=0):
        p = pack(params) if isinstance(params, dict) else np.ascontiguousarray(params, np.float32)
        mm = None if m is None else np.ascontiguousarray(m, np.float32)
        vv = None if v is None else np.ascontiguousarray(v, np.float32)
        self._check(self._lib.v3dm_learned_set_state(self.m._h, p.ctypes.data,
                                                      None if mm is None else mm.ctypes.data,
                                                      None if vv is None else vv.ctypes.data, int(step)))

    def get_state(self):
        p = np.empty(N_PARAMS, np.float32)
        m = np.empty(N_PARAMS, np.float32)
        v = np.empty(N_PARAMS, np.float32)
        step = C.c_int(0)
        self._check(self._lib.v3dm_learned_get_state(self.m._h, p.ctypes.data, m.ctypes.data, v.ctypes.data,
                                                      C.byref(step)))
        return p, m, v, step.value

    def set_code_scale(self, scale):
        self._check(self._lib.v3dm_learned_set_code_scale(self.m._h, C.c_float(scale)))
        self.code_scale = float(scale)

    def save(self, path):
        p, m, v, step = self.get_state()
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=p, adam_m=m, adam_v=v, step=np.int64(step),
                 code_scale=np.float32(self.code_scale))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"], z["adam_m"], z["adam_v"], int(z["step"]))
            scale = float(z["code_scale"])
        if scale > 0:
            self.set_code_scale(scale)
        return self

    # -- training --------------------------------------------------------------------
    def set_training(self, pool, triplets):
        pool = np.ascontiguousarray(pool, np.float32).reshape(-1, 128)
        tri = np.ascontiguousarray(triplets, np.int32).reshape(-1, 3)
        self._check(self._lib.v3dm_learned_set_training(self.m._h, pool.shape[0], pool.ctypes.data, tri.shape[0],
                                                         tri.ctypes.data))
        self.n_triplets = tri.shape[0]

    def train(self, steps, batch=4096, lr=1e-3, tau=DEFAULT_TAU, seed=1):
        losses = np.empty(max(steps, 1), np.float32)
        self._check(self._lib.v3dm_learned_train(self.m._h, int(steps), int(batch), C.c_float(lr), C.c_float(tau),
                                                  C.c_uint64(seed), losses.ctypes.data))
        return losses[:steps]

    def gradients(self, idx, tau=DEFAULT_TAU):
        idx = np.ascontiguousarray(idx, np.int32)
        loss = C.c_float(0)
        g = np.empty(N_PARAMS, np.float32)
        hard = np.empty(idx.shape[0], np.int32)
        self._check(self._lib.v3dm_learned_gradients(self.m._h, idx.shape[0], idx.ctypes.data, C.c_float(tau),
                                                      C.byref(loss), g.ctypes.data, hard.ctypes.data))
        return loss.value, unpack(g), hard

    def embed(self, desc):
        d = np.ascontiguousarray(desc, np.float32).reshape(-1, 128)
        out = np.empty((d.shape[0], 32), np.float32)
        self._check(self._lib.v3dm_learned_embed(self.m._h, d.shape[0], d.ctypes.data, out.ctypes.data))
        return out

    def calibrate_code_scale(self, desc, quantile=0.9995):
        """Scale so that the given quantile of |code component| maps to CODE_LEVELS."""
        z = self.embed(desc)
        c = z / np.maximum(np.linalg.norm(z, axis=1, keepdims=True), 1e-12)
        q = float(np.quantile(np.abs(c), quantile))
        self.set_code_scale(CODE_LEVELS / max(q, 1e-6))
        return self.code_scale


def make_training_set(matcher, n_pairs, n=16384, copy_frac=0.4, config=5, min_score=0.9, min_ratio=0.8,
                      max_triplets=None):
    """Triplets from synthetic pairs labelled by the exact matcher.

    For pair p (seeded by config / p), every exact match (i, j) gives the triplet
    (L_i, R_j, R_s) with s the right feature of second best exact score for left i
    (rescored exactly in the BF16 shortlist form).  Returns (pool float32 (rows, 128),
    triplets int32 (T, 3)).  The matcher must be allocated for n x n."""
    import v3dmatch
    pools, tris = [], []
    rows = 0
    total = 0
    for p in range(n_pairs):
        qL, qR, _ = synth.make_pair(n, n, copy_frac, config=config, pair=p)
        L, R = synth.dequantise(qL), synth.dequantise(qR)
        matcher.set_precision(v3dmatch.PREC_FP32)
        matcher.setLeftFeatures(n, L)
        matcher.setRightFeatures(n, R)
        pairs = matcher.find_putative_matches_array(min_score, min_ratio)
        matcher.set_precision(v3dmatch.PREC_BF16)
        matcher.find_putative_matches_array(min_score, min_ratio)
        second_l, _ = matcher.get_second_best()
        i, j = pairs[:, 0], pairs[:, 1]
        s = second_l[i]
        keep = (s >= 0) & (s != j)
        i, j, s = i[keep], j[keep], s[keep]
        pools.append(L)
        pools.append(R)
        tris.append(np.stack([rows + i, rows + n + j, rows + n + s], axis=1))
        rows += 2 * n
        total += len(i)
        if max_triplets and total >= max_triplets:
            break
    matcher.set_precision(v3dmatch.PREC_FP32)
    tri = np.concatenate(tris).astype(np.int32)
    if max_triplets:
        tri = tri[:max_triplets]
    return np.concatenate(pools), tri


def train_main(argv=None):
    """Train the shortlist network on the GPU and write a checkpoint.

    python -m v3dmatch.learned --pairs 220 --n 16384 --epochs 6 --out <path>
    """
    import argparse
    import sys
    import time

    import v3dmatch

    ap = argparse.ArgumentParser(description=train_main.__doc__)
    ap.add_argument("--pairs", type=int, default=220)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--max-triplets", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=float, default=6.0)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--tau", type=float, default=DEFAULT_TAU)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--resume", default=None, help="checkpoint to continue from")
    ap.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args(argv)

    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(a.n, a.n)
    t0 = time.time()
    pool, tri = make_training_set(m, a.pairs, n=a.n, max_triplets=a.max_triplets)
    print(f"training set: {tri.shape[0]} triplets, {pool.shape[0]} rows ({time.time() - t0:.1f} s)", flush=True)
    net = Network(m)
    if a.resume:
        net.load(a.resume)
    else:
        net.set_params(init_params(a.seed))
    net.set_training(pool, tri)
    steps = int(a.epochs * tri.shape[0] / a.batch)
    t0 = time.time()
    losses = []
    chunk = max(1, steps // 10)
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        lr = a.lr * (0.1 ** (done / max(steps, 1)))  # exponential decay over the run
        losses.extend(net.train(k, batch=a.batch, lr=lr, tau=a.tau, seed=a.seed + done).tolist())
        done += k
        print(f"step {done}/{steps} loss {np.mean(losses[-k:]):.5f}", flush=True)
    print(f"trained {steps} steps in {time.time() - t0:.1f} s", flush=True)
    # code scale from held-out descriptors (a fresh pair of the same generator)
    qL, qR, _ = synth.make_pair(8192, 8192, 0.4, config=6, pair=0)
    scale = net.calibrate_code_scale(np.concatenate([synth.dequantise(qL), synth.dequantise(qR)]))
    print(f"code scale {scale:.3f}", flush=True)
    net.save(a.out)
    print(f"wrote {a.out}", flush=True)
    m.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(train_main())
