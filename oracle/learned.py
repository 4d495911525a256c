"""CPU restatement of the learned shortlist network (test infrastructure only).

The reference has no learned component: this is the oracle of the project's own 128 -> 256
-> 32 ReLU embedding, its contrastive loss and its Adam step, against which the GPU's
BF16 forward / backward kernels are checked.  Only tests, smoke() and bench.py's CPU leg
use it, as the checker.

Model (parameters W1 (256, 128), b1 (256,), W2 (32, 256), b2 (32,)):
    a = W1 x + b1,  h = relu(a),  z = W2 h + b2,  c = z / max(|z|, 1e-12)
Loss of a minibatch of B triplets (anchor, positive, negative descriptor rows), the
positive being the exact matcher's match of the anchor and the negative its second best
(its hardest distractor in descriptor space):
    s_p = c_a . c_p,  s_n = c_a . c_n,  s_h = max_{t' != t} c_a . c_{p_t'}  (hardest in-batch)
    l_t = softplus((s_n - s_p) / tau) + softplus((s_h - s_p) / tau),   L = mean_t l_t
Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on FP32 master weights.

All arithmetic here is float64.  With bf16=True the GEMM operands (x, W1, h, W2, dz, dh)
are rounded to bfloat16 (round to nearest even) first, as the GPU's MFMA operands are, so
that the GPU can be compared at FP32-accumulation tolerance; with bf16=False it is the
exact model, pinned by finite differences in the tests.
"""
import numpy as np

TAU = 0.05
D_IN, D_HID, D_OUT = 128, 256, 32


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    nan = np.isnan(f)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def init_params(seed=0):
    """He-style initialisation (the GPU side is seeded from these arrays)."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((D_HID, D_IN)) * np.sqrt(2.0 / D_IN)
    b1 = np.zeros(D_HID)
    W2 = rng.standard_normal((D_OUT, D_HID)) * np.sqrt(1.0 / D_HID)
    b2 = np.zeros(D_OUT)
    return {k: v.astype(np.float32) for k, v in dict(W1=W1, b1=b1, W2=W2, b2=b2).items()}


def _op(a, bf16):
    return bf16_round(a) if bf16 else np.asarray(a, dtype=np.float64)


def forward(params, X, bf16=False):
    """Returns (z, cache) for descriptor rows X (n, 128)."""
    x = _op(X, bf16)
    W1 = _op(params["W1"], bf16)
    W2 = _op(params["W2"], bf16)
    a = x @ W1.T + np.asarray(params["b1"], np.float64)
    h = np.maximum(a, 0.0)
    hq = _op(h.astype(np.float32), bf16) if bf16 else h
    z = hq @ W2.T + np.asarray(params["b2"], np.float64)
    return z, {"x": x, "a": a, "h": hq, "W1": W1, "W2": W2}


def normalise(z):
    n = np.maximum(np.sqrt((z * z).sum(1)), 1e-12)
    return z / n[:, None], n


def softplus(x):
    return np.logaddexp(0.0, x)


def sigmoid(x):
    return 0.5 * (1.0 + np.tanh(0.5 * x))


def loss_and_code_grads(z, B, tau=TAU, hard=None):
    """z: (3B, 32) codes of [anchors; positives; negatives].  Returns (L, dz (3B, 32), hard
    indices).  `hard` fixes the in-batch hardest positive of each anchor (else argmax)."""
    c, n = normalise(z)
    ca, cp, cn = c[:B], c[B:2 * B], c[2 * B:3 * B]
    sp = (ca * cp).sum(1)
    sn = (ca * cn).sum(1)
    S = ca @ cp.T
    np.fill_diagonal(S, -np.inf)
    if hard is None:
        hard = np.argmax(S, axis=1)
    sh = S[np.arange(B), hard]
    L = (softplus((sn - sp) / tau) + softplus((sh - sp) / tau)).mean()
    gn = sigmoid((sn - sp) / tau) / (tau * B)
    gh = sigmoid((sh - sp) / tau) / (tau * B)
    dc = np.zeros_like(c)
    dc[:B] = gn[:, None] * (cn - cp) + gh[:, None] * (cp[hard] - cp)
    dc[B:2 * B] = -(gn + gh)[:, None] * ca
    dc[2 * B:] = gn[:, None] * ca
    np.add.at(dc, B + hard, gh[:, None] * ca)
    dz = (dc - c * (c * dc).sum(1, keepdims=True)) / n[:, None]
    return L, dz, hard


def backward(cache, dz, bf16=False):
    """Parameter gradients from code gradients dz (n, 32)."""
    dzq = _op(dz.astype(np.float32), bf16) if bf16 else dz
    h = cache["h"]
    dW2 = dzq.T @ h
    db2 = dz.sum(0)
    dh = (dzq @ cache["W2"]) * (cache["a"] > 0)
    dhq = _op(dh.astype(np.float32), bf16) if bf16 else dh
    dW1 = dhq.T @ cache["x"]
    db1 = dh.sum(0)
    return {"W1": dW1, "b1": db1, "W2": dW2, "b2": db2}


def loss_and_grads(params, pool, triplets, tau=TAU, bf16=False, hard=None):
    """Loss and parameter gradients of a minibatch; triplets (B, 3) rows of pool."""
    B = triplets.shape[0]
    X = np.concatenate([pool[triplets[:, 0]], pool[triplets[:, 1]], pool[triplets[:, 2]]])
    z, cache = forward(params, X, bf16)
    L, dz, hard = loss_and_code_grads(z, B, tau, hard)
    return L, backward(cache, dz, bf16), hard


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float64 on copies; state = {"t": int, "m": {...}, "v": {...}}."""
    t = state["t"] + 1
    out, m2, v2 = {}, {}, {}
    for k in params:
        g = np.asarray(grads[k], np.float64)
        m = beta1 * state["m"][k] + (1 - beta1) * g
        v = beta2 * state["v"][k] + (1 - beta2) * g * g
        mh = m / (1 - beta1 ** t)
        vh = v / (1 - beta2 ** t)
        out[k] = np.asarray(params[k], np.float64) - lr * mh / (np.sqrt(vh) + eps)
        m2[k], v2[k] = m, v
    return out, {"t": t, "m": m2, "v": v2}


def adam_init(params):
    return {"t": 0, "m": {k: np.zeros(v.shape) for k, v in params.items()},
            "v": {k: np.zeros(v.shape) for k, v in params.items()}}
