"""GPU tests of the learned network's training kernels against the oracle
(oracle/learned.py): forward codes, loss and gradients of one minibatch (the oracle
rounding the GEMM operands to bf16 as the MFMA kernels do, and using the hardest in-batch
positives the GPU chose), one Adam step, checkpoint round trip, and that training lowers
the loss."""
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import learned as OL
from v3dmatch import learned, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def setup():
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(2048, 2048)
    pool, tri = learned.make_training_set(m, n_pairs=2, n=2048, config=51)
    say(f"training set: {pool.shape[0]} rows, {tri.shape[0]} triplets")
    net = learned.Network(m)
    params = learned.init_params(7)
    params["b1"] = (0.05 * np.random.default_rng(1).standard_normal(256)).astype(np.float32)
    net.set_params(params)
    net.set_training(pool, tri)
    yield m, net, params, pool, tri
    m.close()


def test_training_labels(setup):
    """Positives are exact matches, negatives the exact second best (score below the
    positive's, above every other candidate checked on a sample)."""
    m, net, params, pool, tri = setup
    assert tri.shape[0] > 500
    a, p, n = pool[tri[:, 0]], pool[tri[:, 1]], pool[tri[:, 2]]
    sp = (a.astype(np.float64) * p).sum(1)
    sn = (a.astype(np.float64) * n).sum(1)
    assert np.all(sp > 0.9) and np.all(sn < sp)


def test_forward_matches_oracle(setup):
    m, net, params, pool, tri = setup
    X = pool[:700]
    z = net.embed(X)
    zo, _ = OL.forward(params, X, bf16=True)
    say(f"forward rel err {_rel(z, zo):.2e}")
    assert _rel(z, zo) < 2e-3


@pytest.mark.parametrize("B", [64, 300, 1024])
def test_gradients_match_oracle(setup, B):
    m, net, params, pool, tri = setup
    idx = np.random.default_rng(B).choice(tri.shape[0], size=B, replace=False).astype(np.int32)
    loss, g, hard = net.gradients(idx)
    Lo, go, hard_o = OL.loss_and_grads(params, pool, tri[idx], bf16=True, hard=hard)
    _, _, hard_exact = OL.loss_and_grads(params, pool, tri[idx], bf16=False)
    agree = float(np.mean(hard == hard_exact))
    say(f"B={B} loss gpu {loss:.6f} oracle {Lo:.6f}; hardest agree {agree:.4f}; "
        + " ".join(f"{k} {_rel(g[k], go[k]):.2e}" for k in g))
    assert abs(loss - Lo) <= 1e-3 * abs(Lo) + 1e-5
    assert agree >= 0.95
    for k in g:
        assert _rel(g[k], go[k]) < 1e-2, k


def test_adam_step_matches_oracle(setup):
    m, net, params, pool, tri = setup
    n = tri.shape[0] // 2
    sub_pool, sub_tri = pool, tri[:n]
    net.set_params(params)
    net.set_training(sub_pool, sub_tri)
    _, g, _ = net.gradients(np.arange(n, dtype=np.int32))
    net.train(1, batch=n, lr=1e-3, seed=3)
    p1, mm, vv, step = net.get_state()
    assert step == 1
    want, st = OL.adam_step(params, g, OL.adam_init(params), lr=1e-3)
    got = learned.unpack(p1)
    for k in got:
        # at step 1 Adam moves each weight by lr * g / (|g| + eps): compare where the
        # gradient is well above the FP32 summation noise of the two gradient evaluations
        sig = np.abs(g[k]) > 1e-5 * np.abs(g[k]).max()
        assert sig.mean() > 0.5, k
        assert np.allclose(got[k][sig], want[k][sig], rtol=1e-5, atol=2e-6), k
    net.set_training(pool, tri)


def test_checkpoint_round_trip_and_training(setup, tmp_path):
    m, net, params, pool, tri = setup
    net.set_params(params)
    losses = net.train(60, batch=256, lr=2e-3, seed=5)
    say(f"loss first {losses[:5].mean():.4f} last {losses[-5:].mean():.4f}")
    assert losses[-5:].mean() < 0.8 * losses[:5].mean()
    path = str(tmp_path / "ck.npz")
    net.set_code_scale(40.0)
    net.save(path)
    p, mm, vv, step = net.get_state()
    other = v3dmatch.SIFT_Matching(0)
    n2 = learned.Network(other).load(path)
    p2, m2, v2, s2 = n2.get_state()
    assert s2 == step and np.array_equal(p, p2) and np.array_equal(mm, m2) and np.array_equal(vv, v2)
    assert n2.code_scale == 40.0
    # the same continuation from the original and from the restored state
    n2.set_training(pool, tri)
    a = net.train(5, batch=256, lr=2e-3, seed=9)
    b = n2.train(5, batch=256, lr=2e-3, seed=9)
    pa, pb = net.get_state()[0], n2.get_state()[0]
    # float atomics order the hardest-negative terms differently between runs: equal up
    # to FP32 rounding
    assert np.allclose(a, b, rtol=1e-4) and _rel(pa, pb) < 1e-3
    other.close()
