"""CPU pins of the learned network's oracle (oracle/learned.py): gradients against finite
differences of the float64 loss, Adam against its closed form at step 1, bf16 rounding
against known values.  A planted sign error in the code gradient must be caught."""
import numpy as np
import pytest

from oracle import learned as OL


def _setup(B=6, seed=3):
    rng = np.random.default_rng(seed)
    params = OL.init_params(seed)
    params["b1"] = (0.1 * rng.standard_normal(256)).astype(np.float32)
    params["b2"] = (0.1 * rng.standard_normal(32)).astype(np.float32)
    pool = np.abs(rng.standard_normal((3 * B + 5, 128))).astype(np.float32)
    pool /= np.linalg.norm(pool, axis=1, keepdims=True)
    tri = rng.permutation(pool.shape[0])[: 3 * B].reshape(B, 3).astype(np.int64)
    return params, pool, tri


def _loss(params, pool, tri, hard):
    B = tri.shape[0]
    X = np.concatenate([pool[tri[:, 0]], pool[tri[:, 1]], pool[tri[:, 2]]])
    z, _ = OL.forward(params, X)
    return OL.loss_and_code_grads(z, B, hard=hard)[0]


@pytest.mark.parametrize("name,count", [("W1", 12), ("b1", 6), ("W2", 12), ("b2", 6)])
def test_gradients_match_finite_differences(name, count):
    params, pool, tri = _setup()
    p64 = {k: v.astype(np.float64) for k, v in params.items()}
    L, grads, hard = OL.loss_and_grads(p64, pool, tri)
    rng = np.random.default_rng(11)
    flat = p64[name].reshape(-1)
    for idx in rng.choice(flat.size, size=count, replace=False):
        h = 1e-6
        old = flat[idx]
        flat[idx] = old + h
        lp = _loss(p64, pool, tri, hard)
        flat[idx] = old - h
        lm = _loss(p64, pool, tri, hard)
        flat[idx] = old
        num = (lp - lm) / (2 * h)
        ana = grads[name].reshape(-1)[idx]
        assert abs(num - ana) <= 1e-6 + 1e-4 * abs(num), (name, idx, num, ana)


def test_planted_sign_error_is_caught():
    """A code gradient with the positive term's sign flipped disagrees with finite
    differences (the check above would fail on it)."""
    params, pool, tri = _setup()
    p64 = {k: v.astype(np.float64) for k, v in params.items()}
    B = tri.shape[0]
    X = np.concatenate([pool[tri[:, 0]], pool[tri[:, 1]], pool[tri[:, 2]]])
    z, cache = OL.forward(p64, X)
    L, dz, hard = OL.loss_and_code_grads(z, B)
    bad = dz.copy()
    bad[B:2 * B] *= -1.0
    g_bad = OL.backward(cache, bad)["b2"]
    g_ok = OL.backward(cache, dz)["b2"]
    h = 1e-6
    num = np.empty(32)
    for k in range(32):
        p64["b2"][k] += h
        lp = _loss(p64, pool, tri, hard)
        p64["b2"][k] -= 2 * h
        lm = _loss(p64, pool, tri, hard)
        p64["b2"][k] += h
        num[k] = (lp - lm) / (2 * h)
    assert np.allclose(num, g_ok, rtol=1e-4, atol=1e-7)
    assert not np.allclose(num, g_bad, rtol=1e-2, atol=1e-5)


def test_adam_first_step_closed_form():
    params, pool, tri = _setup()
    L, grads, _ = OL.loss_and_grads(params, pool, tri)
    state = OL.adam_init(params)
    new, st = OL.adam_step(params, grads, state, lr=1e-3)
    assert st["t"] == 1
    for k in params:
        g = grads[k]
        want = params[k].astype(np.float64) - 1e-3 * g / (np.abs(g) + 1e-8)
        assert np.allclose(new[k], want, rtol=0, atol=1e-12)


def test_bf16_rounding_known_values():
    one = np.float32(1.0)
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7 + 2 ** -9, -2.5, 3.0e38, np.nan],
                    dtype=np.float32)
    got = OL.bf16_round(vals)
    # the bf16 ulp at 1 is 2^-7: 1 + 2^-8 ties to even (1.0), 1 + 3*2^-8 ties to even
    # (1 + 2^-6), 1 + 2^-7 + 2^-9 is a quarter ulp above 1 + 2^-7 (rounds down)
    assert got[0] == one and got[1] == one
    assert got[2] == 1.0 + 2 ** -6
    assert got[3] == 1.0 + 2 ** -7
    assert got[4] == -2.5
    assert np.isnan(got[6])
