"""GPU tests of the score error each pass carries against FP64 scores (v3dm_score_block:
the scores of the pass's own instruction, operand rounding and summation order).

* FP32 pass (k_fused_fp32): the certified decisions settle a query from the MFMA scores
  when its decision cannot change within eps = 2^-15 * |q| * max|c| + 1e-30 of the exact
  scores fl32(sum_k q_k c_k) (the oracle's score, V3D/CUDA/v3d_cudamatching.cu:529-530
  computed exactly).  The test measures max |s_mfma - fl32(s_64)| / (|q| * max|c|) and
  requires it below 2^-15: on the generator's descriptors at the three pair sizes, and on
  sets built to stress the bound -- rows scaled by 2^U(-20, 20), heavy-tailed signed
  components, near-cancelling pairs, tiny and huge rows.
* BF16 pass: the shortlist window assumes |s_bf16 - s| <= (2^-8 + 2^-16) * |q| * |c|.
* FP8 pass: its window (2^-5 * |q| * max|c|) is an empirical choice, not a bound (e4m3
  operands err by up to 2^-4 relative each); the measured error is reported, the largest
  error over a block must stay below twice the window and the share of scores beyond the
  window below 10^-4 (a candidate whose score errs by more may be left out of the rescoring;
  the FP8 form's recall is measured in tests/test_gpu_fp8.py).
"""
import sys

import numpy as np
import pytest

import v3dmatch
from v3dmatch import synth

pytestmark = pytest.mark.gpu

EPS_FP32 = 2.0 ** -15
EPS_BF16 = 2.0 ** -8 + 2.0 ** -16
WIN_FP8 = 2.0 ** -5
B = 1024  # block of left x right features measured


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def matcher():
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(B, B)
    yield m
    m.close()


def _errors(m, L, R, mode, window=None):
    """(max |s - fl32(s64)| / (|q| max|c|), max |s - s64| / (|q| |c|)) over the block, and
    with `window` the share of scores whose error exceeds it."""
    m.setLeftFeatures(L.shape[0], L)
    m.setRightFeatures(R.shape[0], R)
    got = m.score_block(mode, L.shape[0], R.shape[0]).astype(np.float64)
    s64 = L.astype(np.float64) @ R.astype(np.float64).T
    ref = s64.astype(np.float32).astype(np.float64)
    nq = np.linalg.norm(L.astype(np.float64), axis=1)
    nc = np.linalg.norm(R.astype(np.float64), axis=1)
    finite = np.isfinite(got) & np.isfinite(ref)
    assert finite.all(), "non-finite scores in a finite case"
    cert = np.abs(got - ref) / (nq[:, None] * nc.max() + 1e-300)
    pair = np.abs(got - s64) / (nq[:, None] * nc[None, :] + 1e-300)
    if window is not None:
        return float(cert.max()), float(pair.max()), float((cert > window).mean())
    return float(cert.max()), float(pair.max())


def _generator(config, n):
    qL, qR, _ = synth.make_pair(n, n, 0.4, config=config, pair=3)
    return synth.dequantise(qL[:B]), synth.dequantise(qR[:B])


def _stress(kind, seed=17):
    rng = np.random.default_rng(seed)
    if kind == "scaled":
        L, R = _generator(1, 4096)
        L = L * np.exp2(rng.uniform(-20, 20, (B, 1))).astype(np.float32)
        R = R * np.exp2(rng.uniform(-20, 20, (B, 1))).astype(np.float32)
    elif kind == "heavy":
        L = (rng.standard_cauchy((B, 128)) * rng.choice([-1, 1], (B, 128))).astype(np.float32)
        R = (rng.standard_cauchy((B, 128)) * rng.choice([-1, 1], (B, 128))).astype(np.float32)
    elif kind == "cancel":
        # right rows nearly opposite to left rows: dot products far below |q| |c|
        L = rng.standard_normal((B, 128)).astype(np.float32)
        R = (-L[rng.permutation(B)] + 1e-3 * rng.standard_normal((B, 128))).astype(np.float32)
    elif kind == "tiny":
        L, R = _generator(1, 4096)
        L, R = L * np.float32(2.0 ** -60), R * np.float32(2.0 ** -50)
    elif kind == "huge":
        L, R = _generator(1, 4096)
        L, R = L * np.float32(2.0 ** 60), R * np.float32(2.0 ** 58)
    else:
        raise ValueError(kind)
    return np.ascontiguousarray(L), np.ascontiguousarray(R)


@pytest.mark.parametrize("config,n", [(0, 4096), (2, 16384), (4, 32768)])
def test_fp32_error_on_generator(matcher, config, n):
    L, R = _generator(config, n)
    cert, pair = _errors(matcher, L, R, v3dmatch.PREC_FP32)
    say(f"case fp32 score error generator {n}^2 block: {cert:.3e} (bound {EPS_FP32:.3e}), per pair {pair:.3e}")
    assert cert < EPS_FP32


@pytest.mark.parametrize("kind", ["scaled", "heavy", "cancel", "tiny", "huge"])
def test_fp32_error_on_stress_sets(matcher, kind):
    L, R = _stress(kind)
    cert, pair = _errors(matcher, L, R, v3dmatch.PREC_FP32)
    say(f"case fp32 score error {kind}: {cert:.3e} (bound {EPS_FP32:.3e}), per pair {pair:.3e}")
    assert cert < EPS_FP32


@pytest.mark.parametrize("config,n", [(0, 4096), (2, 16384)])
def test_bf16_error(matcher, config, n):
    L, R = _generator(config, n)
    cert, pair = _errors(matcher, L, R, v3dmatch.PREC_BF16)
    say(f"case bf16 score error {n}^2 block: per pair {pair:.3e} (bound {EPS_BF16:.3e})")
    assert pair < EPS_BF16
    L, R = _stress("scaled")
    cert, pair = _errors(matcher, L, R, v3dmatch.PREC_BF16)
    say(f"case bf16 score error scaled rows: per pair {pair:.3e}")
    assert pair < EPS_BF16


@pytest.mark.parametrize("config,n", [(0, 4096), (4, 32768)])
def test_fp8_error(matcher, config, n):
    L, R = _generator(config, n)
    cert, pair, beyond = _errors(matcher, L, R, v3dmatch.PREC_FP8, WIN_FP8)
    say(f"case fp8 score error {n}^2 block: {cert:.3e} (window {WIN_FP8:.3e}), per pair {pair:.3e}, "
        f"share beyond the window {beyond:.2e}")
    assert cert < 2 * WIN_FP8 and beyond < 1e-4


def test_score_block_arguments(matcher):
    L, R = _generator(0, 4096)
    matcher.setLeftFeatures(B, L)
    matcher.setRightFeatures(B, R)
    with pytest.raises(v3dmatch.V3DMError):
        matcher.score_block(v3dmatch.PREC_LEARNED, 32, 32)
    with pytest.raises(v3dmatch.V3DMError):
        matcher.score_block(v3dmatch.PREC_FP32, B + 32, 32)
    assert matcher.score_block(v3dmatch.PREC_FP32, 0, 32).shape == (0, 32)
    # a ragged block equals the corner of a larger one
    a = matcher.score_block(v3dmatch.PREC_FP32, 100, 70)
    b = matcher.score_block(v3dmatch.PREC_FP32, 128, 96)
    assert np.array_equal(a, b[:100, :70])
