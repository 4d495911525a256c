"""The synthetic descriptor generator: dequantisation, determinism and the properties
the workloads rely on (CPU only)."""
import numpy as np

from v3dmatch import synth


def np_dequantise(q):
    """SIFT_Feature::normalizeDescriptor (V3D/Base/v3d_feature.h:30-38) in numpy float32."""
    q = q.astype(np.float32)
    out = np.empty_like(q)
    for i in range(q.shape[0]):
        var = np.float32(0.0)
        for k in range(128):
            var = np.float32(var + q[i, k] * q[i, k])
        denom = np.float32(np.float32(1.0) / np.sqrt(var))
        out[i] = q[i] * denom
    return out


def test_dequantise_bit_exact():
    qL, qR, _ = synth.make_pair(64, 64, 0.4, config=0, pair=1)
    assert np.array_equal(synth.dequantise(qL).view(np.uint32), np_dequantise(qL).view(np.uint32))
    assert np.array_equal(synth.dequantise(qR).view(np.uint32), np_dequantise(qR).view(np.uint32))


def test_deterministic_and_seed_sensitive():
    a = synth.make_pair(300, 200, 0.4, config=1, pair=5)
    b = synth.make_pair(300, 200, 0.4, config=1, pair=5)
    c = synth.make_pair(300, 200, 0.4, config=1, pair=6)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    assert synth.seed(1, 2, 3, 0) != synth.seed(1, 2, 3, 1)


def test_descriptor_properties():
    qL, qR, truth = synth.make_pair(2000, 1500, 0.4, config=1, pair=2)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    assert np.allclose(np.linalg.norm(L, axis=1), 1.0, atol=1e-6)
    assert qL.max() <= 255 and qL.min() >= 0
    ncopy = int((truth >= 0).sum())
    assert ncopy == round(0.4 * 1500)
    src = truth[truth >= 0]
    assert len(np.unique(src)) == ncopy  # distinct sources
    cos = np.sum(L[src] * R[truth >= 0], axis=1)
    assert cos.mean() > 0.85
    # fresh distractors are far less similar to their best left descriptor than copies
    dis = R[truth < 0][:200] @ L.T
    assert dis.max(axis=1).mean() < cos.mean() - 0.1


def test_gamma_shape():
    q = synth.fresh(synth.seed(7, 0, 0, 0), 4000).astype(np.float64)
    # clipped, renormalised Gamma(0.5) components: many small values, a few large ones
    x = q / np.linalg.norm(q, axis=1, keepdims=True)
    assert 0.02 < np.median(x) < 0.05
    assert np.mean(x > 0.2) < 0.1


def test_batch_sizes_log_uniform():
    nL, nR, fr = synth.batch_sizes(256)
    n = np.concatenate([nL, nR])
    assert n.min() >= 500 and n.max() <= 16000
    ln = np.log(n)
    assert abs(ln.mean() - (np.log(500) + np.log(16000)) / 2) < 0.3
    assert fr.min() >= 0.3 and fr.max() <= 0.6
