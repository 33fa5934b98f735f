"""The learned-start MLP on the GPU (BF16 operands on the matrix cores, FP32
accumulation) against the oracle's restatement with the same bfloat16
rounding points (orc_mlp_grad), and its training loop."""
import numpy as np
import pytest

from oap_amd.api import MLP
from oracle.pyoracle import mlp_grad

pytestmark = pytest.mark.gpu


def split(dims, g):
    out, off = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        out.append(("W", g[off:off + a * b])); off += a * b
        out.append(("b", g[off:off + b])); off += b
    return out


@pytest.mark.parametrize("dims,B", [([3, 256, 256, 70], 64), ([3, 40, 24, 7], 37),
                                    ([3, 256, 256, 70], 512)])
def test_gradients_match_oracle(oracle, dims, B):
    rng = np.random.default_rng(B)
    m = MLP(dims, max_batch=512)
    p = m.init_params(seed=B)
    X = rng.uniform(-1, 1, (B, dims[0])).astype(np.float32)
    T = rng.standard_normal((B, dims[-1])).astype(np.float32)
    w = rng.uniform(0.5, 2.0, dims[-1]).astype(np.float32)
    m.set_output_weights(w)
    loss, g = m.gradients(X, T)
    lo, go = mlp_grad(oracle, dims, p, X, T, w, round_bf16=True)
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    for (kind, a), (_, b) in zip(split(dims, g), split(dims, go)):
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b) + 1e-7, kind


def test_predict_matches_bf16_forward():
    rng = np.random.default_rng(5)
    dims = [3, 16, 16, 4]
    m = MLP(dims, max_batch=8)
    p = m.init_params(seed=5)
    X = rng.uniform(-1, 1, (19, 3)).astype(np.float32)  # 3 device batches
    Y = m.predict(X)

    def bf(a):
        a = np.asarray(a, dtype=np.float32)
        u = a.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32).astype(np.float64)

    off, h = 0, bf(X)
    for li, (a, b) in enumerate(zip(dims[:-1], dims[1:])):
        W = p[off:off + a * b].reshape(b, a); off += a * b
        bias = p[off:off + b]; off += b
        z = h @ bf(W).T + bias
        h = np.tanh(z) if li < len(dims) - 2 else z
        if li < len(dims) - 2:
            h = bf(h.astype(np.float32))
    assert np.allclose(Y, h, rtol=1e-3, atol=1e-4)


def test_training_reduces_loss():
    rng = np.random.default_rng(7)
    dims = [3, 64, 64, 8]
    A = rng.standard_normal((3, 8))
    X = rng.uniform(-1, 1, (2048, 3)).astype(np.float32)
    T = np.sin(X @ A).astype(np.float32)
    m = MLP(dims, max_batch=256)
    m.init_params(seed=1)
    m.set_dataset(X, T)
    losses = m.train(batch=256, epochs=60, lr=3e-3, seed=3)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0], losses[[0, -1]]
    # the trained parameters are what predict() uses
    Y = m.predict(X[:64])
    assert np.mean((Y - T[:64]) ** 2) < 0.2 * np.mean(T[:64] ** 2)
