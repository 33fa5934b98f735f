"""The oracle's learned-start MLP gradients (orc_mlp_grad) against finite
differences of its own loss (plain FP64 mode), and its bfloat16 rounding
against torch's float32 -> bfloat16 conversion (round to nearest even)."""
import numpy as np
import pytest

from oracle.pyoracle import mlp_grad


def test_bf16_rounding_matches_torch(oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-8, 8, 2000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.0078125, 65504.0, 3.0e38]]).astype(np.float32)
    # values half-way between two bfloat16 numbers round to the even one
    v = np.concatenate([v, (np.float32(1.0) + np.float32(2.0 ** -8) * np.arange(1, 8, 2)).astype(np.float32)])
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([oracle.lib.orc_bf16(float(x)) for x in v], dtype=np.float32)
    assert np.array_equal(got, want)


def test_mlp_gradients_finite_differences(oracle):
    rng = np.random.default_rng(1)
    dims = [3, 6, 5, 4]
    nparam = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = (rng.standard_normal(nparam) * 0.4).astype(np.float32)
    X = rng.standard_normal((5, 3)).astype(np.float32)
    T = rng.standard_normal((5, 4)).astype(np.float32)
    w = np.array([1.0, 0.5, 2.0, 1.0], dtype=np.float32)
    loss, g = mlp_grad(oracle, dims, p, X, T, w, round_bf16=False)
    # parameters are float32: perturb in float64 through a float64 copy is not
    # possible, so use a step that is exact in float32 and a central difference
    h = np.float32(2.0 ** -10)
    for i in range(nparam):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fp, _ = mlp_grad(oracle, dims, pp, X, T, w, round_bf16=False)
        fm, _ = mlp_grad(oracle, dims, pm, X, T, w, round_bf16=False)
        step = float(pp[i]) - float(pm[i])
        fd = (fp - fm) / step
        assert abs(fd - g[i]) <= 2e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_bf16_gradients_close_to_fp64(oracle):
    rng = np.random.default_rng(2)
    dims = [3, 64, 64, 10]
    nparam = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    p = (rng.standard_normal(nparam) * 0.2).astype(np.float32)
    X = rng.uniform(-1, 1, (32, 3)).astype(np.float32)
    T = rng.standard_normal((32, 10)).astype(np.float32)
    w = np.ones(10, dtype=np.float32)
    l64, g64 = mlp_grad(oracle, dims, p, X, T, w, round_bf16=False)
    l16, g16 = mlp_grad(oracle, dims, p, X, T, w, round_bf16=True)
    assert abs(l16 - l64) <= 2e-2 * l64
    assert np.linalg.norm(g16 - g64) <= 3e-2 * np.linalg.norm(g64)
