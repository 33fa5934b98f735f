"""PCA of an image stack on the GPU (EigenCalculator = IEigenCalculator +
MainAPExecutor): data matrix from PNG files, S = M M^T applied through
RecToSquareApi, the largest eigenpairs by the Arnoldi solver; checked against
LAPACK eigvalsh of M M^T formed on the host and by explicit residuals."""
import numpy as np
import pytest

from oap_amd import imaging

pytestmark = pytest.mark.gpu


def test_eigenimages_of_a_synthetic_stack(tmp_path, rng):
    H, W, count = 28, 24, 40
    base = [rng.integers(0, 256, (H, W, 3)).astype(np.float64) for _ in range(3)]
    for i in range(count):
        w = rng.random(3)
        im = sum(wk * b for wk, b in zip(w, base)) / w.sum()
        im = np.clip(im + rng.normal(0, 6, (H, W, 3)), 0, 255).astype(np.uint8)
        imaging.write_png(tmp_path / f"face{i}.png", im)
    loader = imaging.DataLoader.from_dir(tmp_path, "face", count)
    inf = loader.getMatrixInfo()
    M = np.stack([loader.createColumnVector(i) for i in range(count)], axis=1)
    assert M.shape == (inf.rows, inf.columns)
    calc = imaging.EigenCalculator(loader)
    calc.setEigensCount(20, 4)
    out = calc.calculate()
    want = np.sort(np.linalg.eigvalsh(M @ M.T))[::-1][:4]
    got = np.sort(out.eigenvalues)[::-1]
    assert np.all(np.abs(got - want) <= 1e-10 * want[0]), (got, want)
    S = M @ M.T
    for lam, v in zip(out.eigenvalues, out.eigenvectors):
        assert np.linalg.norm(S @ v - lam * v) <= 1e-10 * want[0] * np.linalg.norm(v)
    assert np.all(out.errors <= 1e-10 * np.linalg.norm(S))


@pytest.mark.parametrize("count,H,W", [(3, 5, 7), (37, 33, 29), (64, 40, 40)])
def test_data_matrix_is_the_stacked_columns(tmp_path, rng, count, H, W):
    # createMatrix (image-major upload + device transpose) holds exactly the
    # columns createColumnVector gives, image i in column i
    for i in range(count):
        imaging.write_png(tmp_path / f"img{i}.png", rng.integers(0, 256, (H, W, 3), dtype=np.uint8))
    loader = imaging.DataLoader.from_dir(tmp_path, "img", count)
    want = np.stack([loader.createColumnVector(i) for i in range(count)], axis=1)
    got = loader.createMatrix().to_complex()
    assert got.shape == want.shape
    assert np.array_equal(got.real, want)
