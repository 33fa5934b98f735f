"""The solver at the full C1 size against the committed dense oracle.

Dense complex 8192 x 8192, A = diag(d) + 1e-2 (G_re + i G_im)/sqrt(8192),
d_k = exp(-k/500) e^{2 pi i u_k} (seed 0x5EED0001), the 8 largest-|lambda|
eigenpairs with Krylov dimension 48 at tol 1e-10 ||A||_F.  The fixture
tests/golden/c1_top8.json holds the 16 largest-|lambda| eigenvalues of the
oracle's matrix (ARPACK, two start vectors agreeing to 1.6e-14, residuals
<= 1.1e-14); the 8th and 9th are 0.2 % apart in modulus, so a returned set
holding the 9th in place of the 8th is named by its fixture index.

Bar: the returned eigenvalues are fixture values 0..7, each within
1e-10 |lambda_1|, and every explicit residual is <= 1e-10 ||A||_F, for the
plain FP64 solver and the mixed flow, over several start vectors.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oap_amd._lib import check, load
from oap_amd.api import (ArnUtils, CuHArnoldiCallback, CuProceduresApi, DeviceMatrix, MatrixInfo,
                         Precision, QRType)

pytestmark = pytest.mark.gpu
FIX = json.loads((Path(__file__).resolve().parent / "golden" / "c1_top8.json").read_text())
N, K, M = 8192, 8, 48
WANT = np.array(FIX["re"]) + 1j * np.array(FIX["im"])


@pytest.fixture(scope="module")
def c1():
    A = DeviceMatrix(N, N, True, True)
    check(load().oaph_synth_matrix(A.ptr, 1, 0x5EED0001, 1e-2, 500.0), "synth")
    return A


def fixture_indices(ev):
    """index of the nearest fixture eigenvalue of each returned value, and
    the distance to it"""
    idx = [int(np.argmin(np.abs(WANT - z))) for z in ev]
    err = [float(np.min(np.abs(WANT - z))) for z in ev]
    return idx, err


def test_device_matrix_is_the_fixture_matrix(oracle, c1):
    # the generator in HBM reproduces the oracle's matrix the fixture was
    # computed from (device log/cos may differ from glibc by an ulp)
    fro = CuProceduresApi().magnitude(c1)
    assert abs(fro - FIX["fro"]) <= 1e-13 * FIX["fro"]
    ar, ai = oracle.gen_c1(N, 0x5EED0001, 1e-2, 500.0)
    re, im = c1.download()
    assert np.max(np.abs(re - ar)) <= 1e-14 and np.max(np.abs(im - ai)) <= 1e-14


@pytest.mark.parametrize("precision,projection", [(Precision.FP64, 0), (Precision.MIXED_FP32, 0),
                                                  (Precision.MIXED_FP32, 1)],
                         ids=["fp64", "mixed", "mixed_bf16x3_projection"])
def test_c1_top8_against_fixture(c1, precision, projection):
    s = CuHArnoldiCallback()
    s.setQRType(QRType.QRGR)
    s.setSortType(ArnUtils.SortLargestValues)
    s.setCheckTolerance(1e-10)
    s.setOutputType(ArnUtils.DEVICE)
    s.setMaxRestarts(2000)
    s.setPrecision(precision)
    s.setProjection(projection)  # 1: the first Gram-Schmidt pass from the BF16-split basis
    s.setMatrix(c1)
    re, im = np.zeros(K), np.zeros(K)
    s.setOutputsEigenvalues(re, im)
    vecs = [DeviceMatrix(1, N) for _ in range(K)]
    s.setOutputsEigenvectors(vecs)
    lam1 = abs(WANT[0])
    for seed in (1000, 1001, 1002, 1003):
        s.setSeed(seed)
        s.execute(M, K, MatrixInfo(True, True, N, N))
        ev = re + 1j * im
        idx, err = fixture_indices(ev)
        print(f"precision {precision} seed {seed}: fixture indices {sorted(idx)}, "
              f"max |lambda - fixture| {max(err):.3e}, stats {s.stats()}", flush=True)
        assert sorted(idx) == list(range(K)), (seed, sorted(idx))
        assert max(err) <= 1e-10 * lam1, (seed, max(err))
        if precision == Precision.MIXED_FP32:
            # the FP16 phase ran, and the FP32 cycles' pairs passed the FP64
            # check with room to spare (0.24-0.66 of the tolerance measured):
            # no FP64 cycle
            st = s.stats()
            assert st["matvecs_f16"] > 0 and st["fp64_polish"] == 0, (seed, st)
            assert 0 < st["check_resid"] <= 0.8e-10 * FIX["fro"], (seed, st["check_resid"])
            assert st["matvecs_low"] == st["matvecs"] - K
        for t in range(K):
            assert s.testOutcome(t) <= 1e-10 * FIX["fro"], (seed, t, s.testOutcome(t))
    # the returned vectors of the last solve are eigenvectors of the same
    # operator: one explicit residual through the library's own FP64 matvec
    api = CuProceduresApi()
    y = DeviceMatrix(1, N)
    for t in range(K):
        api.dotProduct(y, c1, vecs[t])
        x = vecs[t].to_complex()[:, 0]
        r = np.linalg.norm(y.to_complex()[:, 0] - ev[t] * x) / np.linalg.norm(x)
        assert r <= 1e-10 * FIX["fro"], (t, r)
