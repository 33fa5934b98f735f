"""Krylov dimensions above 64: the general restart kernel (T in LDS up to
m = 90, in global memory beyond, Z in global memory, single-shift sweeps) and
the column-looping CGS2 chain, against the dense FP64 eigensolver (LAPACK
through numpy), the oracle's IRAM restatement and explicit FP64 residuals.

The reference runs any hdim: its QR falls back to a host-driven path once
m^2 > 1024 (oapMatrixCuda/CuProceduresApi.cpp:439-455); here m <= 64 keeps the
LDS-resident restart and 64 < m <= 256 takes the general kernel.

Bar: eigenvalues within 1e-10 relative of the dense oracle, residuals
||A x - lambda x||_2 <= 1e-10 ||A||_F for every returned pair.
"""
import numpy as np
import pytest

from oap_amd.api import ArnUtils, DeviceMatrix, MatrixInfo, OapError, Precision

from test_iram_gpu import check_residuals, match, run_dense, solver, top_k

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,k", [(65, 8), (80, 12), (128, 16)])
def test_general_restart_complex(oracle, m, k):
    # m = 65: the smallest general launch; 80: T in LDS; 128: T in global memory
    n = 1024
    ar, ai = oracle.gen_c1(n, 0x5EED0001, 1e-2, 500.0)
    A = ar + 1j * ai
    s, ev, vecs = run_dense(A, k, m)
    want = top_k(np.linalg.eigvals(A), k, ArnUtils.SortLargestValues)
    match(ev, want, 1e-10)
    check_residuals(s, A, ev, vecs, k)
    assert s.stats()["restarts"] >= 1


def test_general_restart_mixed_precision(oracle):
    n, m, k = 1024, 96, 8
    ar, ai = oracle.gen_c1(n, 0x5EED0001, 1e-2, 500.0)
    A = ar + 1j * ai
    s, ev, vecs = run_dense(A, k, m, precision=Precision.MIXED_FP32)
    match(ev, top_k(np.linalg.eigvals(A), k, ArnUtils.SortLargestValues), 1e-10)
    check_residuals(s, A, ev, vecs, k)
    assert s.stats()["matvecs_low"] > 0


@pytest.mark.parametrize("which", [ArnUtils.SortSmallestReValues, ArnUtils.SortLargestReValues])
def test_general_restart_real_sort_types(rng, which):
    # a real nonsymmetric operator with a known spectrum: eigenvalues 1..n
    # spread by a similarity transform (complex-conjugate-free, well separated)
    n, m, k = 400, 100, 6
    X = np.eye(n) + 0.05 * rng.standard_normal((n, n)) / np.sqrt(n)
    A = X @ np.diag(np.linspace(1.0, 4.0, n) ** 2) @ np.linalg.inv(X)
    s, ev, vecs = run_dense(A, k, m, which=which)
    match(ev, top_k(np.linalg.eigvals(A), k, which), 1e-10)
    check_residuals(s, A, ev, vecs, k)


def test_general_restart_against_oracle_iram(oracle):
    # the oracle's IRAM restatement (any m) on the same operator: the same
    # eigenvalues
    n, m, k = 600, 72, 6
    ar, ai = oracle.gen_c1(n, 0x5EED0003, 1e-2, 200.0)
    A = ar + 1j * ai
    s, ev, vecs = run_dense(A, k, m)
    st, oev, _, _, _ = oracle.iram(lambda x: A @ x, n, m, k, tol=1e-10,
                                   anorm=float(np.linalg.norm(A)))
    assert st == 0
    match(ev, oev, 1e-10)
    match(oev, ev, 1e-10)
    check_residuals(s, A, ev, vecs, k)


def test_general_restart_with_comparator(oracle):
    # a caller's comparator with m > 64: the first launch stores the general
    # kernel's Schur form, the second goes on from it
    n, m, k = 800, 70, 5
    ar, ai = oracle.gen_c1(n, 0x5EED0005, 1e-2, 300.0)
    A = ar + 1j * ai
    dA = DeviceMatrix.from_numpy(A.real, A.imag)
    s = solver()
    s.setMatrix(dA)
    s.setSortType(lambda a, b: abs(a) > abs(b))  # largest modulus first
    re, im = np.zeros(k), np.zeros(k)
    s.setOutputsEigenvalues(re, im)
    vecs = [DeviceMatrix(1, n) for _ in range(k)]
    s.setOutputsEigenvectors(vecs)
    s.execute(m, k, MatrixInfo(True, True, n, n))
    ev = re + 1j * im
    match(ev, top_k(np.linalg.eigvals(A), k, ArnUtils.SortLargestValues), 1e-10)
    check_residuals(s, A, ev, vecs, k)


def test_general_limit():
    s = solver()
    dA = DeviceMatrix.from_numpy(np.eye(300), None)
    s.setMatrix(dA)
    with pytest.raises(OapError):  # above the general kernel's 256
        s.execute(257, 4, MatrixInfo(True, False, 300, 300))
