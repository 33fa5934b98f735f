"""The oracle against the reference's own golden vectors and known answers
(CPU only).  Each check is chosen so that a plausible mistake in the
restatement (a dropped term, a transposed operand, a wrong sign or index)
fails it."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle.pyoracle import Planes

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return json.loads((GOLDEN / name).read_text())


def test_dot_product_golden_arrays(oracle):
    # oapHostTests/oapDotProductTest.cpp:240-313: CustomDim tests 1-3
    d = load("products.json")
    cases = [("dot_1", (2, 4), (4, 3), [[3, 2], [4, 2], [3, 4]], (2, 3)),
             ("dot_2", (2, 5), (5, 4), [[3, 2], [4, 2], [3, 4]], (2, 3)),
             ("dot_3", (1, 6), (6, 1), [[1, 1], [5, 1], [1, 5]], (1, 2))]
    for name, sa, sb, dims, so in cases:
        a = np.array(d[name]["t_reValues1"]).reshape(sa)
        b = np.array(d[name]["t_reValues2"]).reshape(sb)
        o = Planes(max(so[0], 2), max(so[1], 3), True, False)
        assert oracle.dot_product(o, Planes.of(a, True, False), Planes.of(b, True, False), dims) == 0
        want = np.array(d[name]["t_outputValues"]).reshape(so)
        w, h = dims[0]
        np.testing.assert_allclose(o.re[:h, :w], want[:h, :w], atol=1e-12)


def test_dot_product_constant_closed_forms(oracle):
    # oapDotProductTest.cpp:25-122: constant 2 matrices -> 4 k in every output
    for (m, k, n) in [(4, 1, 4), (1, 4, 1), (4, 4, 4), (10, 7, 3)]:
        a, b = np.full((m, k), 2.0), np.full((k, n), 2.0)
        o = Planes(m, n, True, False)
        oracle.dot_product(o, Planes.of(a, True, False), Planes.of(b, True, False))
        assert np.all(o.re == 4 * k)


def test_tensor_product_golden(oracle):
    # oapHostTests/oapTensorProductTest.cpp:21-73
    d = load("products.json")["tensor_1"]
    a = np.array(d["t_reValues1"]).reshape(2, 1)
    b = np.array(d["t_reValues2"]).reshape(1, 1)
    o = Planes(2, 1, True, False)
    oracle.tensor_product(o, Planes.of(a, True, False), Planes.of(b, True, False),
                          [[1, 2], [1, 2], [1, 1]])
    np.testing.assert_allclose(o.re.ravel(), d["t_outputValues"])


@pytest.mark.parametrize("case", load("qr.json"), ids=lambda c: c["name"])
def test_qrgr_fixtures(oracle, case):
    # A = QR, Q^T Q = I, R upper (oapHostTests/oapQRTests.cpp:104-149); Q/R
    # element equality is disabled upstream (:140-142) except the host
    # fixtures' own refs, which fix the sign convention checked below
    m = case["m"]
    A = np.array(case["matrix"]).reshape(m, m)
    Q, R = Planes(m, m, True, False), Planes(m, m, True, False)
    assert oracle.qrgr(Q, R, Planes.of(A, True, False)) == 0
    np.testing.assert_allclose(Q.re @ R.re, A, atol=1e-9 * max(1, abs(A).max()))
    np.testing.assert_allclose(Q.re.T @ Q.re, np.eye(m), atol=1e-12)
    # every position whose INPUT entry passes the reference's skip test
    # (|A[fb][fa]| >= 1e-5 for m*m <= 1024, CuQRProcedures_GR.h:118-121) is
    # rotated to zero; skipped positions may keep fill-in, as in the reference
    skip = 1e-5 if m * m <= 1024 else 1e-10
    rotated = np.tril(np.abs(A) >= skip, -1)
    assert np.abs(R.re[rotated]).max(initial=0) <= 1e-8 * abs(A).max()
    if "qref" in case and "rref" in case and m <= 4:
        # refs are printed to 5-6 digits; MatrixIsEqual's default tolerance is
        # 1e-4 (oapTests/InfoType.h:28)
        np.testing.assert_allclose(np.abs(R.re), np.abs(np.array(case["rref"]).reshape(m, m)),
                                   atol=1e-4)
        np.testing.assert_allclose(np.abs(Q.re[:, :m - 1]),
                                   np.abs(np.array(case["qref"]).reshape(m, m))[:, :m - 1],
                                   atol=1e-4)


@pytest.mark.parametrize("case", load("schur.json")[:2], ids=lambda c: c["name"])
def test_calc_triangular_h_schur_fixtures(oracle, case):
    # oapArnoldiDeviceTests/oapSchurDecompositionTests.cpp:61-125: diag of the
    # triangular form, e.g. {6.5056, -0.60555, -1} for [[3,2,1],[4,2,1],[4,4,0]]
    m = case["m"]
    H = np.array(case["matrix"]).reshape(m, m)
    Hp, Q = Planes.of(H, True, False), Planes(m, m, True, False)
    oracle.calc_triangular_h(Hp, Q, 10000)
    assert oracle.is_upper_triangular(Hp)
    want = np.diag(np.array(case["eq_matrix"]).reshape(m, m))
    # the fixture is checked upstream with MatrixContainsDiagonalValues, whose
    # Complex comparison tolerance is 0.2 (oapMath/Math.cpp:34-39); schur1's
    # 6.5056 is a typo for 6.60555 (trace 5), so the tight check is against
    # the LAPACK eigenvalues, to the 1e-3 triangularity the iteration stops at
    np.testing.assert_allclose(sorted(np.diag(Hp.re)), sorted(want), atol=0.2)
    np.testing.assert_allclose(sorted(np.diag(Hp.re)), sorted(np.linalg.eigvals(H).real),
                               atol=1e-2 * max(1.0, np.abs(want).max()))
    # similarity: H = Q T Q^T
    np.testing.assert_allclose(Q.re @ Hp.re @ Q.re.T, H, atol=1e-8 * abs(H).max())


@pytest.mark.parametrize("m", [2, 5, 16, 48, 64])
@pytest.mark.parametrize("cplx", [False, True])
def test_hess_schur_against_lapack(oracle, m, cplx):
    rng = np.random.default_rng(m)
    H = np.triu(rng.standard_normal((m, m)), -1)
    if cplx:
        H = H + 1j * np.triu(rng.standard_normal((m, m)), -1)
    st, T, Z, w = oracle.hess_schur(H.astype(complex))
    assert st == 0
    np.testing.assert_allclose(Z @ T @ Z.conj().T, H, atol=1e-12 * m)
    np.testing.assert_allclose(Z.conj().T @ Z, np.eye(m), atol=1e-13 * m)
    assert np.abs(np.tril(T, -1)).max() == 0.0
    ev = np.linalg.eigvals(H)
    for x in w:
        assert np.min(np.abs(ev - x)) <= 1e-10 * np.abs(ev).max()


def test_hess_schur_qrtest_samples(oracle):
    # the reference's recorded Arnoldi Hessenberg matrices (oapQRTestSamples)
    for case in load("qr.json")[:9]:
        m = case["m"]
        # samples 3-9 carry fill below the subdiagonal (up to 5e-2); the
        # Hessenberg solver takes the Hessenberg part
        H = np.triu(np.array(case["matrix"]).reshape(m, m), -1)
        st, T, Z, w = oracle.hess_schur(H.astype(complex))
        assert st == 0
        ev = np.linalg.eigvals(H)
        for x in w:
            assert np.min(np.abs(ev - x)) <= 1e-9 * np.abs(ev).max()


def test_iram_c0_against_fixture(oracle):
    A = oracle.gen_c0(64, 0x5EED0000)
    fix = load("c0_eigs.json")
    assert A.sum() == pytest.approx(fix["sum"], rel=1e-14)  # generator pinned
    want = (np.array(fix["re"]) + 1j * np.array(fix["im"]))[:4]
    st, ev, X, res, info = oracle.iram(lambda x: A @ x, 64, 16, 4, 0, 1e-10, fix["fro"])
    assert st == 0 and info["nconv"] == 4
    for w in want:
        assert np.min(np.abs(ev - w)) <= 1e-10 * abs(want[0])
    assert np.all(res <= 1e-10 * fix["fro"])


def test_iram_cmatrixdata(oracle):
    # oapArnoldiPackageMatricesTests.cpp:224-266 spectra
    for case in load("cmatrixdata.json")["cases"]:
        n = case["rows"]
        A = np.array(case["matrix"]).reshape(n, case["columns"])
        k, m = case["wanted"], case["hdim"]
        st, ev, X, res, info = oracle.iram(lambda x: A @ x, n, m, k, 0, 1e-10,
                                           np.linalg.norm(A), seed=3)
        want = sorted(case["eigenvalues"], key=lambda z: -abs(z))[:k]
        np.testing.assert_allclose(sorted(np.abs(ev)), sorted(np.abs(want)), rtol=1e-9)


def test_generators_pinned(oracle):
    # SplitMix64 (Steele, Lea, Flood 2014): published first output of seed 0
    assert oracle.splitmix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    u = np.array([oracle.uniform(123, i) for i in range(20000)])
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    g = np.array([oracle.normal(7, i) for i in range(20000)])
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1) < 0.03


def test_iram_breakdown_first_value_is_one(oracle):
    # initVvector_fvis1 (ArnoldiProcedures.cpp:279-287) on a diagonal matrix:
    # e0 is an eigenvector, the factorisation breaks down at once; the
    # restart must chase its shifts inside the unreduced trailing block
    n = 64
    d = np.arange(1, n + 1, dtype=float)
    A = np.diag(d)
    v0 = np.zeros(n)
    v0[0] = 1.0
    st, ev, X, res, info = oracle.iram(lambda x: A @ x, n, 12, 3, 0, 1e-10,
                                       np.linalg.norm(A), v0=v0)
    assert st == 0 and info["nconv"] == 3
    np.testing.assert_allclose(sorted(ev.real), [62, 63, 64], rtol=1e-10)
