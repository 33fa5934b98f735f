"""RecToSquareApi on the GPU (liboaph.so oaph_rts_*) against the oracle.

Blocks of S = M M^T are checked against the oracle's dotProduct of the row
block with the oracle's transpose of M (the reference's own construction,
oapMatrixCuda/SquareMatrixApi.cpp:143-181) within gamma_K * sum |a||b|
(fused versus separate multiply-add, K = columns of M).  The implicit
product S v = M (M^T v) is checked against the oracle's two products within
the error bound of two chained sums.  Invariants: S is exactly symmetric for
real M (same products in the same order), a square M is returned as it is,
and the eigenvalues the solver finds through the implicit product match the
dense symmetric eigenvalues of M M^T.
"""
import numpy as np
import pytest

from oap_amd.api import (ArnUtils, CuHArnoldiCallback, DeviceMatrix, MatrixInfo, OapError, QRType,
                         RecToSquareApi)
from oracle.pyoracle import Planes

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def rec(rng, R, C, cplx):
    z = rng.standard_normal((R, C))
    if cplx:
        z = z + 1j * rng.standard_normal((R, C))
    return z


def dev(z, cplx):
    z = np.asarray(z).reshape(z.shape[0], -1)
    return DeviceMatrix.from_numpy(z.real, z.imag if cplx else None)


def host(m):
    re, im = m.download()
    z = np.zeros((m.rows, m.columns), complex)
    if re is not None:
        z.real = re
    if im is not None:
        z.imag = im
    return z


def oracle_block(oracle, M, cplx, r0, n):
    """rows [r0, r0 + n) of M M^T by the oracle's transpose + dotProduct"""
    R, C = M.shape
    MT = Planes(C, R, True, cplx)  # C rows x R columns
    assert oracle.transpose(MT, Planes.of(M, True, cplx)) == 0
    out = Planes(n, R, True, cplx)
    assert oracle.dot_product(out, Planes.of(M[r0:r0 + n], True, cplx), MT) == 0
    return out.z()


def absplanes(z):
    return np.abs(z.real) + np.abs(z.imag)


@pytest.mark.parametrize("R,C", [(1, 17), (5, 9), (64, 300), (130, 70), (200, 1)])
@pytest.mark.parametrize("cplx", [False, True])
def test_sub_blocks_match_oracle(api, oracle, rng, R, C, cplx):
    M = rec(rng, R, C, cplx)
    dM = dev(M, cplx)
    rts = RecToSquareApi(api, dM)
    inf = rts.getMatrixInfo()
    assert (inf.columns, inf.rows, bool(inf.isRe), bool(inf.isIm)) == (R, R, True, cplx)
    bound_full = 2 * C * EPS * absplanes(M) @ absplanes(M).T
    cases = [(0, R), (R // 2, 7), (R - 1, 1), (max(R - 3, 0), 10)]
    for r0, ln in cases:
        n = min(ln, R - r0)
        out = DeviceMatrix(R, 1, True, cplx)
        rts.getDeviceSubMatrix(r0, ln, out)
        assert (out.columns, out.rows) == (R, n)  # clamped length, resized output
        got = host(out)
        want = oracle_block(oracle, M, cplx, r0, n)
        assert np.all(np.abs(got - want) <= bound_full[r0:r0 + n] + 1e-300), (r0, ln)
    api.synchronize()


@pytest.mark.parametrize("cplx", [False, True])
def test_square_and_row_vector(api, oracle, rng, cplx):
    R, C = 96, 257
    M = rec(rng, R, C, cplx)
    rts = RecToSquareApi(api, dev(M, cplx))
    S = DeviceMatrix(R, R, True, cplx)
    rts.getDeviceMatrix(S)
    want = oracle_block(oracle, M, cplx, 0, R)
    bound = 2 * C * EPS * absplanes(M) @ absplanes(M).T
    got = host(S)
    assert np.all(np.abs(got - want) <= bound)
    if not cplx:
        assert np.array_equal(got.real, got.real.T)  # identical products, identical order
    # a real-plane row vector (RecToSquareApi.cpp:117) holds sum re*re only,
    # the dotProduct semantics of a real output
    row = rts.createDeviceRowVector(17)
    assert (row.columns, row.rows, row.is_im) == (R, 1, False)
    if cplx:
        want_row = oracle_block(oracle, M.real, False, 17, 1)[0].real
        assert np.all(np.abs(host(row)[0].real - want_row) <= bound[17])
    else:
        assert np.array_equal(host(row)[0].real, got[17].real)
    sub = rts.createDeviceSubMatrix(90, 50)
    assert (sub.columns, sub.rows) == (R, 6)
    assert np.array_equal(host(sub), got[90:96])


def test_square_input_is_used_as_is(api, rng):
    # RecToSquareApi.cpp:57-63,94-97: a square matrix is the operator itself
    M = rec(rng, 40, 40, True)
    rts = RecToSquareApi(api, dev(M, True))
    out = DeviceMatrix(40, 5, True, True)
    rts.getDeviceSubMatrix(10, 5, out)
    assert np.array_equal(host(out), M[10:15])
    v = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    w = DeviceMatrix(1, 40)
    rts.multiply(w, dev(v.reshape(-1, 1), True))
    assert np.all(np.abs(host(w)[:, 0] - M @ v) <= 1e-12 * (np.abs(M) @ np.abs(v)))


def test_errors(api, rng):
    M = rec(rng, 10, 30, False)
    rts = RecToSquareApi(api, dev(M, False))
    out = DeviceMatrix(10, 2, True, False)
    with pytest.raises(OapError):
        rts.getDeviceSubMatrix(10, 1, out)  # row out of scope
    with pytest.raises(OapError):
        rts.getDeviceSubMatrix(0, 0, out)  # zero length
    with pytest.raises(OapError):
        rts.getDeviceSubMatrix(0, 2, DeviceMatrix(9, 2, True, False))  # wrong width
    with pytest.raises(OapError):
        rts.multiply(DeviceMatrix(1, 9), DeviceMatrix(1, 10))  # wrong length
    with pytest.raises(OapError):
        # a complex product needs an imaginary plane in w
        rts.multiply(DeviceMatrix(1, 10, True, False), DeviceMatrix(1, 10, True, True))


@pytest.mark.parametrize("R,C", [(7, 5), (33, 1001), (300, 64), (4096, 512), (2048, 6000),
                                 (513, 3), (20000, 300), (100, 9999), (64, 70000)])
@pytest.mark.parametrize("mcplx,vcplx", [(False, False), (False, True), (True, False),
                                         (True, True)])
def test_implicit_multiply(api, oracle, rng, R, C, mcplx, vcplx):
    M = rec(rng, R, C, mcplx)
    v = rec(rng, R, 1, vcplx)
    rts = RecToSquareApi(api, dev(M, mcplx))
    cplx = mcplx or vcplx
    w = DeviceMatrix(1, R, True, True)
    w.upload(np.full(R, 7.0), np.full(R, 7.0))  # stale contents must be overwritten
    rts.multiply(w, dev(v, vcplx))
    got = host(w)[:, 0]
    if R * C <= 300_000:
        # oracle: u = M^T v, w = M u (dotProduct with the oracle's transpose)
        MT = Planes(C, R, True, True)
        Mz = M if mcplx else M + 0j
        assert oracle.transpose(MT, Planes.of(Mz, True, True)) == 0
        u = Planes(C, 1, True, True)
        vz = v if vcplx else v + 0j
        assert oracle.dot_product(u, MT, Planes.of(vz, True, True)) == 0
        wo = Planes(R, 1, True, True)
        assert oracle.dot_product(wo, Planes.of(Mz, True, True), u) == 0
        want = wo.z()[:, 0]
    else:
        want = (M @ (M.T @ v))[:, 0]
    # two chained sums: |fl(M fl(M^T v)) - M M^T v| <= gamma_{R+C} |M| |M^T| |v| (planes)
    bound = 2 * (R + C + 4) * EPS * (absplanes(M) @ (absplanes(M).T @ absplanes(v)))[:, 0]
    assert np.all(np.abs(got - want) <= bound + 1e-300)
    if not cplx:
        assert np.all(host(w)[:, 0].imag == 0)


def test_solver_through_implicit_product(api, rng):
    # the reference's PCA caller: CuHArnoldiCallback with a multiply callback
    # over RecToSquareApi (oap2dt3dDevice/MainAPExecutor.cpp:42-127)
    R, C = 400, 3000
    M = rng.standard_normal((R, C)) / np.sqrt(C)
    M[:, :4] *= np.array([40.0, 35.0, 30.0, 20.0])  # four spikes far above the bulk (< 2)
    rts = RecToSquareApi(api, dev(M, False))

    def multiply(w, v, solver_api, user, mt):
        user.multiply(w, v, solver_api)  # ordered on the solver's stream

    s = CuHArnoldiCallback()
    s.setQRType(QRType.QRGR)
    s.setSortType(ArnUtils.SortLargestValues)
    s.setCheckTolerance(1e-10)
    s.setOutputType(ArnUtils.DEVICE)
    s.setSeed(3)
    s.setCallback(multiply, rts)
    re, im = np.zeros(4), np.zeros(4)
    s.setOutputsEigenvalues(re, im)
    s.execute(24, 4, MatrixInfo(True, False, R, R))
    want = np.sort(np.linalg.eigvalsh(M @ M.T))[::-1][:4]
    got = np.sort(re)[::-1]
    assert np.all(np.abs(got - want) <= 1e-10 * want[0]), (got, want)
    assert np.all(np.abs(im) <= 1e-10 * want[0])


def test_solver_large_n_implicit_product(api, rng):
    # n = 40000 unknowns: the Arnoldi steps' CGS2 then runs on 256 blocks of
    # 192 rows each (several 64-row chunks per block, idle blocks past n);
    # S = M M^T has rank 64 and its nonzero eigenvalues are the squared
    # singular values of M
    R, C = 40000, 64
    M = rng.standard_normal((R, C)) / np.sqrt(R)
    M[:, :4] *= np.array([9.0, 8.0, 7.0, 6.0])
    rts = RecToSquareApi(api, dev(M, False))

    def multiply(w, v, solver_api, user, mt):
        user.multiply(w, v, solver_api)

    s = CuHArnoldiCallback()
    s.setQRType(QRType.QRGR)
    s.setSortType(ArnUtils.SortLargestValues)
    s.setCheckTolerance(1e-10)
    s.setOutputType(ArnUtils.DEVICE)
    s.setSeed(5)
    s.setCallback(multiply, rts)
    re, im = np.zeros(4), np.zeros(4)
    s.setOutputsEigenvalues(re, im)
    s.execute(24, 4, MatrixInfo(True, False, R, R))
    want = np.sort(np.linalg.svd(M, compute_uv=False) ** 2)[::-1][:4]
    got = np.sort(re)[::-1]
    assert np.all(np.abs(got - want) <= 1e-10 * want[0]), (got, want)
    assert np.all(np.abs(im) <= 1e-10 * want[0])
