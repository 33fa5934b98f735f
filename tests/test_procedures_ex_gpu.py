"""GPU checks of the dims forms, MatrixEx procedures and matrix utilities of
the procedures API (oapMatrixCuda/CuProceduresApi.h:55-309,
oapCudaMatrixUtils.h:95-200): against the reference's own transposeEx test
vectors, numpy's products (closed forms of the periodic and region
definitions), exact element-wise results, and the oracle's calcTriangularH
and the schur fixtures.  Products are checked within gamma_K sum |a||b|;
element-wise and copy operations bit-exactly."""
import json
from pathlib import Path

import numpy as np
import pytest

from oap_amd.api import CuProceduresApi, DeviceMatrix, OapError
from oracle.pyoracle import Planes

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
EPS = 2.220446049250313e-16


def dev(z, re=True, im=True):
    z = np.asarray(z)
    return DeviceMatrix.from_numpy(z.real if re else None, z.imag if im else None)


def cplx(rng, r, c):
    return rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))


def close_product(got, want, bound_src, K):
    a, b = bound_src
    bound = 2 * K * EPS * (np.abs(a) @ np.abs(b)) + 1e-300
    assert np.all(np.abs(got - want) <= bound)


@pytest.mark.parametrize("case", json.loads((GOLDEN / "transpose_ex.json").read_text())["cases"],
                         ids=lambda c: c["name"])
def test_transpose_ex_reference_vectors(api, case):
    a = np.array(case["input"]).reshape(case["input_rows"], case["input_columns"])
    out = DeviceMatrix(case["output_columns"], case["output_rows"], True, False)
    api.setZeroMatrix(out)
    api.transposeEx(out, dev(a, True, False), case["matrix_ex"])
    want = np.array(case["expected"]).reshape(case["expected_rows"], case["expected_columns"])
    np.testing.assert_array_equal(out.download()[0], want)


def test_transpose_ex_grid_extent(api, oracle, rng):
    # the written region is the reference's launch grid for (columns, rows)
    # (SetThreadsBlocks, which the oracle restates bit-exactly) cut at the
    # MatrixEx bound (column + columns, row + rows)
    for cols, rows, col0, row0 in [(40, 3, 0, 0), (1000, 3, 7, 1), (33, 17, 5, 2)]:
        (bx, by), (tx, ty) = oracle.set_threads_blocks(cols, rows, 1024)
        W, H = min(bx * tx, col0 + cols), min(by * ty, row0 + rows)
        a = cplx(rng, W + 3, H + 2)  # rows >= W, columns >= H
        out = DeviceMatrix(W + 1, H + 1)
        api.setZeroMatrix(out)
        api.transposeEx(out, dev(a), (col0, cols, row0, rows))
        want = np.zeros((H + 1, W + 1), complex)
        want[:H, :W] = a[:W, :H].T
        np.testing.assert_array_equal(out.to_complex(), want)


def test_dot_product_ex_region(api, rng):
    a, b = cplx(rng, 9, 7), cplx(rng, 7, 11)
    out = DeviceMatrix(11, 9)
    api.setZeroMatrix(out)
    api.dotProductEx(out, dev(a), dev(b), (3, 5, 2, 4))
    want = np.zeros((9, 11), complex)
    want[2:6, 3:8] = (a @ b)[2:6, 3:8]
    close_product(out.to_complex(), want, (a, b), 7)
    with pytest.raises(OapError):
        api.dotProductEx(out, dev(a), dev(b), (8, 5, 2, 4))  # beyond the output


def test_dot_product_periodic(api, rng):
    # C = [A B_0; A B_1; ...] with B split into blocks of a.columns rows
    a, b = cplx(rng, 3, 3), cplx(rng, 12, 5)
    out = DeviceMatrix(5, 12)
    api.dotProductPeriodic(out, dev(a), dev(b))
    want = np.vstack([a @ b[3 * i:3 * i + 3] for i in range(4)])
    bound = 6 * EPS * np.vstack([np.abs(a) @ np.abs(b[3 * i:3 * i + 3]) for i in range(4)])
    assert np.all(np.abs(out.to_complex() - want) <= bound)
    # re-only planes
    ar, br = a.real, b.real
    out = DeviceMatrix(5, 12, True, False)
    api.dotProductPeriodic(out, dev(ar, True, False), dev(br, True, False))
    want = np.vstack([ar @ br[3 * i:3 * i + 3] for i in range(4)])
    np.testing.assert_allclose(out.download()[0], want, rtol=0, atol=1e-14 * np.abs(want).max())


def test_dot_product_dim_periodic(api, rng):
    # dims: output block W x H inside each period of P rows, K terms
    W, H, K, P, blocks = 4, 2, 3, 3, 4
    a, b = cplx(rng, P, K + 1), cplx(rng, K * blocks, W + 2)
    out = DeviceMatrix(W + 1, P * blocks)
    api.setZeroMatrix(out)
    api.dotProductDimPeriodic(out, dev(a), dev(b), [[W, H], [K, H], [W, K]], P)
    want = np.zeros((P * blocks, W + 1), complex)
    for y in range(P * blocks):
        if y % P < H:
            blk = (y // P) * K
            want[y, :W] = a[y % P, :K] @ b[blk:blk + K, :W]
    np.testing.assert_allclose(out.to_complex(), want, rtol=0, atol=1e-13)


def test_add_dot_product_writes_the_product(api, rng):
    # the reference kernel assigns (CuAddDotProductProcedures.h:26-84)
    a, b = cplx(rng, 6, 5), cplx(rng, 5, 4)
    out = dev(cplx(rng, 6, 4))
    api.addDotProduct(out, dev(a), dev(b))
    close_product(out.to_complex(), a @ b, (a, b), 5)


def test_region_add_substract(api, rng):
    o0, a, b = cplx(rng, 7, 9), cplx(rng, 7, 9), cplx(rng, 7, 9)
    for name, f in (("addCR", lambda o, x, y: x + y), ("substractCR", lambda o, x, y: x - y),
                    ("addSubstract", lambda o, x, y: o + (x - y))):
        out = dev(o0)
        getattr(api, name)(out, dev(a), dev(b), 5, 4)
        want = o0.copy()
        want[:4, :5] = f(o0[:4, :5], a[:4, :5], b[:4, :5])
        np.testing.assert_array_equal(out.to_complex(), want)
    with pytest.raises(OapError):
        api.addCR(dev(o0), dev(a), dev(b), 10, 4)


def test_hadamard_and_tensor_cr_forms(api, rng):
    a, b, v = cplx(rng, 5, 6), cplx(rng, 5, 6), cplx(rng, 5, 1)
    out = DeviceMatrix(6, 5)
    api.hadamardProductCR(out, dev(a), dev(b), 6, 5)
    z = out.to_complex()
    np.testing.assert_array_equal(z.real, a.real * b.real)  # per plane (Q3)
    np.testing.assert_array_equal(z.imag, a.imag * b.imag)
    api.hadamardProductVecCR(out, dev(a), dev(v), 6, 5)
    np.testing.assert_array_equal(out.to_complex().real, a.real * v.real)
    with pytest.raises(OapError):
        api.hadamardProductCR(out, dev(a), dev(b), 5, 5)
    x, y = rng.standard_normal((2, 3)), rng.standard_normal((4, 2))
    t = DeviceMatrix(6, 8, True, False)
    api.tensorProductCR(t, dev(x, True, False), dev(y, True, False), 6, 8)
    np.testing.assert_array_equal(t.download()[0], np.kron(x, y))


def test_tensor_product_sub_window(api, rng):
    # dims select the leading 2 x 3 block of A, the leading 2 x 2 block of B,
    # and a 4 x 6 window of a larger output (CuTensorProductDimProcedures.h:26-48)
    A, B = rng.standard_normal((5, 4)), rng.standard_normal((3, 3))
    out = DeviceMatrix(7, 9, True, False)
    api.setZeroMatrix(out)
    api.tensorProduct(out, dev(A, True, False), dev(B, True, False),
                      [[4, 6], [2, 3], [2, 2]])
    want = np.zeros((9, 7))
    want[:6, :4] = np.kron(A[:3, :2], B[:2, :2])
    np.testing.assert_array_equal(out.download()[0], want)
    api.tensorProduct(out, dev(A, True, False), dev(B, True, False), [[4, 6], [2, 3], [2, 2]])
    with pytest.raises(OapError):  # window beyond the output
        api.tensorProduct(out, dev(A, True, False), dev(B, True, False),
                          [[8, 6], [4, 3], [2, 2]])


def test_magnitude2_family(api, rng):
    z = cplx(rng, 33, 17)
    d = dev(z)
    m2 = api.magnitude2(d)
    assert abs(m2 - np.sum(np.abs(z) ** 2)) <= 1e-13 * m2
    assert api.magnitude2Opt(d) == m2 and api.magnitude2OptVer2(d) == m2
    assert api.magnitudeOpt(d) == api.magnitude(d) == api.magnitudeOptVer2(d)


@pytest.mark.parametrize("case", json.loads((GOLDEN / "schur.json").read_text())[:2],
                         ids=lambda c: c["name"])
def test_calc_triangular_h_against_oracle_and_fixture(api, oracle, case):
    # oapArnoldiDeviceTests/oapSchurDecompositionTests.cpp:61-125 through the
    # device kernel, against the oracle's restatement on the same input
    m = case["m"]
    H = np.array(case["matrix"]).reshape(m, m)
    dH, dQ = dev(H, True, False), DeviceMatrix(m, m, True, False)
    api.calcTriangularH(dH, dQ)
    T, Q = dH.download()[0], dQ.download()[0]
    Hp, Qp = Planes.of(H, True, False), Planes(m, m, True, False)
    oracle.calc_triangular_h(Hp, Qp, 10000)
    # the same iteration, rounded differently over its steps (unshifted QR to
    # a 1e-3 triangularity: hundreds of steps on schur2)
    np.testing.assert_allclose(T, Hp.re, rtol=0, atol=1e-10 * abs(Hp.re).max())
    np.testing.assert_allclose(Q, Qp.re, rtol=0, atol=1e-10)
    want = np.diag(np.array(case["eq_matrix"]).reshape(m, m))
    np.testing.assert_allclose(sorted(np.diag(T)), sorted(want), atol=0.2)  # the fixture's tolerance
    np.testing.assert_allclose(Q @ T @ Q.T, H, atol=1e-10 * abs(H).max())


@pytest.mark.parametrize("cplx_in", [False, True])
def test_calc_triangular_h_step(api, rng, cplx_in):
    # one unshifted QR step: H1 = R Q with H0 = Q R, R upper triangular
    m = 12
    H = np.triu(cplx(rng, m, m) if cplx_in else rng.standard_normal((m, m)), -1)
    dH = dev(H, True, cplx_in)
    dQ = DeviceMatrix(m, m, True, cplx_in)
    api.calcTriangularHStep(dH, dQ)
    H1, Q = dH.to_complex(), dQ.to_complex()
    np.testing.assert_allclose(Q.conj().T @ Q, np.eye(m), atol=1e-13)
    R = Q.conj().T @ H
    np.testing.assert_allclose(np.tril(R, -1), 0, atol=1e-12 * abs(H).max())
    np.testing.assert_allclose(H1, R @ Q, atol=1e-12 * abs(H).max())
    # an upper triangular H: no step, Q = I
    T = np.triu(rng.standard_normal((m, m)))
    dT = dev(T, True, False)
    dQ = DeviceMatrix(m, m, True, False)
    api.calcTriangularHStep(dT, dQ)
    np.testing.assert_array_equal(dT.download()[0], T)
    np.testing.assert_array_equal(dQ.download()[0], np.eye(m))


def test_matrix_values_rows_and_copies(rng):
    z = cplx(rng, 5, 7)
    d = dev(z)
    assert d.get_value(3 + 2 * 7) == z[2, 3]
    d.set_value(3 + 2 * 7, 1.5, -2.5)
    z[2, 3] = 1.5 - 2.5j
    np.testing.assert_array_equal(d.to_complex(), z)
    d.set_zero_row(4, re=True, im=False)
    z[4] = 1j * z[4].imag
    np.testing.assert_array_equal(d.to_complex(), z)
    d.set_zero_row(1)
    z[1] = 0
    np.testing.assert_array_equal(d.to_complex(), z)
    e = DeviceMatrix(7, 5)
    e.copy_from(d)
    np.testing.assert_array_equal(e.to_complex(), z)
    with pytest.raises(OapError):
        d.get_value(35)
    with pytest.raises(OapError):
        d.set_zero_row(5)


def test_copy_dims_and_set_matrix(rng):
    src = cplx(rng, 6, 8)
    dsrc = dev(src)
    # d2d: the 3 x 4 block at (column 2, row 1) of src into (column 0, row 3) of dst
    dst0 = cplx(rng, 7, 5)
    ddst = dev(dst0)
    dims = [[[0, 4], [3, 3]], [[2, 4], [1, 3]]]
    ddst.copy_from_dims(dsrc, dims)
    want = dst0.copy()
    want[3:6, 0:4] = src[1:4, 2:6]
    np.testing.assert_array_equal(ddst.to_complex(), want)
    # h2d and d2h with host planes
    hre, him = np.ascontiguousarray(src.real), np.ascontiguousarray(src.imag)
    ddst = dev(dst0)
    ddst.upload_dims(hre, him, dims)
    np.testing.assert_array_equal(ddst.to_complex(), want)
    back_re, back_im = np.zeros((6, 8)), np.zeros((6, 8))
    dsrc.download_dims(back_re, back_im, [[[1, 4], [0, 3]], [[2, 4], [1, 3]]])
    wre = np.zeros((6, 8))
    wre[0:3, 1:5] = src.real[1:4, 2:6]
    np.testing.assert_array_equal(back_re, wre)
    with pytest.raises(OapError):  # sizes differ
        ddst.copy_from_dims(dsrc, [[[0, 4], [3, 3]], [[2, 3], [1, 3]]])
    with pytest.raises(OapError):  # beyond the source
        ddst.copy_from_dims(dsrc, [[[0, 4], [0, 3]], [[6, 4], [1, 3]]])
    # SetMatrix / SetReMatrix / SetImMatrix
    small = cplx(rng, 2, 3)
    for planes in (3, 1, 2):
        ddst = dev(dst0)
        ddst.set_matrix(dev(small), 1, 2, planes)  # dst0: 7 rows x 5 columns
        w = dst0.copy()
        if planes & 1:
            w.real[2:4, 1:4] = small.real
        if planes & 2:
            w.imag[2:4, 1:4] = small.imag
        np.testing.assert_array_equal(ddst.to_complex(), w)
    with pytest.raises(OapError):
        dev(dst0).set_matrix(dev(small), 5, 2)


def test_cpp_facade_procedures():
    # tests/cpp/facade_procedures.cpp: the facade's MatrixEx, periodic, dims,
    # addSubstract, magnitude2, calcTriangularHStep and matrix-utility calls
    import subprocess
    exe = Path(__file__).resolve().parent / "cpp" / "facade_procedures"
    assert exe.exists(), "build the library first (python -m oap_amd.build)"
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    # CuHArnoldiDefault with the comparator byRotatedRe (largest re + im):
    # the dense top 3 of that order
    n = 64
    rr, cc = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    off = 0.01 * np.sin(7.0 * rr + 3.0 * cc)
    d = 1.0 / (1.0 + 0.05 * np.arange(n))
    A = off + 0.5j * off
    A[np.arange(n), np.arange(n)] = (np.cos(0.1 * np.arange(n)) + 1j * np.sin(0.1 * np.arange(n))) * d
    want = sorted(np.linalg.eigvals(A), key=lambda z: -(z.real + z.imag))[:3]
    got = np.array(r["eig_rotated"][0::2]) + 1j * np.array(r["eig_rotated"][1::2])
    for w in want:
        assert np.min(np.abs(got - w)) <= 1e-10, (got, want)
    x = np.array([(i % 7) - 2.0 for i in range(40)]).reshape(10, 4)
    tex = np.zeros((4, 10))
    tex[:2] = x[:, :2].T
    np.testing.assert_array_equal(np.array(r["transpose_ex"]).reshape(4, 10), tex)
    a = np.array([[1, 2], [3, 4]], float)
    b = np.array([1, 0, 0, 1, 2, 1, 1, 3, -1, 2, 0.5, 4], float).reshape(6, 2)
    per = np.vstack([a @ b[2 * i:2 * i + 2] for i in range(3)])
    np.testing.assert_allclose(np.array(r["periodic"]).reshape(6, 2), per, rtol=0, atol=1e-15)
    # addSubstract(o, a, a, 2 columns, 1 row) leaves o; add(o, o, a, 1 column, 2 rows)
    o = np.array([[10, 20], [30, 40]], float)
    o[:, :1] += a[:, :1]
    np.testing.assert_array_equal(np.array(r["addsub"]).reshape(2, 2), o)
    assert r["magnitude2"] == 30.0 and abs(r["magnitude"] - np.sqrt(30.0)) <= 1e-15
    h = np.array([4, 1, 2, 3, 5, 1, 0, 2, 6], float).reshape(3, 3)
    q1 = np.array(r["q1"]).reshape(3, 3)
    np.testing.assert_allclose(q1.T @ q1, np.eye(3), atol=1e-14)
    rr = q1.T @ h
    np.testing.assert_allclose(np.tril(rr, -1), 0, atol=1e-13)
    np.testing.assert_allclose(np.array(r["h1"]).reshape(3, 3), rr @ q1, atol=1e-13)
    m = np.zeros((3, 4))
    m[1:3, 1:3] = a
    m[2, 0] = 9.5
    assert r["value21"] == m[1, 2]
    host = -np.ones((3, 4))
    host[0:2, 0:2] = m[1:3, 1:3]
    np.testing.assert_array_equal(np.array(r["dims_host"]).reshape(3, 4), host)
    m[1] = 0
    np.testing.assert_array_equal(np.array(r["set_matrix"]).reshape(3, 4), m)
