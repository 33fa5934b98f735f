"""Every launch shape of the streaming matvecs (rows per wave x loads in
flight x row groups per wave) computes the product to FP64 accuracy (the shapes differ only in how
a lane's partial sums are grouped): dotProduct through each FP64 shape
against the exact product within gamma_n sum |a||x|, and the mixed solver
through each FP32-operand shape against the dense eigenvalues."""
import ctypes as C

import numpy as np
import pytest

from oap_amd._lib import check, load
from oap_amd.api import CuProceduresApi, DeviceMatrix

pytestmark = pytest.mark.gpu


def shapes(fp32):
    lib, out, v = load(), [], 0
    while True:
        r, u, g = C.c_int(), C.c_int(), C.c_int()
        if lib.oaph_matvec_shape_info(fp32, v, C.byref(r), C.byref(u), C.byref(g)) != 0:
            return out
        out.append((r.value, u.value, g.value))
        v += 1


@pytest.mark.parametrize("n,cplx", [(1000, True), (1024, False), (4098, True)])
def test_fp64_shapes_bit_identical(n, cplx):
    lib = load()
    rng = np.random.default_rng(n)
    A = DeviceMatrix.from_numpy(rng.standard_normal((n, n)), rng.standard_normal((n, n)) if cplx else None)
    x = DeviceMatrix.from_numpy(rng.standard_normal((n, 1)), rng.standard_normal((n, 1)))
    api = CuProceduresApi()
    outs = []
    try:
        for v, sh in enumerate(shapes(0)):
            check(lib.oaph_set_matvec_shape(0, v), "shape")
            y = DeviceMatrix(1, n, True, True)
            api.dotProduct(y, A, x)
            outs.append(y.to_complex())
    finally:
        check(lib.oaph_set_matvec_shape(0, 0), "shape")
    Ah = A.to_complex()
    xh = x.to_complex()
    want = Ah @ xh
    bound = 4 * n * 1.1e-16 * (np.abs(Ah) @ np.abs(xh))
    for o in outs:
        assert np.all(np.abs(o - want) <= bound)


def test_fp32_shapes_bit_identical():
    # the C1 construction at n = 2048 with the bench's k = 8, m = 48: plain
    # FP64 and the mixed flow return the dense top 8 for 100 of 100 start
    # vectors there (with k = 4, m = 32 on this spectrum IRAM itself returns a
    # wrong set for ~1 start vector in 5, plain FP64 included)
    lib = load()
    from oap_amd.api import CuHArnoldiCallback, Precision
    n, k, m = 2048, 8, 48
    A = DeviceMatrix(n, n)
    check(lib.oaph_synth_matrix(A.ptr, 1, 0x5EED0001, 1e-2, 500.0), "synth")
    s = CuHArnoldiCallback()
    s.setMatrix(A)
    s.setPrecision(Precision.MIXED_FP32)
    s.setSortType(0)
    ev = []
    try:
        for v, sh in enumerate(shapes(1)):
            check(lib.oaph_set_matvec_shape(1, v), "shape")
            s.setSeed(3)
            re, im = np.zeros(k), np.zeros(k)
            s.setOutputsEigenvalues(re, im)
            from oap_amd.api import MatrixInfo
            s.execute(m, k, MatrixInfo(True, True, n, n))
            st = s.stats()
            ev.append((re + 1j * im, st["matvecs"], st["matvecs_low"]))
    finally:
        check(lib.oaph_set_matvec_shape(1, 0), "shape")
    Ad = A.to_complex()
    want = sorted(np.linalg.eigvals(Ad), key=lambda z: -abs(z))[:k]
    for e, _, _ in ev:
        for w in want:
            assert min(abs(g - w) for g in e) <= 1e-10 * abs(want[0])


@pytest.mark.parametrize("fp32", [0, 1])
def test_shapes_in_solves_at_8192(fp32):
    # n = 8192 runs every shape with its row groups (workgroups of 16 or 32
    # rows), and the Arnoldi steps take the first Gram-Schmidt pass's partial
    # sums from the matvec's epilogue: each shape's solve must converge to
    # the same eigenvalues, with explicit residuals within the tolerance
    from oap_amd.api import CuHArnoldiCallback, CuProceduresApi, MatrixInfo, Precision
    lib = load()
    n, k = 8192, 4
    A = DeviceMatrix(n, n)
    check(lib.oaph_synth_matrix(A.ptr, 1, 0x5EED0001, 1e-2, 500.0), "synth")
    fro = CuProceduresApi().magnitude(A)
    runs = []
    try:
        for v, sh in enumerate(shapes(fp32)):
            check(lib.oaph_set_matvec_shape(fp32, v), "shape")
            s = CuHArnoldiCallback()
            s.setMatrix(A)
            s.setPrecision(Precision.MIXED_FP32 if fp32 else Precision.FP64)
            s.setSortType(0)
            s.setCheckTolerance(1e-10)
            s.setSeed(11)
            re, im = np.zeros(k), np.zeros(k)
            s.setOutputsEigenvalues(re, im)
            s.execute(24, k, MatrixInfo(True, True, n, n))
            for t in range(k):
                assert s.testOutcome(t) <= 1e-10 * fro, (sh, t, s.testOutcome(t))
            runs.append((sh, np.sort_complex(re + 1j * im)))
    finally:
        check(lib.oaph_set_matvec_shape(fp32, 0), "shape")
    ref = runs[0][1]
    for sh, ev in runs[1:]:
        assert np.all(np.abs(ev - ref) <= 1e-10 * np.abs(ref).max()), (sh, ev, ref)
