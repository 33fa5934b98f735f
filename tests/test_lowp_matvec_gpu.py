"""The mixed solver's reduced-precision operator copies and their matvecs.

FP32 copy: numpy's float32 rounding of each plane; its matvec multiplies and
adds in FP64, so y = A32 x up to FP64 summation order (bound 4 n eps
(|A32| |x|) per element against numpy's FP64 product of the same matrix).

FP16 copy: the FP16 rounding (numpy's float64 -> float16 conversion, round
to nearest even, one rounding) of 2^s a, s = 14 - floor(log2 max|a|); its
matvec rounds x to FP32 and forms products and per-lane sums in
FP32 (2n/64 terms a lane), so y = 2^-s A16 x32 within (2n/64 + 2) 2^-24
(|A16| |x32|) per element.  The scaling and rounding are pinned below against
hand-worked values.

Every FP16 launch shape runs at 8192 rows (row groups, the default path of
the C1 operator), the G = 1 fallback at 4096 rows and the two-rows-per-wave
path at 512 rows.
"""
import numpy as np
import pytest

from oap_amd._lib import check, load
from oap_amd.api import CuHArnoldiCallback, DeviceMatrix, OapError

EPS = np.finfo(np.float64).eps
U32 = 2.0 ** -24


def f16_shift(A):
    mx = max(np.max(np.abs(A.real)), np.max(np.abs(A.imag)) if np.iscomplexobj(A) else 0.0)
    return 14 - int(np.floor(np.log2(mx)))


def f16_round(a, shift):
    """the copy's value of each entry, back in A's scale"""
    h = (np.asarray(a, dtype=np.float64) * 2.0 ** shift).astype(np.float16)
    return h.astype(np.float64) * 2.0 ** -shift


def test_f16_scaling_and_rounding_known_values():
    # ties to even at FP16's 11-bit significand, a lower bit breaking a tie,
    # subnormals, and the scale: the largest entry lands in [2^14, 2^15)
    # 1 + 2^-11 + 2^-40 rounds up at once; through FP32 it would first become
    # the tie 1 + 2^-11 and then round down to 1 (the copy rounds once)
    vals = np.array([1.0, 1 + 2**-11, 1 + 3 * 2**-11, 1 + 2**-11 + 2**-20, 1 + 2**-11 + 2**-40,
                     2.0**-24, 2.0**-26])
    want = np.array([1.0, 1.0, 1 + 2**-9, 1 + 2**-10, 1 + 2**-10, 2.0**-24, 0.0])
    assert np.array_equal(f16_round(vals, 0), want)
    A = np.array([[0.75, -3e-3], [1e-9, 2.0**-40]])
    s = f16_shift(A)
    assert s == 15 and 2**14 <= 0.75 * 2**s < 2**15
    r = f16_round(A, s)
    assert r[0, 0] == 0.75 and r[1, 1] == 0.0  # 2^-40 * 2^15 is below FP16's range
    assert abs(r[0, 1] + 3e-3) <= 2**-11 * 3e-3 and r[1, 0] > 0


def solver_for(A):
    s = CuHArnoldiCallback()
    cplx = np.iscomplexobj(A)
    dA = DeviceMatrix.from_numpy(A.real, A.imag if cplx else None)
    s.setMatrix(dA)
    return s, dA


def run_product(s, x, precision):
    n = x.size
    dx = DeviceMatrix.from_numpy(x.real.reshape(n, 1), x.imag.reshape(n, 1))
    dy = DeviceMatrix(1, n)
    s.applyOperator(precision, dx, dy)
    return dy.to_complex()[:, 0]


def within(got, want, bound):
    err = np.abs(got - want)
    return float(np.max(err / (bound + 1e-300)))


def abs_prod(A, x):
    return (np.abs(A.real) + np.abs(A.imag)) @ (np.abs(x.real) + np.abs(x.imag))


def f16_reference(A, x):
    s = f16_shift(A)
    A16 = f16_round(A.real, s) + (1j * f16_round(A.imag, s) if np.iscomplexobj(A) else 0)
    x32 = x.real.astype(np.float32).astype(np.float64) + 1j * x.imag.astype(np.float32).astype(np.float64)
    n = x.size
    return A16 @ x32, (2 * n / 64 + 2) * U32 * abs_prod(A16, x32)


@pytest.mark.gpu
@pytest.mark.parametrize("n,cplx", [(512, True), (520, True), (4096, True), (4096, False)])
def test_lowp_matvec_against_rounded_operator(n, cplx):
    rng = np.random.default_rng(n + cplx)
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    if cplx:
        A = A + 1j * rng.standard_normal((n, n)) / np.sqrt(n)
    np.fill_diagonal(A, np.exp(-np.arange(n) / 500.0))
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    s, _ = solver_for(A)
    Ac = A.astype(np.complex128)
    y64 = run_product(s, x, 0)
    assert within(y64, Ac @ x, 4 * n * EPS * abs_prod(Ac, x)) <= 1
    A32 = (A.real.astype(np.float32).astype(np.float64) +
           (1j * A.imag.astype(np.float32).astype(np.float64) if cplx else 0))
    y32 = run_product(s, x, 1)
    assert within(y32, A32 @ x, 4 * n * EPS * abs_prod(A32, x)) <= 1
    y16 = run_product(s, x, 2)
    want, bound = f16_reference(A, x)
    assert within(y16, want, bound) <= 1
    # the copies are not the operator: their products differ from the FP64 one
    assert np.max(np.abs(y16 - y64)) > np.max(np.abs(y32 - y64)) > 0


@pytest.mark.gpu
def test_f16_copy_needs_rows_of_eight():
    A = np.eye(516) + 0j
    s, _ = solver_for(A)
    x = np.ones(516) + 0j
    assert np.allclose(run_product(s, x, 1), x)
    with pytest.raises(OapError):
        run_product(s, x, 2)


@pytest.mark.gpu
def test_f16_matvec_every_shape_at_8192():
    lib = load()
    n = 8192
    dA = DeviceMatrix(n, n, True, True)
    check(lib.oaph_synth_matrix(dA.ptr, 1, 0x5EED0001, 1e-2, 500.0), "synth")
    re, im = dA.download()
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    want, bound = f16_reference(re + 1j * im, x)
    del re, im
    s = CuHArnoldiCallback()
    s.setMatrix(dA)
    try:
        for v in range(8):
            check(lib.oaph_set_matvec_shape(2, v), "shape")
            assert within(run_product(s, x, 2), want, bound) <= 1, v
    finally:
        check(lib.oaph_set_matvec_shape(2, 0), "shape")
    with pytest.raises(OapError):
        s.applyOperator(3, DeviceMatrix(1, n), DeviceMatrix(1, n))
