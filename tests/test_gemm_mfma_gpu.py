"""The matrix-core complex product (gemm.hip, v_mfma_f64_16x16x4_f64) behind
dotProduct / dotProduct(dims) for W, H >= 256, against numpy's complex128
product and against the vector-FMA tiles on the same operands.

Bar: |C - C_ref| <= 4 K eps (|A| |B|) elementwise (the FP64 summation bound
of a K-term product; the two paths sum in different orders), exact zeros
where an operand plane is absent.  Shapes that are not multiples of the
128 x 128 tile exercise the edge masks.
"""
import numpy as np
import pytest

from oap_amd._lib import check, load
from oap_amd.api import CuProceduresApi, DeviceMatrix

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def bound(a, b):
    return 4 * a.shape[1] * EPS * (np.abs(a) @ np.abs(b)) + 1e-300


def dev(z, re=True, im=True):
    return DeviceMatrix.from_numpy(np.ascontiguousarray(z.real) if re else None,
                                   np.ascontiguousarray(z.imag) if im else None)


@pytest.fixture
def api():
    # every product on the matrix cores (small shapes test the edge masks)
    check(load().oaph_set_gemm_path(2), "gemm path")
    yield CuProceduresApi()
    check(load().oaph_set_gemm_path(1), "gemm path")


@pytest.mark.parametrize("H,K,W", [(256, 256, 256), (300, 197, 333), (512, 64, 1024), (5, 3, 7)])
def test_complex_product(api, rng, H, K, W):
    a = rng.standard_normal((H, K)) + 1j * rng.standard_normal((H, K))
    b = rng.standard_normal((K, W)) + 1j * rng.standard_normal((K, W))
    out = DeviceMatrix(W, H)
    api.dotProduct(out, dev(a), dev(b))
    got = out.to_complex()
    ref = a @ b
    bd = bound(a, b)
    assert np.all(np.abs(got.real - ref.real) <= bd) and np.all(np.abs(got.imag - ref.imag) <= bd)
    # the vector-FMA tiles on the same operands
    check(load().oaph_set_gemm_path(0), "gemm path")
    out2 = DeviceMatrix(W, H)
    api.dotProduct(out2, dev(a), dev(b))
    got2 = out2.to_complex()
    check(load().oaph_set_gemm_path(2), "gemm path")
    assert np.all(np.abs(got - got2) <= 2 * bd)


def test_real_left_operand_complex_right(api, rng):
    # a left operand without an imaginary plane reads as A + 0i
    H, K, W = 384, 160, 256
    a = rng.standard_normal((H, K))
    b = rng.standard_normal((K, W)) + 1j * rng.standard_normal((K, W))
    out = DeviceMatrix(W, H)
    api.dotProduct(out, DeviceMatrix.from_numpy(a, None), dev(b))
    got = out.to_complex()
    ref = a @ b
    bd = bound(a, b)
    assert np.all(np.abs(got.real - ref.real) <= bd) and np.all(np.abs(got.imag - ref.imag) <= bd)


def test_real_planes_only(api, rng):
    H, K, W = 320, 96, 288
    a, b = rng.standard_normal((H, K)), rng.standard_normal((K, W))
    out = DeviceMatrix(W, H, True, False)
    api.dotProduct(out, DeviceMatrix.from_numpy(a, None), DeviceMatrix.from_numpy(b, None))
    re, im = out.download()
    assert im is None
    assert np.all(np.abs(re - a @ b) <= bound(a, b))


def test_dims_window_and_im_only_semantics(api, rng):
    # dims: a W x H window of the output from the leading K columns of A and
    # rows of B (CuDotProductDimProcedures.h:30-48); an im-only output holds
    # -sum a_im b_im, as the reference's kernel writes it
    a = rng.standard_normal((400, 300)) + 1j * rng.standard_normal((400, 300))
    b = rng.standard_normal((300, 500)) + 1j * rng.standard_normal((300, 500))
    W, H, K = 272, 384, 200
    dims = [[W, H], [K, H], [W, K]]
    out = DeviceMatrix(500, 400)
    api.dotProduct(out, dev(a), dev(b), dims)
    got = out.to_complex()
    ref = a[:H, :K] @ b[:K, :W]
    bd = bound(a[:H, :K], b[:K, :W])
    assert np.all(np.abs(got[:H, :W] - ref) <= 2 * bd)
    assert np.all(got[H:, :] == 0) and np.all(got[:, W:] == 0)
    out_im = DeviceMatrix(500, 400, False, True)
    api.dotProduct(out_im, dev(a, False, True), dev(b, False, True), dims)
    _, im = out_im.download()
    ref_im = -(a.imag[:H, :K] @ b.imag[:K, :W])
    assert np.all(np.abs(im[:H, :W] - ref_im) <= bound(a.imag[:H, :K], b.imag[:K, :W]))


def test_full_size_4096(api, rng):
    # the 4096^3 complex product, checked on sampled rows against numpy
    n = 4096
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out = DeviceMatrix(n, n)
    api.dotProduct(out, dev(a), dev(b))
    got = out.to_complex()
    rows = rng.choice(n, 12, replace=False)
    ref = a[rows] @ b
    bd = bound(a[rows], b)
    assert np.all(np.abs(got[rows].real - ref.real) <= bd)
    assert np.all(np.abs(got[rows].imag - ref.imag) <= bd)
