"""HIP procedures (liboaph.so) against the oracle on the same seeded inputs.

Bit-exact for data movement and per-plane element-wise products; for sums
and products with a reduction the bound is gamma_K * sum |a||b| (different
summation order), written per test.
"""
import numpy as np
import pytest

from oap_amd.api import DeviceMatrix, OapError
from oracle.pyoracle import Planes

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def cplx(rng, r, c):
    return rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))


def dev(z, is_re=True, is_im=True):
    z = np.asarray(z)
    if z.ndim == 1:
        z = z.reshape(-1, 1)
    return DeviceMatrix.from_numpy(z.real if is_re else None, z.imag if is_im else None)


def dout(r, c, is_re=True, is_im=True):
    return DeviceMatrix(c, r, is_re, is_im)


def host(m):
    re, im = m.download()
    z = np.zeros((m.rows, m.columns), complex)
    if re is not None:
        z.real = re
    if im is not None:
        z.imag = im
    return z


@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (5, 7, 3), (64, 64, 64), (130, 70, 33),
                                   (257, 129, 65), (8, 300, 1), (300, 8, 1), (1000, 1000, 1),
                                   (4096, 4096, 1)])
def test_dot_product_complex(api, oracle, rng, m, k, n):
    a, b = cplx(rng, m, k), cplx(rng, k, n)
    out = dout(m, n)
    api.dotProduct(out, dev(a), dev(b))
    got = host(out)
    if m * k * n <= 300_000:
        o = Planes(m, n)
        assert oracle.dot_product(o, Planes.of(a), Planes.of(b)) == 0
        want = o.z()
    else:
        want = a @ b
    bound = 2 * k * EPS * (np.abs(a.real) + np.abs(a.imag)) @ (np.abs(b.real) + np.abs(b.imag))
    assert np.all(np.abs(got - want) <= bound + 1e-300)


@pytest.mark.parametrize("m,k,n", [(5, 7, 3), (100, 300, 1), (64, 64, 64)])
def test_dot_product_real(api, oracle, rng, m, k, n):
    a, b = rng.standard_normal((m, k)), rng.standard_normal((k, n))
    out = dout(m, n, True, False)
    api.dotProduct(out, dev(a, True, False), dev(b, True, False))
    o = Planes(m, n, True, False)
    oracle.dot_product(o, Planes.of(a, True, False), Planes.of(b, True, False))
    bound = 2 * k * EPS * np.abs(a) @ np.abs(b)
    assert np.all(np.abs(host(out).real - o.re) <= bound)


def test_dot_product_im_only_rejected(api, rng):
    a = cplx(rng, 3, 3)
    with pytest.raises(OapError):
        api.dotProduct(dout(3, 3, False, True), dev(a, False, True), dev(a, False, True))


def test_dot_product_shape_error(api, rng):
    with pytest.raises(OapError):
        api.dotProduct(dout(3, 3), dev(cplx(rng, 3, 4)), dev(cplx(rng, 3, 3)))


@pytest.mark.parametrize("planes", [(True, True), (True, False), (False, True)])
def test_dot_product_dims(api, oracle, rng, planes):
    a, b = cplx(rng, 6, 9), cplx(rng, 8, 7)
    dims = [[5, 4], [7, 4], [5, 7]]
    out = dout(6, 6, *planes)
    api.setZeroMatrix(out)
    api.dotProduct(out, dev(a, *planes), dev(b, *planes), dims)
    o = Planes(6, 6, *planes)
    oracle.dot_product(o, Planes.of(a, *planes), Planes.of(b, *planes), dims)
    np.testing.assert_allclose(host(out), o.z(), rtol=0, atol=1e-13)


@pytest.mark.parametrize("r,c", [(1, 1), (1, 9), (9, 1), (3, 5), (32, 32), (33, 65), (48, 8192),
                                 (8192, 48), (1000, 999)])
@pytest.mark.parametrize("conj", [False, True])
def test_transpose_bit_exact(api, oracle, rng, r, c, conj):
    a = cplx(rng, r, c)
    out = dout(c, r)
    (api.conjugateTranspose if conj else api.transpose)(out, dev(a))
    o = Planes(c, r)
    oracle.transpose(o, Planes.of(a), conj)
    np.testing.assert_array_equal(host(out), o.z())


@pytest.mark.parametrize("op", ["add", "substract", "hadamardProduct", "hadamardProductVec"])
@pytest.mark.parametrize("planes", [(True, True), (True, False), (False, True)])
def test_binary_bit_exact(api, oracle, rng, op, planes):
    r, c = 37, 129
    a = cplx(rng, r, c)
    b = cplx(rng, r, 1) if op == "hadamardProductVec" else cplx(rng, r, c)
    out = dout(r, c, *planes)
    getattr(api, op)(out, dev(a, *planes), dev(b, *planes))
    o = Planes(r, c, *planes)
    name = {"add": "add", "substract": "substract", "hadamardProduct": "hadamard",
            "hadamardProductVec": "hadamard_vec"}[op]
    oracle.binary(name, o, Planes.of(a, *planes), Planes.of(b, *planes))
    np.testing.assert_array_equal(host(out), o.z())


def test_multiply_constants(api, oracle, rng):
    a = cplx(rng, 50, 3)
    out = dout(50, 3)
    api.setZeroMatrix(out)
    api.multiplyReConstant(out, dev(a), 0.3)
    o = Planes(50, 3)
    oracle.multiply_constant(o, Planes.of(a), 0.3, 0.0, re_only=True)
    np.testing.assert_array_equal(host(out), o.z())
    api.multiplyConstant(out, dev(a), 0.3, -2.0)
    oracle.multiply_constant(o, Planes.of(a), 0.3, -2.0)
    np.testing.assert_array_equal(host(out), o.z())


def test_set_diagonal_identity_zero(api):
    m = dout(7, 5)
    api.setDiagonal(m, 2.0, -3.0)
    z = host(m)
    want = np.zeros((7, 5), complex)
    for i in range(5):
        want[i, i] = 2 - 3j
    np.testing.assert_array_equal(z, want)
    api.setIdentity(m)
    np.testing.assert_array_equal(host(m), np.eye(7, 5))
    api.setZeroMatrix(m)
    assert np.all(host(m) == 0)


def test_set_get_vector(api, oracle, rng):
    V = cplx(rng, 100, 12)
    v = cplx(rng, 100, 1)
    dV = dev(V)
    api.setVector(dV, 5, dev(v), 70)
    o = Planes.of(V)
    oracle.set_vector(o, 5, Planes.of(v), 70)
    np.testing.assert_array_equal(host(dV), o.z())
    g = dout(100, 1)
    api.setZeroMatrix(g)
    api.getVector(g, 100, dV, 11)
    np.testing.assert_array_equal(host(g)[:, 0], o.z()[:, 11])
    with pytest.raises(OapError):
        api.setVector(dV, 12, dev(v), 100)  # column out of range


def test_tensor_product(api, rng):
    a, b = rng.standard_normal((3, 2)), rng.standard_normal((2, 4))
    out = dout(6, 8, True, False)
    api.tensorProduct(out, dev(a, True, False), dev(b, True, False),
                      [[8, 6], [2, 3], [4, 2]])
    np.testing.assert_array_equal(host(out).real, np.kron(a, b))


@pytest.mark.parametrize("n", [1, 2, 63, 64, 1000, 16001, 1 << 20, (1 << 22) + 3])
def test_magnitude_and_sum(api, oracle, rng, n):
    a = cplx(rng, n, 1)
    d = dev(a)
    mag = api.magnitude(d)
    want = np.sqrt(np.sum(np.abs(a) ** 2))
    assert abs(mag - want) <= 4 * np.sqrt(n) * EPS * want + n * EPS * want * 0.01
    s = api.sum(d)
    bound = 2 * np.log2(n + 1) * EPS * np.abs(a).sum() + 1e-300
    assert abs(s[0] - a.real.sum()) <= bound + n * EPS * np.abs(a.real).sum() * 0.01
    assert abs(s[1] - a.imag.sum()) <= bound + n * EPS * np.abs(a.imag).sum() * 0.01


def test_magnitude_and_sum_2p26_exact(api):
    # the C1 standalone size, 2^26 complex elements: integer values whose
    # every partial sum is an integer below 2^53, so any summation order
    # gives the exact result
    n = 1 << 26
    i = np.arange(n, dtype=np.int64)
    re = (i % 97).astype(np.float64)
    im = (i % 5).astype(np.float64)
    d = DeviceMatrix.from_numpy(re.reshape(1 << 13, 1 << 13), im.reshape(1 << 13, 1 << 13))
    assert api.sum(d) == (re.sum(), im.sum())
    sq = float(np.dot(re, re) + np.dot(im, im))
    assert api.magnitude(d) == np.sqrt(sq)


def test_magnitude_and_sum_2p26_random(api, rng):
    # 2^26 + 3 complex elements: the odd tail after the 16-byte pairs
    n = (1 << 26) + 3
    a = cplx(rng, n, 1)
    d = dev(a)
    want = np.sqrt(np.sum(np.abs(a) ** 2))
    assert abs(api.magnitude(d) - want) <= 4 * np.sqrt(n) * EPS * want
    s = api.sum(d)
    bound = 2 * np.log2(n + 1) * EPS * np.abs(a).sum()
    assert abs(s[0] - a.real.sum()) <= bound and abs(s[1] - a.imag.sum()) <= bound


def test_magnitude_closed_forms(api):
    # oapMatrixCudaTest.cpp:418-443 (1..10 -> sqrt(385)), :473-502 (one 3 among 16001)
    v = np.arange(1, 11, dtype=float)
    assert api.magnitude(dev(v, True, False)) == pytest.approx(np.sqrt(385), rel=1e-15)
    w = np.zeros(16001)
    w[1234] = 3.0
    assert api.magnitude(dev(w, True, False)) == 3.0


def test_sum_integers_exact(api):
    # oapSumTests.cpp:67-276: integer-valued matrices sum exactly
    for r, c in [(203, 103), (33, 17), (1024, 2024)]:
        a = np.arange(r * c, dtype=float).reshape(r, c) % 97
        s = api.sum(dev(a + 1j * (a % 5)))
        assert s == (a.sum(), (a % 5).sum())


@pytest.mark.parametrize("m", [3, 4, 16, 32, 48])
@pytest.mark.parametrize("cplx_in", [False, True])
def test_qrgr(api, oracle, rng, m, cplx_in):
    h = np.triu(rng.standard_normal((m, m)), -1)
    if cplx_in:
        h = h + 1j * np.triu(rng.standard_normal((m, m)), -1)
    planes = (True, cplx_in)
    Q, R = dout(m, m, *planes), dout(m, m, *planes)
    api.QRGR(Q, R, dev(h, *planes))
    q, r = host(Q), host(R)
    Qo, Ro = Planes(m, m, *planes), Planes(m, m, *planes)
    oracle.qrgr(Qo, Ro, Planes.of(h, *planes))
    np.testing.assert_allclose(q, Qo.z(), atol=1e-12)
    np.testing.assert_allclose(r, Ro.z(), atol=1e-12)
    np.testing.assert_allclose(q @ r, h, atol=1e-12)
    np.testing.assert_allclose(q.conj().T @ q, np.eye(m), atol=1e-12)


def test_is_upper_triangular(api):
    h = np.triu(np.ones((6, 6)))
    h[3, 2] = 5e-4
    assert api.isUpperTriangular(dev(h, True, False))
    h[4, 3] = 0.01
    assert not api.isUpperTriangular(dev(h, True, False))


def test_threads_blocks_golden():
    from oap_amd.api import set_threads_blocks
    assert set_threads_blocks(1, 16438) == ((1, 514), (1, 32))
