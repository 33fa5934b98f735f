"""The oracle (oracle.c) against the reference's own host code (oracle/_ref).

Pins the C restatement: each procedure must agree bit for bit with the
reference compiled from its sources (HostProcedures on the CPU emulator and
the reference device functions run through its HostKernel emulator), except
where the oracle documents a divergence.
"""
import numpy as np
import pytest

from oracle.pyoracle import Planes

SHAPES = [(1, 1), (3, 5), (7, 4), (16, 16), (33, 17)]


def cplx(rng, r, c):
    return rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))


def pair(z, is_re=True, is_im=True):
    return Planes.of(z, is_re, is_im), Planes.of(z, is_re, is_im)


@pytest.mark.parametrize("w,h", [(20, 20), (256, 256), (1, 16438), (16384, 2), (1024, 1024),
                                 (2048, 2048), (5123, 2133), (400, 20), (1, 1), (31, 33),
                                 (1, 8192), (48, 8192), (8192, 48), (64, 64)])
def test_threads_blocks_matches_reference(oracle, ref, w, h):
    assert oracle.set_threads_blocks(w, h) == ref.set_threads_blocks(w, h)


@pytest.mark.parametrize("planes", [(True, True), (True, False)])
@pytest.mark.parametrize("m,k,n", [(1, 1, 1), (4, 3, 2), (5, 7, 3), (16, 16, 16), (9, 31, 1)])
def test_dot_product_bit_exact(oracle, ref, rng, planes, m, k, n):
    a, b = cplx(rng, m, k), cplx(rng, k, n)
    A, B = Planes.of(a, *planes), Planes.of(b, *planes)
    o1, o2 = Planes(m, n, *planes), Planes(m, n, *planes)
    assert oracle.dot_product(o1, A, B) == 0
    ref.dot_product(o2, A, B)
    np.testing.assert_array_equal(o1.re, o2.re)
    if planes[1]:
        np.testing.assert_array_equal(o1.im, o2.im)


@pytest.mark.parametrize("planes", [(True, True), (True, False), (False, True)])
def test_dot_product_dims_bit_exact(oracle, ref, rng, planes):
    a, b = cplx(rng, 6, 9), cplx(rng, 8, 7)
    A, B = Planes.of(a, *planes), Planes.of(b, *planes)
    dims = [[5, 4], [7, 4], [5, 7]]
    o1, o2 = Planes(6, 6, *planes), Planes(6, 6, *planes)
    assert oracle.dot_product(o1, A, B, dims) == 0
    ref.dot_product(o2, A, B, dims)
    for x, y in ((o1.re, o2.re), (o1.im, o2.im)):
        if x is not None:
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("r,c", SHAPES)
def test_transpose_bit_exact(oracle, ref, rng, r, c):
    a = cplx(rng, r, c)
    o1, o2 = Planes(c, r), Planes(c, r)
    oracle.transpose(o1, Planes.of(a))
    ref.transpose(o2, Planes.of(a))
    np.testing.assert_array_equal(o1.z(), o2.z())


@pytest.mark.parametrize("r,c", [(3, 5), (16, 16), (7, 4)])
def test_conjugate_transpose_bit_exact(oracle, ref, rng, r, c):
    a = cplx(rng, r, c)
    o1, o2 = Planes(c, r), Planes(c, r)
    oracle.transpose(o1, Planes.of(a), conj=True)
    ref.transpose(o2, Planes.of(a), conj=True)
    np.testing.assert_array_equal(o1.z(), o2.z())
    np.testing.assert_array_equal(o1.z(), a.conj().T)


@pytest.mark.parametrize("op", ["add", "substract", "hadamard", "hadamard_vec"])
@pytest.mark.parametrize("planes", [(True, True), (True, False), (False, True)])
def test_binary_bit_exact(oracle, ref, rng, op, planes):
    r, c = 9, 6
    a = cplx(rng, r, c)
    b = cplx(rng, r, 1) if op == "hadamard_vec" else cplx(rng, r, c)
    A, B = Planes.of(a, *planes), Planes.of(b, *planes)
    o1, o2 = Planes(r, c, *planes), Planes(r, c, *planes)
    oracle.binary(op, o1, A, B)
    ref.binary(op, o2, A, B)
    np.testing.assert_array_equal(o1.z(), o2.z())


@pytest.mark.parametrize("re_only", [True, False])
def test_multiply_constant_bit_exact(oracle, ref, rng, re_only):
    a = cplx(rng, 5, 7)
    o1, o2 = Planes(5, 7), Planes(5, 7)
    oracle.multiply_constant(o1, Planes.of(a), 0.37, -1.5, re_only)
    ref.multiply_constant(o2, Planes.of(a), 0.37, -1.5, re_only)
    np.testing.assert_array_equal(o1.z(), o2.z())
    if re_only:  # the im plane is not written (CuMatrixKernels.cu:119-123)
        assert np.all(o1.im == 0)


def test_set_diagonal_and_identity(oracle, ref):
    o1, o2 = Planes(6, 6), Planes(6, 6)
    o1.re[:] = 5
    o2.re[:] = 5
    oracle.set_diagonal(o1, 2.5, -1.0)
    ref.set_diagonal(o2, 2.5, -1.0)
    np.testing.assert_array_equal(o1.z(), o2.z())
    i2 = Planes(5, 5)
    ref.set_identity(i2)
    i1 = Planes(5, 5)
    oracle.set_diagonal(i1, 1.0, 0.0)
    np.testing.assert_array_equal(i1.z(), i2.z())


def test_set_get_vector_bit_exact(oracle, ref, rng):
    V = cplx(rng, 10, 6)
    v = cplx(rng, 10, 1)
    a1, a2 = Planes.of(V), Planes.of(V)
    oracle.set_vector(a1, 3, Planes.of(v), 7)
    ref.set_vector(a2, 3, Planes.of(v), 7)
    np.testing.assert_array_equal(a1.z(), a2.z())
    g1, g2 = Planes(10, 1), Planes(10, 1)
    oracle.get_vector(g1, 8, Planes.of(V), 4)
    ref.get_vector(g2, 8, Planes.of(V), 4)
    np.testing.assert_array_equal(g1.z(), g2.z())


def test_tensor_product_bit_exact(oracle, ref, rng):
    a, b = rng.standard_normal((3, 2)), rng.standard_normal((2, 4))
    dims = [[8, 6], [2, 3], [4, 2]]
    o1, o2 = Planes(6, 8, True, False), Planes(6, 8, True, False)
    oracle.tensor_product(o1, Planes.of(a, True, False), Planes.of(b, True, False), dims)
    ref.tensor_product(o2, Planes.of(a, True, False), Planes.of(b, True, False), dims)
    np.testing.assert_array_equal(o1.re, o2.re)
    np.testing.assert_array_equal(o1.re, np.kron(a, b))


@pytest.mark.parametrize("r,c", [(1, 1), (10, 1), (33, 17), (64, 64)])
def test_magnitude_and_sum(oracle, ref, rng, r, c):
    a = cplx(rng, r, c)
    m1 = oracle.magnitude(Planes.of(a))
    m2 = ref.magnitude(Planes.of(a))
    n = r * c
    assert abs(m1 - m2) <= 4 * n * 2.2e-16 * m1  # block-tree order vs index order
    assert abs(m1 - np.linalg.norm(a)) <= 4 * n * 2.2e-16 * m1
    s1 = oracle.sum(Planes.of(a))
    s2 = ref.sum(Planes.of(a))
    bound = 4 * n * 2.2e-16 * np.abs(a).sum()
    assert abs(s1[0] - s2[0]) <= bound and abs(s1[1] - s2[1]) <= bound


def test_is_upper_triangular(oracle, ref):
    h = np.triu(np.ones((5, 5))) + 0j
    h[2, 1] = 5e-4
    assert oracle.is_upper_triangular(Planes.of(h)) and ref.is_upper_triangular(Planes.of(h))
    h[3, 2] = 2e-3
    assert not oracle.is_upper_triangular(Planes.of(h))
    assert not ref.is_upper_triangular(Planes.of(h))


@pytest.mark.parametrize("m", [3, 5, 8])
def test_qrgr_real_matches_reference_odd_rotation_count(oracle, ref, rng, m):
    # Hessenberg input: m-1 rotations; m even gives an odd count, where the
    # reference's ping-pong keeps the final rotation (for even counts it
    # drops it -- a divergence the oracle fixes)
    h = np.triu(rng.standard_normal((m, m)), -1)
    Q1, R1 = Planes(m, m, True, False), Planes(m, m, True, False)
    assert oracle.qrgr(Q1, R1, Planes.of(h, True, False)) == 0
    np.testing.assert_allclose(Q1.re @ R1.re, h, atol=1e-13)
    np.testing.assert_allclose(Q1.re.T @ Q1.re, np.eye(m), atol=1e-13)
    assert np.allclose(np.tril(R1.re, -1), 0, atol=1e-13)
    if (m - 1) % 2 == 1:
        Q2, R2 = Planes(m, m, True, False), Planes(m, m, True, False)
        ref.qrgr(Q2, R2, Planes.of(h, True, False))
        # the reference's device function uses sqrtf for the rotation norm
        np.testing.assert_allclose(R1.re, R2.re, atol=1e-6)
        np.testing.assert_allclose(Q1.re, Q2.re, atol=1e-6)
