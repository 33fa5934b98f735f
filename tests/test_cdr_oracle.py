"""The C4 operator generator of the oracle (orc_gen_cdr) against the closed
form of its spectrum: the operator is the Kronecker sum of two tridiagonal
Toeplitz matrices, so its eigenvalues are
c + a_x + 2 sqrt(b_x d_x) cos(j pi h) + a_y + 2 sqrt(b_y d_y) cos(k pi h)
(FP32 rounding of the coefficients aside)."""
import numpy as np
import pytest


def closed_form(N, pe, c, alpha):
    h = 1.0 / (N + 1)
    j = np.arange(1, N + 1)

    def ev1(a):
        lo, up = -a / h**2 - pe / (2 * h), -a / h**2 + pe / (2 * h)
        return 2 * a / h**2 + 2 * np.sqrt(lo * up) * np.cos(j * np.pi * h)

    return np.sort((ev1(alpha)[:, None] + ev1(1.0)[None, :]).ravel() + c)


@pytest.mark.parametrize("pe,c,alpha", [(0.0, 0.0, 1.0), (0.0, 3.0, 0.5), (6.0, 2.0, 1.5),
                                        (10.0, 7.0, 2.0)])
def test_cdr_spectrum_closed_form(oracle, pe, c, alpha):
    N = 12
    A = oracle.gen_cdr(N, pe, c, alpha)
    ev = np.sort(np.linalg.eigvals(A).real)
    want = closed_form(N, pe, c, alpha)
    assert np.max(np.abs(ev - want)) <= 1e-5 * np.max(np.abs(want))


def test_cdr_structure(oracle):
    N = 8
    A = oracle.gen_cdr(N, 20.0, 5.0, 0.75)
    n = N * N
    assert np.count_nonzero(A) == n + 2 * (N - 1) * N * 2
    # every coefficient is an FP32 value
    assert np.array_equal(A, A.astype(np.float32).astype(np.float64))
    # x neighbours never wrap between grid rows
    for r in range(N, n, N):
        assert A[r, r - 1] == 0 and A[r - 1, r] == 0
