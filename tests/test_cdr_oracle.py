"""The C4 operator generator of the oracle (orc_gen_cdr) against the closed
form of its spectrum: the operator is the Kronecker sum of two tridiagonal
Toeplitz matrices, so its eigenvalues are
d + 2 sqrt(xm xp) cos(i pi h) + 2 sqrt(ym yp) cos(j pi h) with the
generator's FP32-rounded coefficients (learned.closed_form_smallest)."""
import numpy as np
import pytest

from oap_amd.learned import closed_form_smallest


@pytest.mark.parametrize("pe,c,alpha", [(0.0, 0.0, 1.0), (0.0, 3.0, 0.5), (6.0, 2.0, 1.5),
                                        (10.0, 7.0, 2.0), (3.0, 5.0, 0.6)])
def test_cdr_spectrum_closed_form(oracle, pe, c, alpha):
    N = 12
    A = oracle.gen_cdr(N, pe, c, alpha)
    ev = np.sort(np.linalg.eigvals(A).real)
    want = closed_form_smallest(pe, c, alpha, N * N, N)
    assert np.max(np.abs(ev - want)) <= 1e-12 * np.max(np.abs(want))
    # exact coefficients would miss by the FP32 rounding of the diagonal
    h = 1.0 / (N + 1)
    j = np.arange(1, N + 1)

    def ev1(a):
        lo, up = -a / h**2 - pe / (2 * h), -a / h**2 + pe / (2 * h)
        return 2 * a / h**2 + 2 * np.sqrt(lo * up) * np.cos(j * np.pi * h)

    exact = np.sort((ev1(alpha)[:, None] + ev1(1.0)[None, :]).ravel() + c)
    assert np.max(np.abs(exact - want)) <= 1e-5 * np.max(np.abs(want))


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
