"""The in-HBM synthetic generators (oaph_synth_matrix) reproduce the
oracle's generators (same counter-based SplitMix64 + Box-Muller); the device
libm may differ from glibc by an ulp in log/cos, hence the 1e-14 bound."""
import numpy as np
import pytest

from oap_amd._lib import check, load
from oap_amd.api import DeviceMatrix

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [64, 257])
def test_c0_generator(oracle, n):
    A = DeviceMatrix(n, n, True, False)
    check(load().oaph_synth_matrix(A.ptr, 0, 0x5EED0000, 0.05, 0.0), "synth")
    got = A.download()[0]
    want = oracle.gen_c0(n, 0x5EED0000)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("n", [128, 1000])
def test_c1_generator(oracle, n):
    A = DeviceMatrix(n, n, True, True)
    check(load().oaph_synth_matrix(A.ptr, 1, 0x5EED0001, 1e-2, 500.0), "synth")
    got = A.to_complex()
    ar, ai = oracle.gen_c1(n, 0x5EED0001, 1e-2, 500.0)
    np.testing.assert_allclose(got, ar + 1j * ai, rtol=0, atol=1e-14)
