"""oaph_magnitude_noalloc (the allocation-free magnitude) against
oaph_magnitude on the same matrices, through the C ABI: two calls in a row,
the second on a smaller matrix, then the first again (the static scratch
keeps the previous call's contents and must not leak into the next).  The
device's free memory is the same before and after the allocation-free
calls."""
import ctypes as C

import numpy as np
import pytest

from oap_amd._lib import check, load
from oap_amd.api import CuProceduresApi, DeviceMatrix

pytestmark = pytest.mark.gpu


def free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_magnitude_noalloc_matches_magnitude():
    lib = load()
    rng = np.random.default_rng(77)
    big = DeviceMatrix(4096, 4096, True, True)
    check(lib.oaph_synth_matrix(big.ptr, 1, 0x5EED0001, 1e-2, 500.0), "synth")
    small_h = rng.standard_normal((333, 1001)) + 1j * rng.standard_normal((333, 1001))
    small = DeviceMatrix.from_numpy(small_h.real, small_h.imag)  # 1001 x 333, odd size
    re_only = DeviceMatrix.from_numpy(rng.standard_normal((17, 5)), None)
    out = DeviceMatrix(1, 1, True, False)
    out_ptr = C.c_void_p(out.ptr.contents.reValues)
    ref_api = CuProceduresApi()
    want = [ref_api.magnitude(m) for m in (big, small, re_only, big)]
    api = CuProceduresApi()  # created (and its buffers allocated) before the check
    api.synchronize()
    before = free_device_bytes()
    got = []
    for m in (big, small, re_only, big):
        check(lib.oaph_magnitude_noalloc(api._h, out_ptr, m.ptr), "magnitude_noalloc")
        api.synchronize()
        got.append(float(out.download()[0][0, 0]))
    after = free_device_bytes()
    for g, w in zip(got, want):
        # same grid and summation order; the square root on the device is
        # within one ulp of the host's
        assert abs(g - w) <= 2.0 ** -52 * w, (g, w)
    assert got[0] == got[3]
    assert after == before, (before, after)
    # the magnitude against the numpy value of the host matrix
    assert abs(got[1] - np.linalg.norm(small_h)) <= 1e-13 * np.linalg.norm(small_h)


def test_magnitude_noalloc_argument_errors():
    lib = load()
    api = CuProceduresApi()
    m = DeviceMatrix.from_numpy(np.ones((4, 4)), None)
    assert lib.oaph_magnitude_noalloc(api._h, None, m.ptr) != 0
    assert lib.oaph_magnitude_noalloc(None, C.c_void_p(1), m.ptr) != 0
