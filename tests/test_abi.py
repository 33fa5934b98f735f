"""CPU checks of the C ABI: liboaph.so loads and exports every symbol that
include/oaph.h declares; the host-only launch-geometry map is bit-exact with
the oracle.  No GPU compute is called."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "oaph.h").read_text()
    return sorted(set(re.findall(r"\b(oaph_[a-z0-9_]+)\s*\(", text)) -
                  {"oaph_multiply_fn", "oaph_check_fn", "oaph_status"})


def test_library_exports_every_declared_symbol():
    from oap_amd._lib import declared_symbols, load
    lib = load()
    names = declared()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding describes every declared entry point
    assert sorted(declared_symbols()) == names


def test_status_strings():
    from oap_amd._lib import load
    lib = load()
    assert lib.oaph_status_string(0) == b"ok"
    assert lib.oaph_status_string(2) == b"matrix dimensions do not match"


@pytest.mark.parametrize("w,h,limit", [(20, 20, 1024), (256, 256, 1024), (1, 16438, 1024),
                                       (16384, 2, 1024), (1024, 1024, 1024), (5123, 2133, 1024),
                                       (400, 20, 1024), (7, 9, 4), (203, 103, 9), (33, 17, 36)])
def test_threads_blocks_library_matches_oracle(oracle, w, h, limit):
    from oap_amd.api import set_threads_blocks
    assert set_threads_blocks(w, h, limit) == oracle.set_threads_blocks(w, h, limit)


def test_threads_blocks_golden_table():
    # oapHostTests/oapThreadsMapperTest.cpp:58-167
    from oap_amd.api import set_threads_blocks
    assert set_threads_blocks(20, 20) == ((1, 1), (20, 20))
    assert set_threads_blocks(256, 256) == ((8, 8), (32, 32))
    assert set_threads_blocks(1, 16438) == ((1, 514), (1, 32))
    b, t = set_threads_blocks(16384, 2)
    assert t == (32, 2) and b[0] * t[0] >= 16384
    assert set_threads_blocks(1024, 1024) == ((32, 32), (32, 32))
    assert set_threads_blocks(2048, 2048) == ((64, 64), (32, 32))
    b, t = set_threads_blocks(5123, 2133)
    assert t[0] * t[1] == 1024 and b[0] * t[0] >= 5123 and b[1] * t[1] >= 2133
    b, t = set_threads_blocks(400, 20)
    assert b[0] * t[0] >= 400 and b[1] * t[1] >= 20


def test_cpp_facade_example_is_built():
    # the facade header compiles as C++17 against the C ABI only
    exe = ROOT / "tests" / "cpp" / "facade_example"
    assert exe.exists()


def test_batch_variant_shapes():
    # the batched solver's launch variants (no GPU needed to list them):
    # threads per workgroup, rows of X per wave step, depth of the row ring
    from oap_amd.api import CovarianceBatch
    shapes = CovarianceBatch.variants()
    assert len(shapes) >= 7 and len(set(shapes)) == len(shapes)
    for threads, rows, depth in shapes:
        assert threads in (128, 256, 512, 1024) and rows in (1, 2) and depth in (2, 3, 4)
    import ctypes as C
    from oap_amd._lib import load
    lib = load()
    t, r = C.c_int(), C.c_int()
    assert lib.oaph_batch_variant_info(0, C.byref(t), C.byref(r)) == 0
    assert (t.value, r.value) == shapes[0][:2]
    assert lib.oaph_batch_variant_shape(len(shapes), None, None, None) != 0
