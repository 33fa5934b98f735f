"""Matrix text and binary formats of oap (host code in liboaph.so).

Text: ``oap::host::ToString`` / ``NewMatrix(text)``
(oapMatrixCpu/oapHostMatrixUtils.cpp:223-274,576-588; oapMatrix/MatrixPrinter.h,
oapMatrix/MatrixParser.cpp).  Binary: ``SaveMatrix`` / ``LoadMatrix`` records
and the ``utils::ByteBuffer`` file framing (oapMatrixCpu/oapHostMatrixUtils.cpp
:1043-1104, oapUtils/Buffer.h).  The device forms copy the matrix once.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DP, MP, check, load, oaph_matrix_info
from .api import DeviceMatrix

_SZ = C.c_size_t


def _d(a):
    return None if a is None else a.ctypes.data_as(DP)


def _plane(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    if a.size != n:
        raise ValueError("plane size does not match columns x rows")
    return a


def to_string(columns: int, rows: int, re=None, im=None) -> str:
    """ToString of a host matrix given as row-major planes."""
    lib = load()
    n = columns * rows
    r, i = _plane(re, n), _plane(im, n)
    ln = _SZ()
    check(lib.oaph_host_matrix_to_string(columns, rows, _d(r), _d(i), None, 0, C.byref(ln)),
          "to_string")
    buf = C.create_string_buffer(ln.value + 1)
    check(lib.oaph_host_matrix_to_string(columns, rows, _d(r), _d(i), buf, ln.value + 1,
                                         C.byref(ln)), "to_string")
    return buf.value.decode()


def matrix_to_string(m: DeviceMatrix | None) -> str:
    lib = load()
    p = None if m is None else m.ptr
    ln = _SZ()
    check(lib.oaph_matrix_to_string(p, None, 0, C.byref(ln)), "matrix_to_string")
    buf = C.create_string_buffer(ln.value + 1)
    check(lib.oaph_matrix_to_string(p, buf, ln.value + 1, C.byref(ln)), "matrix_to_string")
    return buf.value.decode()


def parse_array(text: str, which: int = 1) -> np.ndarray:
    """CreateArray(text, which)."""
    lib = load()
    t = text.encode()
    ln = _SZ()
    check(lib.oaph_host_parse_array(t, which, None, 0, C.byref(ln)), "parse_array")
    out = np.zeros(ln.value)
    check(lib.oaph_host_parse_array(t, which, _d(out), out.size, C.byref(ln)), "parse_array")
    return out


def parse_host(text: str):
    """NewMatrix(text) into host planes: (columns, rows, re | None, im | None)."""
    lib = load()
    t = text.encode()
    info = oaph_matrix_info()
    check(lib.oaph_host_parse_matrix(t, C.byref(info), None, None, 0), "parse")
    n = info.columns * info.rows
    re = np.zeros(n) if info.isRe else None
    im = np.zeros(n) if info.isIm else None
    check(lib.oaph_host_parse_matrix(t, C.byref(info), _d(re), _d(im), n), "parse")
    shape = (info.rows, info.columns)
    return (info.columns, info.rows, None if re is None else re.reshape(shape),
            None if im is None else im.reshape(shape))


def parse(text: str) -> DeviceMatrix:
    """NewMatrix(text) as a device matrix."""
    p = MP()
    check(load().oaph_matrix_parse(text.encode(), C.byref(p)), "matrix_parse")
    return DeviceMatrix._adopt(p)


def save_host(columns: int, rows: int, re=None, im=None, present: bool = True) -> bytes:
    lib = load()
    n = columns * rows
    r, i = _plane(re, n), _plane(im, n)
    ln = _SZ()
    check(lib.oaph_host_matrix_save(int(present), columns, rows, _d(r), _d(i), None, 0,
                                    C.byref(ln)), "save")
    buf = (C.c_uint8 * ln.value)()
    check(lib.oaph_host_matrix_save(int(present), columns, rows, _d(r), _d(i), buf, ln.value,
                                    C.byref(ln)), "save")
    return bytes(buf)


def save(m: DeviceMatrix | None) -> bytes:
    """SaveMatrix of a device matrix (None: the null record)."""
    lib = load()
    p = None if m is None else m.ptr
    ln = _SZ()
    check(lib.oaph_matrix_save(p, None, 0, C.byref(ln)), "matrix_save")
    buf = (C.c_uint8 * ln.value)()
    check(lib.oaph_matrix_save(p, buf, ln.value, C.byref(ln)), "matrix_save")
    return bytes(buf)


def load_host(data: bytes, offset: int = 0):
    """One record at `offset`: ((columns, rows, re, im) or None, bytes consumed)."""
    lib = load()
    view = (C.c_uint8 * (len(data) - offset)).from_buffer_copy(data, offset)
    info, present, used = oaph_matrix_info(), C.c_int(), _SZ()
    check(lib.oaph_host_matrix_load(view, len(view), C.byref(used), C.byref(present),
                                    C.byref(info), None, None, 0), "load")
    if not present.value:
        return None, used.value
    n = info.columns * info.rows
    re = np.zeros(n) if info.isRe else None
    im = np.zeros(n) if info.isIm else None
    check(lib.oaph_host_matrix_load(view, len(view), C.byref(used), C.byref(present),
                                    C.byref(info), _d(re), _d(im), n), "load")
    shape = (info.rows, info.columns)
    return ((info.columns, info.rows, None if re is None else re.reshape(shape),
             None if im is None else im.reshape(shape)), used.value)


def load_device(data: bytes, offset: int = 0):
    """LoadMatrix: (DeviceMatrix or None, bytes consumed)."""
    view = (C.c_uint8 * (len(data) - offset)).from_buffer_copy(data, offset)
    p, used = MP(), _SZ()
    check(load().oaph_matrix_load(C.byref(p), view, len(view), C.byref(used)), "matrix_load")
    return (DeviceMatrix._adopt(p) if p else None), used.value


def write_file(path: str, data: bytes) -> None:
    """ByteBuffer::fwrite framing."""
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data) if data else None
    check(load().oaph_buffer_write_file(str(path).encode(), buf, len(data)), "write_file")


def read_file(path: str) -> bytes:
    """ByteBuffer::fread framing."""
    lib = load()
    ln = _SZ()
    check(lib.oaph_buffer_read_file(str(path).encode(), None, 0, C.byref(ln)), "read_file")
    buf = (C.c_uint8 * ln.value)()
    check(lib.oaph_buffer_read_file(str(path).encode(), buf, ln.value, C.byref(ln)), "read_file")
    return bytes(buf)
