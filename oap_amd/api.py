"""Python mirror of oap's procedures API over the C ABI.

`CuProceduresApi` keeps the reference's method names and argument order
(oapMatrixCuda/CuProceduresApi.h:45-309); matrices are `DeviceMatrix`
handles (math::Matrix with planes in HBM).  Scalars come back as return
values instead of `floatt&` out-parameters.  Every call goes through
liboaph.so; a non-zero status raises OapError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DIMS, DP, MP, U32, OapError, check, load, oaph_matrix, oaph_matrix_info


def _dptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(DP)


class DeviceMatrix:
    """A device matrix: math::Matrix with its re/im planes in HBM.

    Mirrors oap::cuda::NewDeviceMatrix / DeleteDeviceMatrix
    (oapMatrixCuda/oapCudaMatrixUtils.h:38-74).
    """

    def __init__(self, columns: int, rows: int, is_re: bool = True, is_im: bool = True):
        lib = load()
        self._p = MP()
        check(lib.oaph_matrix_new(C.byref(self._p), int(is_re), int(is_im), columns, rows),
              "oaph_matrix_new")

    # ---- construction helpers
    @classmethod
    def _adopt(cls, p) -> "DeviceMatrix":
        """Takes ownership of a library-allocated matrix (MP)."""
        m = cls.__new__(cls)
        m._p = p
        return m

    @classmethod
    def from_numpy(cls, re: np.ndarray | None, im: np.ndarray | None = None) -> "DeviceMatrix":
        """re/im are (rows, columns) arrays; either may be None (absent plane)."""
        ref = re if re is not None else im
        if ref is None:
            raise ValueError("need at least one plane")
        ref = np.asarray(ref)
        if ref.ndim == 1:
            ref = ref.reshape(-1, 1)
        rows, cols = ref.shape
        m = cls(cols, rows, re is not None, im is not None)
        m.upload(re, im)
        return m

    @classmethod
    def zeros_like_info(cls, info) -> "DeviceMatrix":
        return cls(info.columns, info.rows, bool(info.isRe), bool(info.isIm))

    def upload(self, re: np.ndarray | None, im: np.ndarray | None = None) -> None:
        r = None if re is None else np.ascontiguousarray(np.asarray(re, dtype=np.float64).reshape(-1))
        i = None if im is None else np.ascontiguousarray(np.asarray(im, dtype=np.float64).reshape(-1))
        for a in (r, i):
            if a is not None and a.size != self.columns * self.rows:
                raise ValueError("plane size does not match the matrix")
        check(load().oaph_matrix_copy_h2d(self._p, _dptr(r), _dptr(i)), "oaph_matrix_copy_h2d")

    def download(self) -> tuple[np.ndarray | None, np.ndarray | None]:
        n = self.columns * self.rows
        re = np.empty(n) if self.is_re else None
        im = np.empty(n) if self.is_im else None
        check(load().oaph_matrix_copy_d2h(_dptr(re), _dptr(im), self._p), "oaph_matrix_copy_d2h")
        shape = (self.rows, self.columns)
        return (None if re is None else re.reshape(shape), None if im is None else im.reshape(shape))

    def to_complex(self) -> np.ndarray:
        re, im = self.download()
        z = np.zeros((self.rows, self.columns), dtype=np.complex128)
        if re is not None:
            z.real = re
        if im is not None:
            z.imag = im
        return z

    # ---- info
    @property
    def ptr(self):
        return self._p

    @property
    def columns(self) -> int:
        return self._p.contents.columns

    @property
    def rows(self) -> int:
        return self._p.contents.rows

    @property
    def is_re(self) -> bool:
        return bool(self._p.contents.reValues)

    @property
    def is_im(self) -> bool:
        return bool(self._p.contents.imValues)

    def info(self) -> oaph_matrix_info:
        inf = oaph_matrix_info()
        check(load().oaph_matrix_info_get(self._p, C.byref(inf)), "oaph_matrix_info_get")
        return inf

    def get_value(self, index: int) -> complex:
        re, im = C.c_double(), C.c_double()
        check(load().oaph_matrix_get_value(self._p, index, C.byref(re), C.byref(im)),
              "oaph_matrix_get_value")
        return complex(re.value, im.value)

    def set_value(self, index: int, re: float, im: float = 0.0) -> None:
        check(load().oaph_matrix_set_value(self._p, index, re, im), "oaph_matrix_set_value")

    def set_zero_row(self, row: int, re: bool = True, im: bool = True) -> None:
        """SetZeroRow (oapCudaMatrixUtils.h:186)"""
        check(load().oaph_matrix_set_zero_row(self._p, row, int(re), int(im)),
              "oaph_matrix_set_zero_row")

    def copy_from(self, src: "DeviceMatrix") -> None:
        """CopyDeviceMatrixToDeviceMatrix (oapCudaMatrixUtils.h:95)"""
        check(load().oaph_matrix_copy_d2d(self._p, src.ptr), "oaph_matrix_copy_d2d")

    def copy_from_dims(self, src: "DeviceMatrix", dims) -> None:
        """CopyDeviceMatrixToDeviceMatrixDims: dims[0] = [[column, columns],
        [row, rows]] of this matrix, dims[1] of src"""
        check(load().oaph_matrix_copy_d2d_dims(self._p, src.ptr, _cdims(dims)),
              "oaph_matrix_copy_d2d_dims")

    def upload_dims(self, re, im, dims) -> None:
        """CopyHostMatrixToDeviceMatrixDims from host planes (rows x columns)"""
        hm, keep = _host_matrix(re, im)
        check(load().oaph_matrix_copy_h2d_dims(self._p, C.byref(hm), _cdims(dims)),
              "oaph_matrix_copy_h2d_dims")
        del keep

    def download_dims(self, re, im, dims) -> None:
        """CopyDeviceMatrixToHostMatrixDims into host planes (rows x columns,
        C-contiguous float64 arrays, written in place)"""
        hm, keep = _host_matrix(re, im)
        check(load().oaph_matrix_copy_d2h_dims(C.byref(hm), self._p, _cdims(dims)),
              "oaph_matrix_copy_d2h_dims")
        del keep

    def set_matrix(self, src: "DeviceMatrix", column: int, row: int, planes: int = 3) -> None:
        """SetMatrix (planes 3) / SetReMatrix (1) / SetImMatrix (2)
        (oapCudaMatrixUtils.h:142-146): src written at (column, row)"""
        check(load().oaph_matrix_set_submatrix(self._p, src.ptr, column, row, planes),
              "oaph_matrix_set_submatrix")

    def free(self) -> None:
        if getattr(self, "_p", None):
            load().oaph_matrix_free(self._p)
            self._p = MP()

    def __del__(self):  # pragma: no cover - interpreter teardown order varies
        try:
            self.free()
        except Exception:
            pass


def _dims(d) -> DIMS:
    arr = DIMS()
    for i in range(3):
        for j in range(2):
            arr[i][j] = int(d[i][j])
    return arr


def _m(x):
    return None if x is None else x.ptr


def _cdims(d):
    from ._lib import CDIMS
    arr = CDIMS()
    for i in range(2):
        for j in range(2):
            for k in range(2):
                arr[i][j][k] = int(d[i][j][k])
    return arr


def _host_matrix(re, im):
    """An oaph_matrix over host planes (rows x columns float64, C-contiguous;
    either may be None); returns it with the arrays it points to."""
    ref = re if re is not None else im
    for a in (re, im):
        if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous or a.ndim != 2):
            raise ValueError("host planes must be C-contiguous 2-D float64 arrays")
    hm = oaph_matrix()
    hm.rows, hm.columns = ref.shape
    hm.realRows, hm.realColumns = ref.shape
    hm.reValues = None if re is None else re.ctypes.data
    hm.imValues = None if im is None else im.ctypes.data
    return hm, (re, im)


class CuProceduresApi:
    """oap::CuProceduresApi (oapMatrixCuda/CuProceduresApi.h:45-309) over liboaph.so."""

    def __init__(self, handle: int | None = None):
        self._lib = load()
        self._owned = handle is None
        if handle is None:
            h = C.c_void_p()
            check(self._lib.oaph_api_create(C.byref(h)), "oaph_api_create")
            self._h = h
        else:
            self._h = C.c_void_p(handle)

    @property
    def handle(self):
        return self._h

    def synchronize(self) -> None:
        check(self._lib.oaph_api_synchronize(self._h), "oaph_api_synchronize")

    def destroy(self) -> None:
        if self._owned and self._h:
            self._lib.oaph_api_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass

    # ---- products
    def dotProduct(self, out, a, b, dims=None):
        if dims is None:
            check(self._lib.oaph_dot_product(self._h, out.ptr, a.ptr, b.ptr), "dotProduct")
        else:
            check(self._lib.oaph_dot_product_dims(self._h, out.ptr, a.ptr, b.ptr, _dims(dims)),
                  "dotProduct(dims)")

    def dotProductShared(self, out, a, b):
        check(self._lib.oaph_dot_product_shared(self._h, out.ptr, a.ptr, b.ptr), "dotProductShared")

    def tensorProduct(self, out, a, b, dims=None):
        if dims is None:
            dims = [[out.columns, out.rows], [a.columns, a.rows], [b.columns, b.rows]]
        check(self._lib.oaph_tensor_product(self._h, out.ptr, a.ptr, b.ptr, _dims(dims)),
              "tensorProduct")

    def hadamardProduct(self, out, a, b):
        check(self._lib.oaph_hadamard_product(self._h, out.ptr, a.ptr, b.ptr), "hadamardProduct")

    def hadamardProductVec(self, out, a, b):
        check(self._lib.oaph_hadamard_product_vec(self._h, out.ptr, a.ptr, b.ptr),
              "hadamardProductVec")

    def tensorProductCR(self, out, a, b, columns: int, rows: int):
        """tensorProduct(output, a, b, columns, rows) (CuProceduresApi.cpp:136-151)"""
        check(self._lib.oaph_tensor_product_cr(self._h, out.ptr, a.ptr, b.ptr, columns, rows),
              "tensorProduct(columns, rows)")

    def hadamardProductCR(self, out, a, b, columns: int, rows: int):
        check(self._lib.oaph_hadamard_product_cr(self._h, out.ptr, a.ptr, b.ptr, columns, rows),
              "hadamardProduct(columns, rows)")

    def hadamardProductVecCR(self, out, a, b, columns: int, rows: int):
        check(self._lib.oaph_hadamard_product_vec_cr(self._h, out.ptr, a.ptr, b.ptr, columns,
                                                     rows), "hadamardProductVec(columns, rows)")

    def addDotProduct(self, out, a, b):
        """the reference kernel writes output = a b (not +=)"""
        check(self._lib.oaph_add_dot_product(self._h, out.ptr, a.ptr, b.ptr), "addDotProduct")

    def dotProductPeriodic(self, out, a, b):
        check(self._lib.oaph_dot_product_periodic(self._h, out.ptr, a.ptr, b.ptr),
              "dotProductPeriodic")

    def dotProductDimPeriodic(self, out, a, b, dims, periodic_rows=None):
        if periodic_rows is None:
            periodic_rows = a.rows  # CuProceduresApi.h:149-153
        check(self._lib.oaph_dot_product_dim_periodic(self._h, out.ptr, a.ptr, b.ptr,
                                                      _dims(dims), periodic_rows),
              "dotProductDimPeriodic")

    def dotProductEx(self, out, a, b, ex):
        check(self._lib.oaph_dot_product_ex(self._h, out.ptr, a.ptr, b.ptr, C.byref(_mex(ex))),
              "dotProductEx")

    dotProductExOpt = dotProductEx

    def transposeEx(self, out, a, ex):
        check(self._lib.oaph_transpose_ex(self._h, out.ptr, a.ptr, C.byref(_mex(ex))),
              "transposeEx")

    # ---- element-wise
    def transpose(self, out, a):
        check(self._lib.oaph_transpose(self._h, out.ptr, a.ptr), "transpose")

    def conjugateTranspose(self, out, a):
        check(self._lib.oaph_conjugate_transpose(self._h, out.ptr, a.ptr), "conjugateTranspose")

    def add(self, out, a, b):
        check(self._lib.oaph_add(self._h, out.ptr, a.ptr, b.ptr), "add")

    def substract(self, out, a, b):
        check(self._lib.oaph_substract(self._h, out.ptr, a.ptr, b.ptr), "substract")

    def addCR(self, out, a, b, columns: int, rows: int):
        """add(output, a, b, columns, rows): the columns x rows region"""
        check(self._lib.oaph_add_cr(self._h, out.ptr, a.ptr, b.ptr, columns, rows), "add(cr)")

    def substractCR(self, out, a, b, columns: int, rows: int):
        check(self._lib.oaph_substract_cr(self._h, out.ptr, a.ptr, b.ptr, columns, rows),
              "substract(cr)")

    def addSubstract(self, out, a, b, columns=None, rows=None):
        """output += a - b over the columns x rows region (default: all of out)"""
        columns = out.columns if columns is None else columns
        rows = out.rows if rows is None else rows
        check(self._lib.oaph_add_substract(self._h, out.ptr, a.ptr, b.ptr, columns, rows),
              "addSubstract")

    def multiplyReConstant(self, out, a, re: float):
        check(self._lib.oaph_multiply_re_constant(self._h, out.ptr, a.ptr, re),
              "multiplyReConstant")

    def multiplyConstant(self, out, a, re: float, im: float):
        check(self._lib.oaph_multiply_constant(self._h, out.ptr, a.ptr, re, im), "multiplyConstant")

    def setDiagonal(self, m, re: float, im: float):
        check(self._lib.oaph_set_diagonal(self._h, m.ptr, re, im), "setDiagonal")

    def setIdentity(self, m):
        check(self._lib.oaph_set_identity(self._h, m.ptr), "setIdentity")

    def setZeroMatrix(self, m):
        check(self._lib.oaph_set_zero_matrix(self._h, m.ptr), "setZeroMatrix")

    def setVector(self, V, column: int, v, length: int):
        check(self._lib.oaph_set_vector(self._h, V.ptr, column, v.ptr, length), "setVector")

    def getVector(self, v, length: int, V, column: int):
        check(self._lib.oaph_get_vector(self._h, v.ptr, length, V.ptr, column), "getVector")

    # ---- reductions
    def magnitude(self, m) -> float:
        out = C.c_double()
        check(self._lib.oaph_magnitude(self._h, C.byref(out), m.ptr), "magnitude")
        return out.value

    def magnitude2(self, m) -> float:
        out = C.c_double()
        check(self._lib.oaph_magnitude2(self._h, C.byref(out), m.ptr), "magnitude2")
        return out.value

    magnitude2Opt = magnitude2
    magnitude2OptVer2 = magnitude2
    magnitudeOpt = magnitude
    magnitudeOptVer2 = magnitude

    def sum(self, m) -> tuple[float, float]:
        re, im = C.c_double(), C.c_double()
        check(self._lib.oaph_sum(self._h, C.byref(re), C.byref(im), m.ptr), "sum")
        return re.value, im.value

    # ---- small dense
    def QRGR(self, Q, R, H, aux0=None, aux1=None, aux2=None, aux3=None):
        check(self._lib.oaph_qrgr(self._h, Q.ptr, R.ptr, H.ptr, _m(aux0), _m(aux1), _m(aux2),
                                  _m(aux3)), "QRGR")

    def isUpperTriangular(self, m) -> bool:
        r = C.c_int()
        check(self._lib.oaph_is_upper_triangular(self._h, m.ptr, C.byref(r)), "isUpperTriangular")
        return r.value == 1

    def calcTriangularH(self, H, Q, R=None, *aux):
        a = list(aux) + [None] * (6 - len(aux))
        check(self._lib.oaph_calc_triangular_h(self._h, H.ptr, Q.ptr, _m(R), *[_m(x) for x in a]),
              "calcTriangularH")


    def calcTriangularHStep(self, H, Q, R=None, *aux):
        check(self._lib.oaph_calc_triangular_h_step(self._h, H.ptr, Q.ptr), "calcTriangularHStep")


def _mex(ex):
    """MatrixEx from a (column, columns, row, rows) sequence or struct"""
    from ._lib import oaph_matrix_ex
    if isinstance(ex, oaph_matrix_ex):
        return ex
    return oaph_matrix_ex(*[int(v) for v in ex])


class Activation:
    """OAPH_FN_* codes (CuProceduresApi.h:239-290); the layer activations of
    oap's Network are SIGMOID, LINEAR, TANH and SIN."""
    SIGMOID, DSIGMOID, MULTIPLY_DSIGMOID = 0, 1, 2
    LINEAR, DLINEAR = 3, 4
    TANH, DTANH, MULTIPLY_DTANH = 5, 6, 7
    SIN, DSIN, MULTIPLY_DSIN = 8, 9, 10


def _activation(api, fn, out, m, dims):
    lib = api._lib
    if dims is None:
        check(lib.oaph_activation(api._h, fn, _m(out), _m(m)), "activation")
        return
    d = np.asarray(dims, dtype=np.uint32)
    if d.shape == (2,):
        arr = (C.c_uint32 * 2)(*d)
        check(lib.oaph_activation_dims(api._h, fn, _m(out), _m(m), arr), "activation_dims")
    elif d.shape == (2, 2):
        arr = (C.c_uint32 * 4)(*d.ravel())
        check(lib.oaph_activation_dims_periodic(api._h, fn, _m(out), _m(m), arr),
              "activation_dims_periodic")
    else:
        raise ValueError("dims must be [2] or [2][2]")


def _add_activation_methods():
    names = {"sigmoid": Activation.SIGMOID, "dsigmoid": Activation.DSIGMOID,
             "multiplyDSigmoid": Activation.MULTIPLY_DSIGMOID, "linear": Activation.LINEAR,
             "dlinear": Activation.DLINEAR, "tanh": Activation.TANH, "dtanh": Activation.DTANH,
             "multiplyDTanh": Activation.MULTIPLY_DTANH, "sin": Activation.SIN,
             "dsin": Activation.DSIN, "cos": Activation.DSIN,
             "multiplyDSin": Activation.MULTIPLY_DSIN}
    for name, fn in names.items():
        def method(self, out, m, dims=None, _fn=fn):
            _activation(self, _fn, out, m, dims)
        method.__name__ = name
        setattr(CuProceduresApi, name, method)


_add_activation_methods()


def set_threads_blocks(w: int, h: int, limit: int = 1024):
    """utils::mapper::SetThreadsBlocks (oapUtils/ThreadsMapper.cpp:36-74) -> (blocks, threads)."""
    lib = load()
    b = (U32 * 2)()
    t = (U32 * 2)()
    check(lib.oaph_set_threads_blocks(b, t, w, h, limit), "SetThreadsBlocks")
    return (b[0], b[1]), (t[0], t[1])


__all__ = ["DeviceMatrix", "CuProceduresApi", "OapError", "set_threads_blocks", "oaph_matrix"]


# ---------------------------------------------------------------------------
# Arnoldi: CuHArnoldi / CuHArnoldiCallback (ArnoldiPackage/ArnoldiProcedures.h:33-241,
# ArnoldiProceduresImpl.h:45-74)
# ---------------------------------------------------------------------------
class QRType:
    """oap::QRType (ArnoldiPackage/oapGenericArnoldiApi.h:31-36)"""
    NONE, QRGR, QRHT = -1, 0, 1


class VecMultiplicationType:
    """oap::VecMultiplicationType (oapGenericArnoldiApi.h:38-42)"""
    TYPE_EIGENVECTOR, TYPE_WV = 0, 1


class InitVVectorType:
    """oap::InitVVectorType (oapGenericArnoldiApi.h:44-48)"""
    RANDOM, FIRST_VALUE_IS_ONE = 0, 1


class ArnUtils:
    """ArnUtils sort functions, CheckType, TriangularHProcedureType, Type (ArnoldiUtils.h:58-87)"""
    SortLargestValues, SortLargestReValues, SortLargestImValues = 0, 1, 2
    SortSmallestValues, SortSmallestReValues, SortSmallestImValues = 3, 4, 5
    CHECK_INTERNAL, CHECK_FIRST_STOP, CHECK_COUNTER = 0, 1, 2
    CALC_IN_DEVICE, CALC_IN_HOST, CALC_IN_DEVICE_STEP = 0, 1, 2
    UNDEFINED, DEVICE, HOST = 0, 1, 2


class Precision:
    FP64, MIXED_FP32 = 0, 1


class Projection:
    """first Gram-Schmidt pass of each Arnoldi step (OAPH_PROJ_*)"""
    FP64, BF16X3 = 0, 1


class MatrixInfo:
    """math::MatrixInfo (oapMatrix/MatrixInfo.h:16-67)."""

    def __init__(self, is_re: bool, is_im: bool, columns: int, rows: int):
        self.c = oaph_matrix_info(int(is_re), int(is_im), columns, rows)

    @property
    def rows(self):
        return self.c.rows


class CuHArnoldiCallback:
    """CuHArnoldiCallback over liboaph.so.

    multiply(w, v, api, user_data, mt): w = A v, where w and v are
    single-column DeviceMatrixView objects owned by the solver and api is the
    solver's CuProceduresApi (same stream).  Set either a callback or a dense
    device operator (setMatrix, the CuHArnoldiDefault behaviour).
    """

    def __init__(self):
        from ._lib import CHECK_FN, MULTIPLY_FN
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.oaph_arnoldi_create(C.byref(h)), "oaph_arnoldi_create")
        self._h = h
        self.api = CuProceduresApi(self._lib.oaph_arnoldi_api(h))
        self._MULTIPLY_FN = MULTIPLY_FN
        self._CHECK_FN = CHECK_FN
        self._cb = None
        self._ccb = None
        self._outs = None
        self._error = None

    def __del__(self):  # pragma: no cover
        try:
            if self._h:
                self._lib.oaph_arnoldi_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    # ---- setters (ArnoldiProcedures.h:41-62)
    def setRho(self, rho=1.0 / 3.14):
        check(self._lib.oaph_arnoldi_set_rho(self._h, rho), "setRho")

    def setBLimit(self, blimit):
        check(self._lib.oaph_arnoldi_set_blimit(self._h, blimit), "setBLimit")

    def setQRType(self, qrtype=QRType.QRGR):
        check(self._lib.oaph_arnoldi_set_qr_type(self._h, qrtype), "setQRType")

    def setVecInitType(self, t):
        check(self._lib.oaph_arnoldi_set_vec_init_type(self._h, t), "setVecInitType")

    def setSortType(self, t):
        """a built-in order (ArnUtils.Sort*: applied on the device) or a
        comparator before(i, j) -> bool over complex Ritz values (the
        reference's ArnUtils::SortType, ArnoldiUtils.h:63-72: ordered on the
        host at each restart)"""
        if callable(t):
            from ._lib import SORT_FN

            def before(ri, ii, rj, ij, _user):
                return 1 if t(complex(ri, ii), complex(rj, ij)) else 0

            self._sort_cb = SORT_FN(before)  # kept alive with the solver
            check(self._lib.oaph_arnoldi_set_sort_fn(self._h, self._sort_cb, None), "setSortType")
        else:
            check(self._lib.oaph_arnoldi_set_sort_type(self._h, t), "setSortType")
            self._sort_cb = None

    def setCheckType(self, t):
        check(self._lib.oaph_arnoldi_set_check_type(self._h, t), "setCheckType")

    def setCheckCounts(self, n):
        check(self._lib.oaph_arnoldi_set_check_counts(self._h, n), "setCheckCounts")

    def setCheckTolerance(self, tol):
        check(self._lib.oaph_arnoldi_set_check_tolerance(self._h, tol), "setCheckTolerance")

    def setOutputType(self, t):
        check(self._lib.oaph_arnoldi_set_output_type(self._h, t), "setOutputType")

    def setCalcTraingularHType(self, t):
        check(self._lib.oaph_arnoldi_set_calc_triangular_h_type(self._h, t),
              "setCalcTraingularHType")

    # extensions
    def setSeed(self, seed):
        check(self._lib.oaph_arnoldi_set_seed(self._h, seed), "setSeed")

    def setNormA(self, norm):
        check(self._lib.oaph_arnoldi_set_norm_a(self._h, norm), "setNormA")

    def setMaxRestarts(self, n):
        check(self._lib.oaph_arnoldi_set_max_restarts(self._h, n), "setMaxRestarts")

    def setPrecision(self, p):
        check(self._lib.oaph_arnoldi_set_precision(self._h, p), "setPrecision")

    def setProjection(self, mode):
        """Projection.FP64 (default) or Projection.BF16X3: the first
        Gram-Schmidt pass on the matrix cores from a BF16-split basis"""
        check(self._lib.oaph_arnoldi_set_projection(self._h, mode), "setProjection")

    def setRestartStages(self, tols):
        """restart from the sum of the wanted (and guard) Ritz vectors at each
        tolerance (opt-in: such restarts can settle on a wrong wanted set when
        the k-th eigenvalue is close to the next)"""
        t = np.ascontiguousarray(tols, dtype=np.float64)
        check(self._lib.oaph_arnoldi_set_restart_stages(self._h, _dptr(t) if t.size else None,
                                                         t.size), "setRestartStages")

    def setRestartGuard(self, guard):
        """Ritz vectors after the wanted ones kept in each Ritz-sum restart's
        sum (the mixed flow's FP64 start and the stages; default 4)"""
        check(self._lib.oaph_arnoldi_set_restart_guard(self._h, guard), "setRestartGuard")

    def setRestartGuardWeight(self, weight):
        """weight of the guard Ritz vectors in a Ritz-sum restart (wanted: 1)"""
        check(self._lib.oaph_arnoldi_set_restart_guard_weight(self._h, weight),
              "setRestartGuardWeight")

    KEEP_AUTO = 0xFFFFFFFF

    def setRestartKeep(self, keep):
        """smallest number of Ritz vectors an exact-shift restart keeps: 0 is
        the reference's k plus min(nconv, (m - k) / 2) (ARPACK's adjustment);
        KEEP_AUTO (the default) is m / 2 in the mixed flow and 0 in FP64 and
        for an operator whose FP32 copy is exact"""
        check(self._lib.oaph_arnoldi_set_restart_keep(self._h, keep), "setRestartKeep")

    def setSchurMethod(self, method):
        """How the restart finds the eigenpairs of H.  2 (default): the
        Aberth eigenvalues and the wanted eigenvectors by inverse iteration,
        no Schur form where every vector is accepted; 1: the Schur form from
        perfect shifts (the Aberth eigenvalues); 0: Wilkinson shifts only.
        Each falls back to the next lower method where it does not succeed."""
        check(self._lib.oaph_arnoldi_set_schur_method(self._h, method), "setSchurMethod")

    def setMixedTolerance(self, tol_low):
        check(self._lib.oaph_arnoldi_set_mixed_tolerance(self._h, tol_low), "set_mixed_tolerance")

    def setF16Phase(self, on=True, tol=2e-3):
        """mixed flow: cycles on an FP16 copy of the operator (scaled by a
        power of two) until the Ritz estimates reach tol (relative to the
        largest Ritz value), then the FP32 phase from their Ritz-vector sum"""
        check(self._lib.oaph_arnoldi_set_f16_phase(self._h, int(bool(on)), tol), "set_f16_phase")

    def setMixedCheck(self, on=True, fraction=0.5):
        """mixed flow after the FP32 phase: True (default) runs FP32 cycles
        down to fraction x the tolerance and accepts the pairs on an explicit
        FP64 residual check, with FP64 cycles only where it fails; False
        always polishes with FP64 cycles"""
        check(self._lib.oaph_arnoldi_set_mixed_check(self._h, int(bool(on)), fraction),
              "set_mixed_check")

    def setStartVector(self, v):
        if v is None:
            check(self._lib.oaph_arnoldi_set_start_vector(self._h, None, None, 0), "setStartVector")
            return
        v = np.asarray(v)
        re = np.ascontiguousarray(v.real, dtype=np.float64)
        im = np.ascontiguousarray(v.imag, dtype=np.float64)
        check(self._lib.oaph_arnoldi_set_start_vector(self._h, _dptr(re), _dptr(im), v.size),
              "setStartVector")

    def setMatrix(self, A: DeviceMatrix):
        self._A = A
        check(self._lib.oaph_arnoldi_set_matrix(self._h, A.ptr), "setMatrix")

    def setCallback(self, fn, user_data=None):
        """fn(w, v, api, user_data, mt) with w, v DeviceMatrixView."""
        if fn is None:
            self._cb = None
            check(self._lib.oaph_arnoldi_set_callback(self._h, self._MULTIPLY_FN(), None),
                  "setCallback")
            return

        def tramp(wp, vp, _api, _u, mt):
            try:
                fn(DeviceMatrixView(wp), DeviceMatrixView(vp), self.api, user_data, mt)
            except Exception as e:  # surface after execute returns
                self._error = e

        self._cb = self._MULTIPLY_FN(tramp)
        check(self._lib.oaph_arnoldi_set_callback(self._h, self._cb, None), "setCallback")

    def setCheckCallback(self, fn, user_data=None):
        def tramp(re, im, vp, index, mx, _u):
            return int(bool(fn(re, im, DeviceMatrixView(vp), index, mx, user_data)))

        self._ccb = self._CHECK_FN(tramp)
        check(self._lib.oaph_arnoldi_set_check_callback(self._h, self._ccb, None),
              "setCheckCallback")

    def setOutputsEigenvalues(self, re: np.ndarray, im: np.ndarray):
        self._ev = (re, im)
        check(self._lib.oaph_arnoldi_set_outputs_eigenvalues(self._h, _dptr(re), _dptr(im)),
              "setOutputsEigenvalues")

    def setOutputsEigenvectors(self, vectors):
        """vectors: list of DeviceMatrix (n x 1) or None."""
        arr = (MP * len(vectors))(*[v.ptr if v is not None else MP() for v in vectors])
        self._outs = (arr, vectors)
        check(self._lib.oaph_arnoldi_set_outputs_eigenvectors(self._h, arr),
              "setOutputsEigenvectors")

    # ---- run
    def execute(self, hdim: int, wanted: int, info: MatrixInfo, mtype=ArnUtils.DEVICE,
                raise_on_not_converged=True) -> int:
        self._error = None
        st = self._lib.oaph_arnoldi_execute(self._h, hdim, wanted, C.byref(info.c), mtype)
        if self._error is not None:
            raise self._error
        if st == 7 and not raise_on_not_converged:
            return st
        check(st, "execute")
        return st

    def applyOperator(self, precision: int, x: "DeviceMatrix", y: "DeviceMatrix"):
        """y = A x with the dense operator at a precision: 0 FP64, 1 the FP32
        copy (FP64 products and sums), 2 the FP16 copy (FP32 products and
        per-lane sums)"""
        check(self._lib.oaph_arnoldi_apply_operator(self._h, precision, x.ptr, y.ptr),
              "apply_operator")

    def testOutcome(self, index: int) -> float:
        out = C.c_double()
        check(self._lib.oaph_arnoldi_test_outcome(self._h, index, C.byref(out)), "testOutcome")
        return out.value

    def stats(self) -> dict:
        from ._lib import oaph_arnoldi_stats
        s = oaph_arnoldi_stats()
        check(self._lib.oaph_arnoldi_get_stats(self._h, C.byref(s)), "get_stats")
        return {f: getattr(s, f) for f, _ in s._fields_}


class DeviceMatrixView:
    """A non-owning view of a solver-owned device matrix (the w/v of the callback)."""

    def __init__(self, p):
        self._p = p

    @property
    def ptr(self):
        return self._p

    @property
    def columns(self):
        return self._p.contents.columns

    @property
    def rows(self):
        return self._p.contents.rows

    @property
    def is_re(self):
        return bool(self._p.contents.reValues)

    @property
    def is_im(self):
        return bool(self._p.contents.imValues)

    download = DeviceMatrix.download
    upload = DeviceMatrix.upload
    to_complex = DeviceMatrix.to_complex


__all__ += ["CuHArnoldiCallback", "MatrixInfo", "QRType", "VecMultiplicationType",
            "InitVVectorType", "ArnUtils", "Precision", "DeviceMatrixView"]


class RecToSquareApi:
    """The square operator S = M M^T of a rectangular device matrix M.

    Mirrors oap::RecToSquareApi (oapMatrixCuda/RecToSquareApi.h:33-58): S is
    R x R for M with R rows; a matrix that is already square is used as it
    is (RecToSquareApi.cpp:57-63).  Blocks of S come from a tiled product
    that reads M transposed in place (SquareMatrixApi.cpp:143-181 forms a
    transposed copy instead); multiply(w, v) applies S without forming it.
    M must outlive this object.
    """

    def __init__(self, api: "CuProceduresApi", rec: DeviceMatrix):
        self._lib = load()
        self._rec = rec
        h = C.c_void_p()
        check(self._lib.oaph_rts_create(C.byref(h), api.handle, rec.ptr), "oaph_rts_create")
        self._h = h

    def getMatrixInfo(self) -> oaph_matrix_info:
        inf = oaph_matrix_info()
        check(self._lib.oaph_rts_info(self._h, C.byref(inf)), "oaph_rts_info")
        return inf

    def createDeviceMatrix(self) -> DeviceMatrix:
        inf = self.getMatrixInfo()
        out = DeviceMatrix(inf.columns, inf.rows, True, False)
        return self.getDeviceMatrix(out)

    def getDeviceMatrix(self, out: DeviceMatrix) -> DeviceMatrix:
        check(self._lib.oaph_rts_square(self._h, out.ptr), "oaph_rts_square")
        return out

    def createDeviceSubMatrix(self, rindex: int, rlength: int) -> DeviceMatrix:
        inf = self.getMatrixInfo()
        rows = max(1, min(rlength, inf.rows - rindex)) if rindex < inf.rows else 1
        out = DeviceMatrix(inf.columns, rows, bool(inf.isRe), bool(inf.isIm))
        return self.getDeviceSubMatrix(rindex, rlength, out)

    def getDeviceSubMatrix(self, rindex: int, rlength: int, out: DeviceMatrix | None) -> DeviceMatrix:
        if out is None:
            return self.createDeviceSubMatrix(rindex, rlength)
        check(self._lib.oaph_rts_sub_matrix(self._h, rindex, rlength, out.ptr),
              "oaph_rts_sub_matrix")
        return out

    def createDeviceRowVector(self, index: int) -> DeviceMatrix:
        inf = self.getMatrixInfo()
        return self.getDeviceRowVector(index, DeviceMatrix(inf.columns, 1, True, False))

    def getDeviceRowVector(self, index: int, out: DeviceMatrix | None) -> DeviceMatrix:
        if out is None:
            return self.createDeviceRowVector(index)
        check(self._lib.oaph_rts_row_vector(self._h, index, out.ptr), "oaph_rts_row_vector")
        return out

    def multiply(self, w: DeviceMatrix, v: DeviceMatrix, api: "CuProceduresApi | None" = None):
        """w = S v without forming S; `api` selects the stream (inside a
        solver callback, the api the callback receives)."""
        check(self._lib.oaph_rts_multiply(self._h, None if api is None else api.handle, w.ptr,
                                          v.ptr), "oaph_rts_multiply")

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            self._lib.oaph_rts_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass


class CovarianceBatch:
    """Many real symmetric eigenproblems C_i = X_i^T X_i / s solved together
    on the GPU (one persistent launch, one workgroup per problem).

    The operator is the reference's M M^T product (RecToSquareApi,
    oapMatrixCuda/RecToSquareApi.cpp:57-141) applied implicitly; the solver
    is the implicitly restarted Arnoldi iteration of CuHArnoldi
    (ArnoldiPackage/ArnoldiProcedures.cpp:129-140) with a symmetric
    tridiagonal projection.
    """

    def __init__(self, sizes, samples: int):
        self._lib = load()
        self.n = np.ascontiguousarray(sizes, dtype=np.uint32)
        self.s = int(samples)
        h = C.c_void_p()
        check(self._lib.oaph_batch_create(C.byref(h), len(self.n),
                                          self.n.ctypes.data_as(C.POINTER(C.c_uint32)), self.s),
              "batch_create")
        self._h = h
        self.k = 0

    @staticmethod
    def c3_sizes(count: int, seed: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint32)
        check(load().oaph_batch_c3_sizes(count, seed, out.ctypes.data_as(C.POINTER(C.c_uint32))),
              "batch_c3_sizes")
        return out

    def upload(self, i: int, X: np.ndarray) -> None:
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.shape != (self.s, int(self.n[i])):
            raise ValueError(f"X_{i} must be {self.s} x {int(self.n[i])}, got {X.shape}")
        check(self._lib.oaph_batch_upload_x(self._h, i, X.ctypes.data_as(C.POINTER(C.c_float))),
              "batch_upload_x")

    def download(self, i: int) -> np.ndarray:
        X = np.zeros((self.s, int(self.n[i])), dtype=np.float32)
        check(self._lib.oaph_batch_download_x(self._h, i, X.ctypes.data_as(C.POINTER(C.c_float))),
              "batch_download_x")
        return X

    def synth_c3(self, seed: int) -> None:
        check(self._lib.oaph_batch_synth_c3(self._h, seed), "batch_synth_c3")

    def set_variant(self, v: int) -> None:
        check(self._lib.oaph_batch_set_variant(self._h, v), "batch_set_variant")

    @staticmethod
    def variants() -> list[tuple[int, int, int]]:
        """(threads per workgroup, rows per wave step, row-ring depth) of each
        launch variant"""
        out, v = [], 0
        lib = load()
        while True:
            t, r, d = C.c_int(), C.c_int(), C.c_int()
            if lib.oaph_batch_variant_shape(v, C.byref(t), C.byref(r), C.byref(d)) != 0:
                return out
            out.append((t.value, r.value, d.value))
            v += 1

    def set_graph(self, enable: bool) -> None:
        check(self._lib.oaph_batch_set_graph(self._h, int(enable)), "batch_set_graph")

    def solve(self, hdim: int, wanted: int, which: int = 0, tol: float = 1e-10,
              max_restarts: int = 1000, seed: int = 0x5EED5EED) -> None:
        check(self._lib.oaph_batch_solve(self._h, hdim, wanted, which, tol, max_restarts, seed),
              "batch_solve")
        self.k = wanted

    def autotune(self, hdim: int, wanted: int, which: int = 0, tol: float = 1e-10,
                 max_restarts: int = 1000, seed: int = 0x5EED5EED, reps: int = 2,
                 cache: str | None = None, measure: bool = True):
        """Returns (chosen variant, per-variant ms or None when taken from the cache)."""
        nv = len(self.variants())
        times = np.zeros(nv)
        ch = C.c_int()
        check(self._lib.oaph_batch_autotune(
            self._h, hdim, wanted, which, tol, max_restarts, seed, reps,
            cache.encode() if cache else None, C.byref(ch),
            times.ctypes.data_as(C.POINTER(C.c_double)) if measure else None), "batch_autotune")
        return ch.value, (times if measure else None)

    def results(self) -> dict:
        cnt, k = len(self.n), self.k
        ev, rs = np.zeros((cnt, k)), np.zeros((cnt, k))
        nconv = np.zeros(cnt, dtype=np.uint32)
        rst = np.zeros(cnt, dtype=np.uint32)
        mv = np.zeros(cnt, dtype=np.uint32)
        ms = C.c_double()
        u32 = C.POINTER(C.c_uint32)
        check(self._lib.oaph_batch_results(
            self._h, ev.ctypes.data_as(C.POINTER(C.c_double)),
            rs.ctypes.data_as(C.POINTER(C.c_double)), nconv.ctypes.data_as(u32),
            rst.ctypes.data_as(u32), mv.ctypes.data_as(u32), C.byref(ms)), "batch_results")
        return {"evals": ev, "resid": rs, "nconv": nconv, "restarts": rst, "matvecs": mv,
                "kernel_ms": ms.value}

    def phase_cycles(self) -> np.ndarray:
        """count x 5 shader-clock cycles: operator, orthogonalisation,
        restarts, basis update, Ritz vectors."""
        c = np.zeros((len(self.n), 5), dtype=np.int64)
        check(self._lib.oaph_batch_phase_cycles(self._h, c.ctypes.data_as(C.POINTER(C.c_longlong))),
              "batch_phase_cycles")
        return c

    def eigenvectors(self, i: int) -> np.ndarray:
        x = np.zeros((self.k, int(self.n[i])))
        check(self._lib.oaph_batch_eigenvectors(self._h, i, x.ctypes.data_as(C.POINTER(C.c_double))),
              "batch_eigenvectors")
        return x

    def destroy(self) -> None:
        if getattr(self, "_h", None):
            self._lib.oaph_batch_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass


__all__ += ["CovarianceBatch"]


class Network:
    """oap::Network (oapNeuralDevice/oapNetwork.h:33-120) in FP64 on the GPU:
    per-sample forwardPropagation / accumulateErrors(MEAN_SQUARE_ERROR) /
    backPropagation and updateWeights after a batch."""

    def __init__(self, api: "CuProceduresApi | None" = None):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.oaph_network_create(C.byref(h), api._h if api else None), "network_create")
        self._h = h
        self._layers = []

    def createLayer(self, neurons: int, addBias: bool = False, activation: int = Activation.SIGMOID):
        check(self._lib.oaph_network_add_layer(self._h, neurons, int(addBias), activation),
              "network_add_layer")
        self._layers.append((neurons, bool(addBias), activation))

    def getLayersCount(self) -> int:
        return len(self._layers)

    def weightsShape(self, layer: int):
        n, b, _ = self._layers[layer]
        return (self._layers[layer + 1][0], n + int(b))

    def setLearningRate(self, lr: float):
        check(self._lib.oaph_network_set_learning_rate(self._h, lr), "set_learning_rate")

    def setHostWeights(self, w, layer: int):
        w = np.ascontiguousarray(w, dtype=np.float64).ravel()
        if w.size != int(np.prod(self.weightsShape(layer))):
            raise ValueError("weights size")
        check(self._lib.oaph_network_set_host_weights(self._h, layer, w.ctypes.data_as(DP)),
              "set_host_weights")

    def getHostWeights(self, layer: int) -> np.ndarray:
        w = np.zeros(int(np.prod(self.weightsShape(layer))))
        check(self._lib.oaph_network_get_host_weights(self._h, layer, w.ctypes.data_as(DP)),
              "get_host_weights")
        return w

    def setInputs(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(self._lib.oaph_network_set_inputs(self._h, x.ctypes.data_as(DP)), "set_inputs")

    def setExpected(self, t):
        t = np.ascontiguousarray(np.atleast_1d(t), dtype=np.float64)
        check(self._lib.oaph_network_set_expected(self._h, t.ctypes.data_as(DP)), "set_expected")

    def forwardPropagation(self):
        check(self._lib.oaph_network_forward(self._h), "forward")

    def getOutputs(self) -> np.ndarray:
        y = np.zeros(self._layers[-1][0])
        check(self._lib.oaph_network_get_outputs(self._h, y.ctypes.data_as(DP)), "get_outputs")
        return y

    def accumulateErrors(self):
        check(self._lib.oaph_network_accumulate_errors(self._h), "accumulate_errors")

    def backPropagation(self):
        check(self._lib.oaph_network_backward(self._h), "backward")

    def updateWeights(self):
        check(self._lib.oaph_network_update_weights(self._h), "update_weights")

    def postStep(self):
        check(self._lib.oaph_network_post_step(self._h), "post_step")

    def calculateError(self) -> float:
        v = C.c_double()
        check(self._lib.oaph_network_calculate_error(self._h, C.byref(v)), "calculate_error")
        return v.value

    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.oaph_network_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass


__all__ += ["Activation", "Network"]


class MLP:
    """The learned-start network on the GPU (oaph_mlp_*): tanh hidden layers,
    linear output, BF16 operands / FP32 accumulation, FP32 master weights,
    Adam."""

    def __init__(self, dims, max_batch: int = 512):
        self._lib = load()
        self.dims = [int(d) for d in dims]
        arr = (C.c_uint32 * len(self.dims))(*self.dims)
        h = C.c_void_p()
        check(self._lib.oaph_mlp_create(C.byref(h), arr, len(self.dims), max_batch), "mlp_create")
        self._h = h
        self.max_batch = max_batch
        n = C.c_size_t()
        check(self._lib.oaph_mlp_num_params(self._h, C.byref(n)), "mlp_num_params")
        self.num_params = n.value

    @staticmethod
    def _f(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    @staticmethod
    def _fp(a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def init_params(self, seed: int = 0) -> np.ndarray:
        """Glorot-uniform weights, zero biases."""
        rng = np.random.default_rng(seed)
        parts = []
        for din, dout in zip(self.dims[:-1], self.dims[1:]):
            lim = np.sqrt(6.0 / (din + dout))
            parts += [rng.uniform(-lim, lim, dout * din), np.zeros(dout)]
        p = self._f(np.concatenate(parts))
        self.set_params(p)
        return p

    def set_params(self, p):
        p = self._f(p)
        if p.size != self.num_params:
            raise ValueError("parameter count")
        check(self._lib.oaph_mlp_set_params(self._h, self._fp(p)), "mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.num_params, dtype=np.float32)
        check(self._lib.oaph_mlp_get_params(self._h, self._fp(p)), "mlp_get_params")
        return p

    def set_output_weights(self, w):
        w = self._f(w)
        check(self._lib.oaph_mlp_set_output_weights(self._h, self._fp(w)), "mlp_set_output_weights")

    def predict(self, X) -> np.ndarray:
        X = self._f(X)
        Y = np.zeros((X.shape[0], self.dims[-1]), dtype=np.float32)
        check(self._lib.oaph_mlp_predict(self._h, self._fp(X), X.shape[0], self._fp(Y)), "mlp_predict")
        return Y

    def gradients(self, X, T):
        X, T = self._f(X), self._f(T)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = C.c_float()
        check(self._lib.oaph_mlp_gradients(self._h, self._fp(X), self._fp(T), X.shape[0],
                                           self._fp(g), C.byref(loss)), "mlp_gradients")
        return float(loss.value), g

    def set_dataset(self, X, T):
        X, T = self._f(X), self._f(T)
        check(self._lib.oaph_mlp_set_dataset(self._h, self._fp(X), self._fp(T), X.shape[0]),
              "mlp_set_dataset")

    def train(self, batch: int, epochs: int, lr: float, seed: int = 1) -> np.ndarray:
        losses = np.zeros(epochs, dtype=np.float32)
        check(self._lib.oaph_mlp_train(self._h, batch, epochs, lr, seed, self._fp(losses)),
              "mlp_train")
        return losses

    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.oaph_mlp_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.destroy()
        except Exception:
            pass


__all__ += ["MLP"]
