"""ctypes wrappers of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  `Oracle` wraps liboracle.so (the C restatement
in oracle.c); `Ref` wraps _ref/libref.so (the reference's own host sources,
built only where the reference tree is present).

Matrices are passed as complex numpy arrays of shape (rows, columns) plus
flags for which planes exist (a plane that does not exist is NULL).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DP = C.POINTER(C.c_double)
U32 = C.c_uint32


class orc_matrix(C.Structure):
    _fields_ = [("columns", U32), ("rows", U32), ("re", DP), ("im", DP)]


class orc_iram_opts(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("k", C.c_int), ("which", C.c_int),
                ("tol", C.c_double), ("anorm", C.c_double), ("max_restarts", C.c_int),
                ("seed", C.c_uint64), ("real_start", C.c_int), ("v0re", DP), ("v0im", DP)]


class orc_iram_result(C.Structure):
    _fields_ = [("nconv", C.c_int), ("restarts", C.c_int), ("matvecs", C.c_int)]


MATVEC_FN = C.CFUNCTYPE(None, C.c_void_p, DP, DP, DP, DP)


class Planes:
    """Host matrix with optional planes (keeps the numpy buffers alive)."""

    def __init__(self, rows: int, cols: int, is_re=True, is_im=True, z=None):
        self.rows, self.cols = rows, cols
        self.re = np.zeros((rows, cols)) if is_re else None
        self.im = np.zeros((rows, cols)) if is_im else None
        if z is not None:
            z = np.asarray(z).reshape(rows, cols)
            if self.re is not None:
                self.re[:] = np.real(z)
            if self.im is not None:
                self.im[:] = np.imag(z)

    @classmethod
    def of(cls, z, is_re=True, is_im=True):
        z = np.asarray(z)
        if z.ndim == 1:
            z = z.reshape(-1, 1)
        return cls(z.shape[0], z.shape[1], is_re, is_im, z)

    def c(self) -> orc_matrix:
        return orc_matrix(self.cols, self.rows, _p(self.re), _p(self.im))

    def z(self) -> np.ndarray:
        out = np.zeros((self.rows, self.cols), dtype=np.complex128)
        if self.re is not None:
            out.real = self.re
        if self.im is not None:
            out.imag = self.im
        return out


def _p(a):
    return None if a is None else a.ctypes.data_as(DP)


def _dims(d):
    arr = ((U32 * 2) * 3)()
    for i in range(3):
        for j in range(2):
            arr[i][j] = int(d[i][j])
    return arr


class Oracle:
    def __init__(self, path: str | Path | None = None):
        p = Path(path) if path else HERE / "liboracle.so"
        if not p.exists():
            raise FileNotFoundError(f"{p} missing: run `make -C oracle`")
        self.lib = C.CDLL(str(p))
        L = self.lib
        MP = C.POINTER(orc_matrix)
        for name in ("orc_dot_product", "orc_transpose", "orc_conjugate_transpose",
                     "orc_hadamard_product", "orc_hadamard_product_vec", "orc_add", "orc_substract"):
            getattr(L, name).argtypes = [MP, MP] + ([MP] if name not in (
                "orc_transpose", "orc_conjugate_transpose") else [])
            getattr(L, name).restype = C.c_int
        L.orc_dot_product_dims.argtypes = [MP, MP, MP, (U32 * 2) * 3]
        L.orc_tensor_product.argtypes = [MP, MP, MP, (U32 * 2) * 3]
        L.orc_multiply_re_constant.argtypes = [MP, MP, C.c_double]
        L.orc_multiply_constant.argtypes = [MP, MP, C.c_double, C.c_double]
        L.orc_set_diagonal.argtypes = [MP, C.c_double, C.c_double]
        L.orc_set_identity.argtypes = [MP]
        L.orc_set_vector.argtypes = [MP, U32, MP, U32]
        L.orc_get_vector.argtypes = [MP, U32, MP, U32]
        L.orc_magnitude.argtypes = [MP]
        L.orc_magnitude.restype = C.c_double
        L.orc_sum.argtypes = [MP, DP, DP]
        L.orc_qrgr.argtypes = [MP, MP, MP]
        L.orc_is_upper_triangular.argtypes = [MP]
        L.orc_calc_triangular_h.argtypes = [MP, MP, U32]
        L.orc_set_threads_blocks.argtypes = [C.POINTER(U32), C.POINTER(U32), U32, U32, U32]
        L.orc_hess_schur.argtypes = [C.c_int] + [DP] * 6
        L.orc_trevc.argtypes = [C.c_int] + [DP] * 4
        L.orc_iram.argtypes = [C.POINTER(orc_iram_opts), MATVEC_FN, C.c_void_p] + [DP] * 5 + [
            C.POINTER(orc_iram_result)]
        L.orc_dense_matvec.argtypes = [C.c_int] + [DP] * 6 + [C.c_int]
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_uniform.restype = C.c_double
        L.orc_normal.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_normal.restype = C.c_double
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_gen_c0.argtypes = [C.c_int, C.c_uint64, DP]
        L.orc_gen_c1.argtypes = [C.c_int, C.c_uint64, C.c_double, C.c_double, DP, DP]
        L.orc_gen_c3_problem.argtypes = [C.c_int, C.c_uint64, C.c_int, C.POINTER(C.c_float)]
        L.orc_gen_c3_problem.restype = C.c_int
        IP = C.POINTER(C.c_int)
        L.orc_net_create.argtypes = [C.c_int, IP, IP, IP]
        L.orc_net_create.restype = C.c_void_p
        L.orc_net_free.argtypes = [C.c_void_p]
        for nm in ("orc_net_set_weights", "orc_net_get_weights", "orc_net_gradient"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int, DP]
        L.orc_net_forward.argtypes = [C.c_void_p, DP, DP]
        L.orc_net_accumulate.argtypes = [C.c_void_p, DP]
        L.orc_net_backward.argtypes = [C.c_void_p]
        L.orc_net_update.argtypes = [C.c_void_p, C.c_double]
        L.orc_net_post_step.argtypes = [C.c_void_p]
        L.orc_net_error.argtypes = [C.c_void_p]
        L.orc_net_error.restype = C.c_double
        L.orc_gen_cdr.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, DP]
        FP = C.POINTER(C.c_float)
        L.orc_bf16.argtypes = [C.c_double]
        L.orc_bf16.restype = C.c_double
        L.orc_mlp_grad.argtypes = [C.POINTER(C.c_uint32), C.c_int, FP, FP, FP, FP, C.c_int, DP,
                                   C.c_int]
        L.orc_mlp_grad.restype = C.c_double

    # ---- procedures (return the output Planes)
    def dot_product(self, out: Planes, a: Planes, b: Planes, dims=None) -> int:
        if dims is None:
            return self.lib.orc_dot_product(C.byref(out.c()), C.byref(a.c()), C.byref(b.c()))
        return self.lib.orc_dot_product_dims(C.byref(out.c()), C.byref(a.c()), C.byref(b.c()),
                                             _dims(dims))

    def transpose(self, out: Planes, a: Planes, conj=False) -> int:
        f = self.lib.orc_conjugate_transpose if conj else self.lib.orc_transpose
        return f(C.byref(out.c()), C.byref(a.c()))

    def tensor_product(self, out, a, b, dims) -> int:
        return self.lib.orc_tensor_product(C.byref(out.c()), C.byref(a.c()), C.byref(b.c()),
                                           _dims(dims))

    def binary(self, op: str, out, a, b) -> int:
        f = {"add": self.lib.orc_add, "substract": self.lib.orc_substract,
             "hadamard": self.lib.orc_hadamard_product,
             "hadamard_vec": self.lib.orc_hadamard_product_vec}[op]
        return f(C.byref(out.c()), C.byref(a.c()), C.byref(b.c()))

    def multiply_constant(self, out, a, re, im=0.0, re_only=False) -> int:
        if re_only:
            return self.lib.orc_multiply_re_constant(C.byref(out.c()), C.byref(a.c()), re)
        return self.lib.orc_multiply_constant(C.byref(out.c()), C.byref(a.c()), re, im)

    def set_diagonal(self, m, re, im) -> int:
        return self.lib.orc_set_diagonal(C.byref(m.c()), re, im)

    def set_vector(self, V, column, v, length) -> int:
        return self.lib.orc_set_vector(C.byref(V.c()), column, C.byref(v.c()), length)

    def get_vector(self, v, length, V, column) -> int:
        return self.lib.orc_get_vector(C.byref(v.c()), length, C.byref(V.c()), column)

    def magnitude(self, m) -> float:
        return self.lib.orc_magnitude(C.byref(m.c()))

    def sum(self, m):
        r, i = C.c_double(), C.c_double()
        self.lib.orc_sum(C.byref(m.c()), C.byref(r), C.byref(i))
        return r.value, i.value

    def qrgr(self, Q, R, A) -> int:
        return self.lib.orc_qrgr(C.byref(Q.c()), C.byref(R.c()), C.byref(A.c()))

    def is_upper_triangular(self, m) -> bool:
        return bool(self.lib.orc_is_upper_triangular(C.byref(m.c())))

    def calc_triangular_h(self, H, Q, max_steps=10000) -> int:
        return self.lib.orc_calc_triangular_h(C.byref(H.c()), C.byref(Q.c()), max_steps)

    def set_threads_blocks(self, w, h, limit=1024):
        b, t = (U32 * 2)(), (U32 * 2)()
        self.lib.orc_set_threads_blocks(b, t, w, h, limit)
        return (b[0], b[1]), (t[0], t[1])

    # ---- eigen-machinery
    def hess_schur(self, H: np.ndarray):
        m = H.shape[0]
        hr = np.ascontiguousarray(H.real, dtype=np.float64).copy()
        hi = np.ascontiguousarray(H.imag, dtype=np.float64).copy()
        zr, zi = np.zeros((m, m)), np.zeros((m, m))
        wr, wi = np.zeros(m), np.zeros(m)
        st = self.lib.orc_hess_schur(m, _p(hr), _p(hi), _p(zr), _p(zi), _p(wr), _p(wi))
        return st, hr + 1j * hi, zr + 1j * zi, wr + 1j * wi

    def iram(self, matvec, n, m, k, which=0, tol=1e-10, anorm=1.0, max_restarts=1000, seed=1,
             real_start=True, v0=None):
        """matvec(x complex (n,)) -> complex (n,).  Returns (evals, evecs (k, n), resid, stats)."""

        def cb(_u, xr, xi, yr, yi):
            x = np.ctypeslib.as_array(xr, (n,)) + 1j * np.ctypeslib.as_array(xi, (n,))
            y = matvec(x)
            np.ctypeslib.as_array(yr, (n,))[:] = y.real
            np.ctypeslib.as_array(yi, (n,))[:] = y.imag

        fn = MATVEC_FN(cb)
        o = orc_iram_opts(n, m, k, which, tol, anorm, max_restarts, seed, int(real_start))
        if v0 is not None:
            v0r = np.ascontiguousarray(np.real(v0), dtype=np.float64)
            v0i = np.ascontiguousarray(np.imag(v0), dtype=np.float64)
            o.v0re, o.v0im = _p(v0r), _p(v0i)
        er, ei = np.zeros(k), np.zeros(k)
        xr, xi = np.zeros(k * n), np.zeros(k * n)
        rs = np.zeros(k)
        res = orc_iram_result()
        st = self.lib.orc_iram(C.byref(o), fn, None, _p(er), _p(ei), _p(xr), _p(xi), _p(rs),
                               C.byref(res))
        return (st, er + 1j * ei, (xr + 1j * xi).reshape(k, n), rs,
                {"nconv": res.nconv, "restarts": res.restarts, "matvecs": res.matvecs})

    def dense_matvec(self, are, aim, x, threads=1):
        n = are.shape[0]
        xr = np.ascontiguousarray(x.real)
        xi = np.ascontiguousarray(x.imag)
        yr, yi = np.zeros(n), np.zeros(n)
        self.lib.orc_dense_matvec(n, _p(are), _p(aim), _p(xr), _p(xi), _p(yr), _p(yi), threads)
        return yr + 1j * yi

    # ---- generators
    def gen_c0(self, n=64, seed=0x5EED0000):
        a = np.zeros((n, n))
        self.lib.orc_gen_c0(n, seed, _p(a))
        return a

    def gen_cdr(self, N, pe, c, alpha):
        """C4 operator (N^2 x N^2 dense, row-major)."""
        a = np.zeros((N * N, N * N))
        self.lib.orc_gen_cdr(N, pe, c, alpha, _p(a))
        return a

    def gen_c3(self, i, seed=0x5EED0003, s=256):
        """Problem i of the seeded covariance batch: X (s x n float32)."""
        n = self.lib.orc_gen_c3_problem(i, seed, s, None)
        X = np.zeros((s, n), dtype=np.float32)
        self.lib.orc_gen_c3_problem(i, seed, s, X.ctypes.data_as(C.POINTER(C.c_float)))
        return X

    def gen_c1(self, n=8192, seed=0x5EED0001, eps=1e-2, decay=500.0):
        ar, ai = np.zeros((n, n)), np.zeros((n, n))
        self.lib.orc_gen_c1(n, seed, eps, decay, _p(ar), _p(ai))
        return ar, ai

    def uniform(self, seed, idx) -> float:
        return self.lib.orc_uniform(seed, idx)

    def normal(self, seed, idx) -> float:
        return self.lib.orc_normal(seed, idx)

    def splitmix64(self, x) -> int:
        return self.lib.orc_splitmix64(x)


class Ref:
    """The reference's own host code (oracle/_ref/libref.so)."""

    PATH = HERE / "_ref" / "libref.so"

    @classmethod
    def available(cls) -> bool:
        return cls.PATH.exists()

    def __init__(self):
        self.lib = C.CDLL(str(self.PATH))

    @staticmethod
    def _args(p: Planes):
        return [int(p.re is not None), int(p.im is not None), p.cols, p.rows, _p(p.re), _p(p.im)]

    # ---- matrix formats (the reference's own ToString / NewMatrix / SaveMatrix)
    def to_string(self, a: Planes) -> str:
        f = self.lib.ref_to_string
        f.restype = C.c_size_t
        n = f(*self._args(a), None, C.c_size_t(0))
        buf = C.create_string_buffer(n + 1)
        f(*self._args(a), buf, C.c_size_t(n + 1))
        return buf.value.decode()

    def parse_matrix(self, text: str, cap: int = 1 << 20):
        """NewMatrix(text): (columns, rows, re | None, im | None) or None."""
        isre, isim, c, r = C.c_int(), C.c_int(), U32(), U32()
        re, im = np.zeros(cap), np.zeros(cap)
        ok = self.lib.ref_parse_matrix(text.encode(), C.byref(isre), C.byref(isim), C.byref(c),
                                       C.byref(r), _p(re), _p(im), C.c_size_t(cap))
        if not ok:
            return None
        n = c.value * r.value
        assert n <= cap
        shape = (r.value, c.value)
        return (c.value, r.value, re[:n].reshape(shape) if isre.value else None,
                im[:n].reshape(shape) if isim.value else None)

    def create_array(self, text: str, which: int, cap: int = 1 << 20):
        f = self.lib.ref_create_array
        f.restype = C.c_long
        out = np.zeros(cap)
        n = f(text.encode(), C.c_uint(which), _p(out), C.c_size_t(cap))
        return None if n < 0 else out[:n].copy()

    def save_matrix(self, a: Planes | None) -> bytes:
        f = self.lib.ref_save_matrix
        f.restype = C.c_size_t
        args = [1, 0, 0, 0, 0, None, None] if a is None else [0] + self._args(a)
        n = f(*args, None, C.c_size_t(0))
        buf = (C.c_uint8 * n)()
        f(*args, buf, C.c_size_t(n))
        return bytes(buf)

    def load_matrix(self, data: bytes, cap: int = 1 << 20):
        isre, isim, c, r = C.c_int(), C.c_int(), U32(), U32()
        re, im = np.zeros(cap), np.zeros(cap)
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
        ok = self.lib.ref_load_matrix(buf, C.c_size_t(len(data)), C.byref(isre), C.byref(isim),
                                      C.byref(c), C.byref(r), _p(re), _p(im), C.c_size_t(cap))
        if not ok:
            return None
        n = c.value * r.value
        shape = (r.value, c.value)
        return (c.value, r.value, re[:n].reshape(shape) if isre.value else None,
                im[:n].reshape(shape) if isim.value else None)

    def opt_size(self, bitmap: np.ndarray):
        a = np.ascontiguousarray(bitmap, dtype=np.uint8)
        h, w = a.shape[:2]
        c = 1 if a.ndim == 2 else a.shape[2]
        out = (U32 * 4)()
        self.lib.ref_opt_size(a.ctypes.data_as(C.c_void_p), U32(w), U32(h), U32(c), out)
        return tuple(out)

    def buffer_fwrite(self, path: str, data: bytes) -> None:
        buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data + b"\0")
        self.lib.ref_buffer_fwrite(str(path).encode(), buf, C.c_size_t(len(data)))

    def set_threads_blocks(self, w, h, limit=1024):
        b, t = (U32 * 2)(), (U32 * 2)()
        self.lib.ref_set_threads_blocks(b, t, U32(w), U32(h), U32(limit))
        return (b[0], b[1]), (t[0], t[1])

    def dot_product(self, out, a, b, dims=None):
        if dims is None:
            self.lib.ref_dot_product(*self._args(out), *self._args(a), *self._args(b))
        else:
            d = (U32 * 6)(*[int(v) for row in dims for v in row])
            self.lib.ref_dot_product_dims(*self._args(out), *self._args(a), *self._args(b), d)

    def transpose(self, out, a, conj=False):
        f = self.lib.ref_conjugate_transpose if conj else self.lib.ref_transpose
        f(*self._args(out), *self._args(a))

    def tensor_product(self, out, a, b, dims):
        d = (U32 * 6)(*[int(v) for row in dims for v in row])
        self.lib.ref_tensor_product(*self._args(out), *self._args(a), *self._args(b), d)

    def sum(self, a):
        r, i = C.c_double(), C.c_double()
        self.lib.ref_sum(*self._args(a), C.byref(r), C.byref(i))
        return r.value, i.value

    def set_vector(self, V, column, v, length):
        self.lib.ref_set_vector(*self._args(V), U32(column), *self._args(v), U32(length))

    def get_vector(self, v, length, V, column):
        self.lib.ref_get_vector(*self._args(v), U32(length), *self._args(V), U32(column))

    def set_identity(self, m):
        self.lib.ref_set_identity(*self._args(m))

    def binary(self, op, out, a, b):
        code = {"add": 0, "substract": 1, "hadamard": 2, "hadamard_vec": 3}[op]
        self.lib.ref_binary(C.c_int(code), *self._args(out), *self._args(a), *self._args(b))

    def multiply_constant(self, out, a, re, im=0.0, re_only=False):
        self.lib.ref_multiply_constant(C.c_int(int(re_only)), *self._args(out), *self._args(a),
                                       C.c_double(re), C.c_double(im))

    def set_diagonal(self, m, re, im):
        self.lib.ref_set_diagonal(*self._args(m), C.c_double(re), C.c_double(im))

    def magnitude(self, a) -> float:
        self.lib.ref_magnitude.restype = C.c_double
        return self.lib.ref_magnitude(*self._args(a))

    def qrgr(self, Q, R, A):
        self.lib.ref_qrgr(*self._args(Q), *self._args(R), *self._args(A))

    def is_upper_triangular(self, a) -> bool:
        return bool(self.lib.ref_is_upper_triangular(*self._args(a)))


def mlp_grad(oracle, dims, params, X, T, w, round_bf16=True):
    """(loss, flat gradients) of the learned-start MLP on one batch."""
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    params, X, T, w = f(params), f(X), f(T), f(w)
    d = (C.c_uint32 * len(dims))(*dims)
    g = np.zeros(params.size)
    FP = C.POINTER(C.c_float)
    loss = oracle.lib.orc_mlp_grad(d, len(dims), params.ctypes.data_as(FP), X.ctypes.data_as(FP),
                                   T.ctypes.data_as(FP), w.ctypes.data_as(FP), X.shape[0], _p(g),
                                   int(round_bf16))
    return loss, g


class OracleNet:
    """The oracle's FP64 network (orc_net_*): layers (neurons, bias, activation)."""

    def __init__(self, oracle: "Oracle", layers):
        self.lib = oracle.lib
        self.shape = [(int(n), bool(b), int(a)) for n, b, a in layers]
        arr = lambda v: (C.c_int * len(v))(*v)
        self.h = self.lib.orc_net_create(len(layers), arr([x[0] for x in self.shape]),
                                         arr([int(x[1]) for x in self.shape]),
                                         arr([x[2] for x in self.shape]))

    def wshape(self, l):
        n, b, _ = self.shape[l]
        return (self.shape[l + 1][0], n + int(b))

    def set_weights(self, l, w):
        w = np.ascontiguousarray(w, dtype=np.float64).ravel()
        assert w.size == np.prod(self.wshape(l))
        self.lib.orc_net_set_weights(self.h, l, _p(w))

    def get_weights(self, l):
        w = np.zeros(self.wshape(l))
        self.lib.orc_net_get_weights(self.h, l, _p(w))
        return w.ravel()

    def gradient(self, l):
        g = np.zeros(self.wshape(l))
        self.lib.orc_net_gradient(self.h, l, _p(g))
        return g.ravel()

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.shape[-1][0])
        self.lib.orc_net_forward(self.h, _p(x), _p(y))
        return y

    def accumulate(self, t):
        t = np.ascontiguousarray(np.atleast_1d(t), dtype=np.float64)
        self.lib.orc_net_accumulate(self.h, _p(t))

    def backward(self):
        self.lib.orc_net_backward(self.h)

    def update(self, lr):
        self.lib.orc_net_update(self.h, lr)

    def post_step(self):
        self.lib.orc_net_post_step(self.h)

    def error(self):
        return self.lib.orc_net_error(self.h)

    def __del__(self):  # pragma: no cover
        try:
            self.lib.orc_net_free(self.h)
        except Exception:
            pass
