"""The PCA data path of oap's oap2dt3d application on this library.

PNG images -> cropped pixel columns (``oap::DataLoader`` over ``PngFile``,
oapAppUtils/DataLoader.cpp, oapAppUtils/PngFile.cpp) -> the data matrix M
(one image per column) -> S = M M^T through ``RecToSquareApi`` -> the largest
eigenpairs by ``CuHArnoldiCallback`` with a multiply callback
(``IEigenCalculator`` / ``MainAPExecutor``, oap2dt3dDevice/IEigenCalculator.h,
oap2dt3dDevice/MainAPExecutor.cpp:129-174).  Image decoding, cropping and the
load process are host code in liboaph.so; the products run on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import DP, check, load, oaph_matrix_info
from .api import (ArnUtils, CuHArnoldiCallback, DeviceMatrix, MatrixInfo, QRType,
                  RecToSquareApi)

U32 = C.c_uint32


def write_png(path, pixels: np.ndarray) -> None:
    """pixels: uint8 (h, w) grey or (h, w, c) with c in 1..4"""
    a = np.ascontiguousarray(pixels, dtype=np.uint8)
    h, w = a.shape[:2]
    c = 1 if a.ndim == 2 else a.shape[2]
    check(load().oaph_png_write(str(path).encode(), w, h, c, a.ctypes.data_as(C.c_void_p)),
          "png_write")


def read_png(path) -> np.ndarray:
    """raw 8-bit rows as (h, w, colors)"""
    lib = load()
    w, h, c = U32(), U32(), U32()
    p = str(path).encode()
    check(lib.oaph_png_read(p, C.byref(w), C.byref(h), C.byref(c), None, 0), "png_read")
    out = np.zeros((h.value, w.value, c.value), dtype=np.uint8)
    check(lib.oaph_png_read(p, C.byref(w), C.byref(h), C.byref(c),
                            out.ctypes.data_as(C.c_void_p), out.size), "png_read")
    return out


def opt_size(bitmap: np.ndarray):
    """GetOptWidth / GetOptHeight: (width, begin column, height, begin row)"""
    a = np.ascontiguousarray(bitmap, dtype=np.uint8)
    h, w = a.shape[:2]
    c = 1 if a.ndim == 2 else a.shape[2]
    out = (U32 * 4)()
    check(load().oaph_image_opt_size(a.ctypes.data_as(C.c_void_p), w, h, c, out), "opt_size")
    return tuple(out)


def load_plan(windows):
    """DataLoader's load process on (w, wbegin, h, hbegin) windows:
    (windows used, bitmap loads per image)"""
    s = np.ascontiguousarray(np.asarray(windows, dtype=np.uint32).reshape(-1))
    n = s.size // 4
    out = np.zeros(4 * n, dtype=np.uint32)
    loads = np.zeros(n, dtype=np.uint32)
    P = C.POINTER(U32)
    check(load().oaph_dataloader_plan(s.ctypes.data_as(P), n, out.ctypes.data_as(P),
                                      loads.ctypes.data_as(P)), "dataloader_plan")
    return out.reshape(n, 4), loads


class DataLoader:
    """oap::DataLoader over PNG files (one column per image)."""

    def __init__(self, paths, truncate: bool = True):
        self._lib = load()
        enc = [str(p).encode() for p in paths]
        arr = (C.c_char_p * len(enc))(*enc)
        h = C.c_void_p()
        check(self._lib.oaph_dataloader_create(C.byref(h), arr, len(enc), int(truncate)),
              "dataloader_create")
        self._h = h
        self.count = len(enc)

    @classmethod
    def from_dir(cls, dir_path, name_base: str, count: int, truncate: bool = True):
        """createDataLoader(dirPath, nameBase, count): files nameBase<i>.png"""
        from pathlib import Path
        d = Path(dir_path)
        return cls([d / f"{name_base}{i}.png" for i in range(count)], truncate)

    def getMatrixInfo(self) -> oaph_matrix_info:
        inf = oaph_matrix_info()
        check(self._lib.oaph_dataloader_info(self._h, C.byref(inf)), "dataloader_info")
        return inf

    def window(self, i: int):
        out, loads = (U32 * 4)(), U32()
        check(self._lib.oaph_dataloader_window(self._h, i, out, C.byref(loads)), "window")
        return tuple(out), loads.value

    def createColumnVector(self, i: int) -> np.ndarray:
        v = np.zeros(self.getMatrixInfo().rows)
        check(self._lib.oaph_dataloader_column(self._h, i, v.ctypes.data_as(DP)), "column")
        return v

    def createMatrix(self) -> DeviceMatrix:
        inf = self.getMatrixInfo()
        m = DeviceMatrix(inf.columns, inf.rows, True, False)
        check(self._lib.oaph_dataloader_matrix(self._h, m.ptr), "dataloader_matrix")
        return m

    def __del__(self):  # pragma: no cover
        try:
            if getattr(self, "_h", None):
                self._lib.oaph_dataloader_destroy(self._h)
                self._h = None
        except Exception:
            pass


@dataclass
class Outcome:
    """oap2dt3dDevice/Outcome.h: eigenvalues, residual errors, eigenvectors"""
    eigenvalues: np.ndarray
    errors: np.ndarray
    eigenvectors: list


class EigenCalculator:
    """IEigenCalculator + MainAPExecutor: the largest eigenpairs of M M^T for
    the data matrix M of a DataLoader, the product applied without forming
    M M^T (RecToSquareApi implicit multiply on the solver's stream)."""

    def __init__(self, data_loader: DataLoader, tol: float = 1e-10, seed: int = 1):
        self.loader = data_loader
        self.tol, self.seed = tol, seed
        self.eigens, self.wanted = 0, 0

    def setEigensCount(self, eigens: int, wanted: int) -> None:
        if wanted > eigens:
            raise ValueError("wanted eigenpairs exceed the Krylov size")
        self.eigens, self.wanted = eigens, wanted

    def getMatrixInfo(self):
        return self.loader.getMatrixInfo()

    def calculate(self) -> Outcome:
        if not self.eigens:
            raise ValueError("setEigensCount first")
        M = self.loader.createMatrix()
        n = M.rows
        s = CuHArnoldiCallback()
        s.setQRType(QRType.QRGR)
        s.setSortType(ArnUtils.SortLargestValues)
        s.setCheckTolerance(self.tol)
        s.setOutputType(ArnUtils.DEVICE)
        s.setSeed(self.seed)
        rts = RecToSquareApi(s.api, M)

        def multiply(w, v, api, user, mt):
            user.multiply(w, v, api)

        s.setCallback(multiply, rts)
        re, im = np.zeros(self.wanted), np.zeros(self.wanted)
        s.setOutputsEigenvalues(re, im)
        vecs = [DeviceMatrix(1, n, True, True) for _ in range(self.wanted)]
        s.setOutputsEigenvectors(vecs)
        s.execute(self.eigens, self.wanted, MatrixInfo(True, False, n, n))
        errors = np.array([s.testOutcome(t) for t in range(self.wanted)])
        # S is real symmetric: remove the Ritz vector's arbitrary complex phase
        vectors = []
        for v in vecs:
            z = v.to_complex()[:, 0]
            p = z[np.argmax(np.abs(z))]
            vectors.append((z * (np.conj(p) / abs(p))).real if p != 0 else z.real)
        self._keep = (M, rts, s)
        return Outcome(re.copy(), errors, vectors)
