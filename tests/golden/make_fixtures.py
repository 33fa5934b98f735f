"""Writes the committed golden fixtures under tests/golden/.

Run in the build container (where the reference tree and numpy/scipy are
present):  python tests/golden/make_fixtures.py [--c1]

Two kinds of fixtures:
  * data the reference's own tests hold, extracted as numbers from its test
    headers (read as text, never compiled or executed here):
      - oapQRTestSamples/qrtest1-9.h, oapHostTests/host_qrtest1-4.h  -> qr.json
      - oapArnoldiDeviceTests/schur1-3.h                             -> schur.json
      - oapCMatrixDataTests/CMatrixData1-4.h with the (wanted, hdim)
        pairs of oapArnoldiPackageMatricesTests.cpp:224-266          -> cmatrixdata.json
      - oapTestsData/oapDotProductTests_Data_1-4.h and
        oapTensorProductTests_Data_1.h with the shapes their tests use -> products.json
      - the expected eigenvalues of the data1-7 callback tests
        (oapArnoldiPackageCallbackTest.cpp:153-187; operators absent)  -> callback_eigs.json
      - the TransponseReMatrixExTest1-3 arrays of
        oapDeviceTests/oapMatrixCudaTest.cpp:280-375                  -> transpose_ex.json
  * dense LAPACK eigenvalues (numpy) of the seeded synthetic matrices the
    oracle generates: c0_eigs.json (64 x 64 real), c1_top8.json (8192 x 8192
    complex, optional, slow: --c1).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = Path("/root/reference")
sys.path.insert(0, str(ROOT))

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def numbers(text: str) -> list[float]:
    return [float(x) for x in re.findall(NUM, text)]


def c_string(src: str, name: str) -> str:
    """concatenated C string literal assigned to `name`"""
    m = re.search(r"\b" + name + r"\s*=\s*((?:\"[^\"]*\"\s*)+);", src, re.S)
    if not m:
        raise KeyError(name)
    return "".join(re.findall(r"\"([^\"]*)\"", m.group(1)))


def c_array(src: str, name: str) -> list[float]:
    m = re.search(r"\b" + name + r"\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}", src, re.S)
    if not m:
        raise KeyError(name)
    return numbers(m.group(1))


def matrix_string(s: str) -> list[float]:
    """oap's text matrix "[a, b, ...]" (square when no dimensions are given);
    "x <repeats N times>" (gdb's run-length form, used by schur3.h) expands to N copies"""
    body = s[s.index("[") + 1:s.rindex("]")]
    body = re.sub(r"(" + NUM + r")\s*<repeats\s+(\d+)\s+times>",
                  lambda m: ",".join([m.group(1)] * int(m.group(2))), body)
    return numbers(body)


def qr_fixtures():
    out = []
    for i in range(1, 10):
        p = REF / "oapQRTestSamples" / f"qrtest{i}.h"
        src = p.read_text()
        vals = matrix_string(c_string(src, "matrix"))
        m = int(round(len(vals) ** 0.5))
        out.append({"name": f"qrtest{i}", "m": m, "matrix": vals})
    for i in range(1, 5):
        p = REF / "oapHostTests" / f"host_qrtest{i}.h"
        src = p.read_text()
        vals = matrix_string(c_string(src, "matrix"))
        m = int(round(len(vals) ** 0.5))
        rec = {"name": f"host_qrtest{i}", "m": m, "matrix": vals}
        for key in ("qref", "rref"):
            try:
                rec[key] = matrix_string(c_string(src, key))
            except KeyError:
                pass
        out.append(rec)
    return out


def schur_fixtures():
    out = []
    for i in range(1, 4):
        src = (REF / "oapArnoldiDeviceTests" / f"schur{i}.h").read_text()
        a = matrix_string(c_string(src, "matrix"))
        e = matrix_string(c_string(src, "eq_matrix"))
        m = int(round(len(a) ** 0.5))
        out.append({"name": f"schur{i}", "m": m, "matrix": a, "eq_matrix": e})
    return out


def cmatrixdata_fixtures():
    # (data file, wanted, hdim, tolerance) as oapArnoldiPackageMatricesTests.cpp:224-266 runs them
    runs = [("CMatrixData1", 2, 6), ("CMatrixData2", 3, 6), ("CMatrixData3", 6, 12),
            ("CMatrixData4", 6, 12)]
    cases = []
    for name, wanted, hdim in runs:
        src = (REF / "oapCMatrixDataTests" / f"{name}.h").read_text()
        cols = int(re.search(r"columns\s*=\s*(\d+)", src).group(1))
        rows = int(re.search(r"rows\s*=\s*(\d+)", src).group(1))
        cases.append({"name": name, "columns": cols, "rows": rows, "wanted": wanted,
                      "hdim": hdim, "matrix": c_array(src, "matrix"),
                      "eigenvalues": c_array(src, "eigenvalues")})
    return {"cases": cases}


def products_fixtures():
    d = {}
    # Data_4 holds inputs only (no expected output): not a golden vector
    for i in range(1, 4):
        src = (REF / "oapTestsData" / f"oapDotProductTests_Data_{i}.h").read_text()
        d[f"dot_{i}"] = {k: c_array(src, k) for k in ("t_reValues1", "t_reValues2",
                                                      "t_outputValues")}
    src = (REF / "oapTestsData" / "oapTensorProductTests_Data_1.h").read_text()
    d["tensor_1"] = {k: c_array(src, k) for k in ("t_reValues1", "t_reValues2",
                                                  "t_outputValues")}
    return d


def transpose_ex_fixtures():
    """The TransponseReMatrixExTest1-3 cases of oapDeviceTests/oapMatrixCudaTest.cpp
    :280-375: input array, output and input sizes, the MatrixEx and the expected
    output (the device output matrix starts zeroed)."""
    src = (REF / "oapDeviceTests" / "oapMatrixCudaTest.cpp").read_text()
    cases = []
    for i in (1, 2, 3):
        start = src.index(f"TransponseReMatrixExTest{i})")
        body = src[start:src.index("TEST_F", start + 10)]
        arr = numbers(body[body.index("hArray[] = {") + 11:body.index("};", body.index("hArray[]"))])
        out = numbers(body[body.index("hOutputArray[] = {") + 17:
                           body.index("};", body.index("hOutputArray[]"))])
        dev = re.findall(r"NewDeviceMatrix\s*\(output,\s*(\d+),\s*(\d+)\)", body)
        ex = [int(v) for v in re.search(r"hMatrixEx\s*=\s*\{([^}]*)\}", body).group(1).split(",")]
        eq = re.search(r"eq_output\s*=\s*oap::host::NewReMatrixCopy\((\d+),\s*(\d+)", body)
        cases.append({"name": f"TransponseReMatrixExTest{i}", "input": arr,
                      "output_columns": int(dev[0][0]), "output_rows": int(dev[0][1]),
                      "input_columns": int(dev[1][0]), "input_rows": int(dev[1][1]),
                      "matrix_ex": ex, "expected": out,
                      "expected_columns": int(eq.group(1)), "expected_rows": int(eq.group(2))})
    return {"cases": cases}


def callback_eigs():
    src = (REF / "oapArnoldiDeviceTests" / "oapArnoldiPackageCallbackTest.cpp").read_text()
    vals = re.findall(r"executeArnoldiTest\s*\(\s*(" + NUM + r")\s*,\s*\"([^\"]+)\"", src)
    return {"expected": [{"value": float(v), "data": d} for v, d in vals],
            "note": "operators (.tdata) are absent offline; values only"}


def c0_eigs():
    from oracle.pyoracle import Oracle
    A = Oracle().gen_c0(64, 0x5EED0000)
    ev = np.linalg.eigvals(A)
    ev = ev[np.argsort(-np.abs(ev), kind="stable")]
    return {"n": 64, "seed": 0x5EED0000, "re": ev.real.tolist(), "im": ev.imag.tolist(),
            "fro": float(np.linalg.norm(A)), "sum": float(A.sum())}


def c1_top8():
    """The 16 largest-|lambda| eigenvalues of the dense complex 8192 x 8192
    C1 matrix (the solver wants 8; the next 8 name a wrong wanted set).  Two
    ARPACK runs from different start vectors must agree, and every pair's
    explicit residual is recorded."""
    import scipy.sparse.linalg as sla
    from oracle.pyoracle import Oracle
    n, kk = 8192, 16
    ar, ai = Oracle().gen_c1(n, 0x5EED0001, 1e-2, 500.0)
    A = ar + 1j * ai
    del ar, ai
    runs = []
    for seed in (1, 2):
        v0 = np.random.default_rng(seed).standard_normal(n) + 0j
        ev, vec = sla.eigs(A, k=kk, which="LM", ncv=64, tol=1e-14, maxiter=50000, v0=v0)
        order = np.argsort(-np.abs(ev), kind="stable")
        runs.append((ev[order], vec[:, order]))
    (ev, vec), (ev2, _) = runs
    agree = max(float(np.min(np.abs(ev2 - z))) for z in ev)
    resid = [float(np.linalg.norm(A @ vec[:, t] - ev[t] * vec[:, t]) / np.linalg.norm(vec[:, t]))
             for t in range(kk)]
    fro = float(np.linalg.norm(A))
    assert agree <= 1e-13 and max(resid) <= 1e-13 * fro, (agree, max(resid))
    return {"n": n, "seed": 0x5EED0001, "k": 8, "re": ev.real.tolist(), "im": ev.imag.tolist(),
            "fro": fro, "residuals": resid, "two_start_agreement": agree,
            "method": "ARPACK zneupd (scipy.sparse.linalg.eigs), k=16, ncv=64, tol 1e-14, "
                      "two start vectors"}


def main():
    HERE.mkdir(exist_ok=True)
    if REF.exists():
        (HERE / "qr.json").write_text(json.dumps(qr_fixtures()))
        (HERE / "schur.json").write_text(json.dumps(schur_fixtures()))
        (HERE / "cmatrixdata.json").write_text(json.dumps(cmatrixdata_fixtures()))
        (HERE / "products.json").write_text(json.dumps(products_fixtures()))
        (HERE / "callback_eigs.json").write_text(json.dumps(callback_eigs()))
        (HERE / "transpose_ex.json").write_text(json.dumps(transpose_ex_fixtures()))
    (HERE / "c0_eigs.json").write_text(json.dumps(c0_eigs()))
    if "--c1" in sys.argv:
        (HERE / "c1_top8.json").write_text(json.dumps(c1_top8()))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
