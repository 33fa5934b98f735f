"""Writes tests/golden/formats.json: expected results of the reference's own
matrix format code for the matrix texts its tests hold, and for a few small
synthetic matrices.

Inputs (data read as text from the reference's test headers, never compiled
from them): the string matrices of oapArnoldiDeviceTests/matrix1-7.h and
schur1-3.h, oapQRTestSamples/qrtest1.h and qrtest8.h and
oapHostTests/host_qrtest1-4.h.  Planes of more than 64 values are stored as
the SHA-256 of their float64 bytes (bit-exact check, small file).
Expected outputs come from the reference's own code compiled from its
sources by oracle/Makefile into oracle/_ref/libref.so: oap::host::NewMatrix
(text) (oapMatrixCpu/oapHostMatrixUtils.cpp:223-274), oap::host::ToString
(:576-588) and oap::host::SaveMatrix (:1050-1070).

Run in the build container (reference tree present, `make -C oracle` done):
    python tests/golden/make_format_fixtures.py
"""
import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
REF = Path("/root/reference")
sys.path.insert(0, str(ROOT))

from oracle.pyoracle import Planes, Ref  # noqa: E402

SOURCES = ([REF / "oapArnoldiDeviceTests" / f"matrix{i}.h" for i in range(1, 8)] +
           [REF / "oapArnoldiDeviceTests" / f"schur{i}.h" for i in range(1, 4)] +
           [REF / "oapQRTestSamples" / f"qrtest{i}.h" for i in (1, 8)] +
           [REF / "oapHostTests" / f"host_qrtest{i}.h" for i in range(1, 5)])


def literals(path: Path):
    """(name, value) of every `const char* name = "..." "...";` in a header"""
    src = path.read_text().replace("\\\n", "")  # C line splicing
    out = []
    for m in re.finditer(r"const\s+char\s*\*\s*(\w+)\s*=\s*((?:\"(?:[^\"\\]|\\.)*\"\s*)+);", src):
        parts = re.findall(r"\"((?:[^\"\\]|\\.)*)\"", m.group(2))
        text = "".join(parts)
        text = text.replace("\\n", "\n").replace("\\\"", "\"").replace("\\\\", "\\")
        out.append((m.group(1), text))
    return out


def digest(a):
    """full values for small planes, else the SHA-256 of the float64 bytes"""
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype="<f8").reshape(-1)
    if a.size <= 64:
        return a.tolist()
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(), "head": a[:4].tolist()}


def planes(c, r, re_, im_):
    return {"columns": c, "rows": r, "re": digest(re_), "im": digest(im_)}


def main():
    ref = Ref()
    texts = []
    for p in SOURCES:
        for name, text in literals(p):
            entry = {"source": f"{p.relative_to(REF)}:{name}", "text": text}
            a1, a2 = ref.create_array(text, 1), ref.create_array(text, 2)
            dims = re.search(r"columns=(\d+).*rows=(\d+)", text, re.S)
            if a1 is not None and dims and len(a1) < int(dims.group(1)) * int(dims.group(2)):
                # NewMatrix would adopt an array shorter than columns x rows
                # (and read past it); recorded, not parsed
                entry["expected"] = {"columns": int(dims.group(1)), "rows": int(dims.group(2)),
                                     "array_length": len(a1), "mismatch": True}
            else:
                got = ref.parse_matrix(text)
                entry["expected"] = None if got is None else planes(*got)
            entry["arrays"] = [None if a is None else digest(a) for a in (a1, a2)]
            texts.append(entry)
    synth = []
    cases = [
        ("real 3x2, negatives, runs", np.array([[1.0, 1.0, -2.5], [0.0, 0.0, 0.0]]), None),
        ("complex 2x2", np.array([[1.5, -0.25], [3.0, 4.0]]), np.array([[0.0, 2.0], [-1.0, 0.0]])),
        ("1x1 zero", np.zeros((1, 1)), None),
        ("4x1 mixed", np.array([[-0.0], [1e-10], [123456.5], [7.0]]), None),
        ("3x5 one value", np.full((5, 3), 2.0), None),
        ("im only 2x3", None, np.array([[1.0, -1.0, 0.5], [0.5, 0.5, 3.0]])),
    ]
    for label, re_, im_ in cases:
        z = re_ if re_ is not None else im_
        p = Planes(z.shape[0], z.shape[1], re_ is not None, im_ is not None)
        if re_ is not None:
            p.re[:] = re_
        if im_ is not None:
            p.im[:] = im_
        synth.append({"label": label, "matrix": planes(z.shape[1], z.shape[0], re_, im_),
                      "to_string": ref.to_string(p), "save_hex": ref.save_matrix(p).hex()})
    synth.append({"label": "null", "matrix": None, "to_string": None,
                  "save_hex": ref.save_matrix(None).hex()})
    out = ROOT / "tests" / "golden" / "formats.json"
    out.write_text(json.dumps({"texts": texts, "synthetic": synth}, indent=1))
    print(f"wrote {out}: {len(texts)} texts, {len(synth)} synthetic")


if __name__ == "__main__":
    main()
