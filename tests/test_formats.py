"""Matrix text and binary formats (host code of liboaph.so; no GPU needed).

Pinned by the reference's own format code: the expected results of
oap::host::NewMatrix(text), ToString and SaveMatrix for the matrix texts the
reference's tests hold and for small synthetic matrices are committed in
tests/golden/formats.json (written by tests/golden/make_format_fixtures.py
from oracle/_ref, the reference compiled from its sources).  Where oracle/_ref
is present, random matrices are also compared with it directly.  Bit-exact:
text byte for byte, parsed values and record bytes exactly.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from oap_amd import formats
from oap_amd.api import OapError
from oracle.pyoracle import Planes

FIX = json.loads((Path(__file__).resolve().parent / "golden" / "formats.json").read_text())


def same(got, want):
    """a plane against a fixture entry (values or digest), bit for bit"""
    if want is None:
        return got is None
    g = np.ascontiguousarray(got, dtype="<f8").reshape(-1)
    if isinstance(want, dict):
        return hashlib.sha256(g.tobytes()).hexdigest() == want["sha256"]
    w = np.array(want, dtype="<f8")
    return g.shape == w.shape and g.tobytes() == w.tobytes()


@pytest.mark.parametrize("case", FIX["texts"], ids=[t["source"] for t in FIX["texts"]])
def test_parse_reference_fixture_texts(case):
    exp = case["expected"]
    if exp is None:
        # the reference's parser yields no matrix for this text (schur3.h: rows
        # run together without a separator after "<repeats n times>")
        with pytest.raises(OapError):
            formats.parse_host(case["text"])
        return
    for which, want in ((1, case["arrays"][0]), (2, case["arrays"][1])):
        if want is None:
            with pytest.raises(OapError):
                formats.parse_array(case["text"], which)
        else:
            assert same(formats.parse_array(case["text"], which), want)
    if exp.get("mismatch"):
        # the reference adopts an array shorter than columns x rows
        # (matrix7.h: 32 x 100 declared, 1600 values given): rejected here
        with pytest.raises(OapError) as e:
            formats.parse_host(case["text"])
        assert e.value.status == 2
        return
    c, r, re, im = formats.parse_host(case["text"])
    assert (c, r) == (exp["columns"], exp["rows"])
    assert same(re, exp["re"]) and same(im, exp["im"])


@pytest.mark.parametrize("case", FIX["synthetic"], ids=[s["label"] for s in FIX["synthetic"]])
def test_print_and_save_match_reference(case):
    m = case["matrix"]
    if m is None:
        assert formats.save_host(0, 0, present=False).hex() == case["save_hex"]
        return
    c, r = m["columns"], m["rows"]
    re = None if m["re"] is None else np.array(m["re"]).reshape(r, c)
    im = None if m["im"] is None else np.array(m["im"]).reshape(r, c)
    assert formats.to_string(c, r, re, im) == case["to_string"]
    data = formats.save_host(c, r, re, im)
    assert data.hex() == case["save_hex"]
    (c2, r2, re2, im2), used = formats.load_host(data)
    assert used == len(data) and (c2, r2) == (c, r)
    assert same(re2, m["re"]) and same(im2, m["im"])


def test_record_layout():
    data = formats.save_host(3, 2, np.arange(6.0), None)
    assert data[:3] == bytes([1, 1, 0])
    assert int.from_bytes(data[3:7], "little") == 3 and int.from_bytes(data[7:11], "little") == 2
    assert np.frombuffer(data[11:], dtype="<f8").tolist() == list(range(6))
    assert formats.load_host(b"\x00") == (None, 1)
    with pytest.raises(OapError):
        formats.load_host(data[:-1])  # truncated record


def test_oversized_header_rejected():
    # columns = rows = 2^31: 2^62 elements, whose byte count wraps a 64-bit size
    # to 11; the header must be refused before any size is formed from it
    head = bytes([1, 1, 1]) + (1 << 31).to_bytes(4, "little") * 2
    for tail in (b"", bytes(64)):
        with pytest.raises(OapError):
            formats.load_host(head + tail)
    # 2^32 elements exactly is the first size refused; one less passes the
    # header check and fails only for want of data
    head = bytes([1, 1, 0]) + (1 << 16).to_bytes(4, "little") * 2
    with pytest.raises(OapError):
        formats.load_host(head)


def test_records_concatenate(tmp_path):
    a = formats.save_host(2, 2, np.eye(2), np.ones((2, 2)))
    b = formats.save_host(0, 0, present=False)
    c = formats.save_host(1, 3, np.array([1.0, -2.0, 3.0]))
    blob = a + b + c
    off = 0
    out = []
    while off < len(blob):
        m, used = formats.load_host(blob, off)
        out.append(m)
        off += used
    assert out[1] is None and out[0][0:2] == (2, 2) and out[2][0:2] == (1, 3)
    p = tmp_path / "m.bin"
    formats.write_file(p, blob)
    raw = p.read_bytes()
    assert int.from_bytes(raw[:8], "little") == 1
    assert int.from_bytes(raw[8:16], "little") == len(blob) and raw[16:] == blob
    assert formats.read_file(p) == blob


def test_parser_semantics():
    # dims given explicitly; '|' and ',' separate; runs expand; '-' negates
    c, r, re, im = formats.parse_host("(columns=3, rows=2) [1, -2, 3 | 4, 5 <repeats 2 times>]")
    assert (c, r) == (3, 2) and re.tolist() == [[1, -2, 3], [4, 5, 5]] and im is None
    # no dims: square of side floor(sqrt(n)); a second array is the im plane
    c, r, re, im = formats.parse_host("[1,2,3,4] [0, 1, 0, -1]")
    assert (c, r) == (2, 2) and im.tolist() == [[0, 1], [0, -1]]
    assert formats.parse_array("x [1.5e-3, 2] y [7]", 2).tolist() == [7.0]
    assert formats.parse_array("[1.5e-3, 2]", 1).tolist() == [1.5e-3, 2.0]
    for bad in ("no array here", "[1, 2, a]", "[1, 2 <repeats 3 times> 4]"):
        with pytest.raises(OapError):
            formats.parse_host(bad)
    with pytest.raises(OapError):
        formats.parse_host("(columns=3, rows=3) [1, 2, 3]")  # shorter than columns x rows
    # values past columns x rows are not read (a trailing ',' adds one)
    assert formats.parse_host("[1, 2, 3, 4, 5,]")[2].tolist() == [[1, 2], [3, 4]]


def test_text_round_trip():
    rng = np.random.default_rng(3)
    re = rng.integers(-8, 8, (5, 7)) / 4.0  # values that print exactly with 9 decimals
    im = rng.integers(-8, 8, (5, 7)) / 8.0
    s = formats.to_string(7, 5, re, im)
    c, r, re2, im2 = formats.parse_host(f"(columns=7, rows=5) {s}")
    assert (c, r) == (7, 5) and np.array_equal(re2, re) and np.array_equal(im2, im)


@pytest.mark.parametrize("seed", range(6))
def test_against_reference_build(ref, seed, tmp_path):
    rng = np.random.default_rng(seed)
    r, c = int(rng.integers(1, 9)), int(rng.integers(1, 9))
    kind = seed % 3
    vals = lambda: np.where(rng.random((r, c)) < 0.3, 1.0,  # noqa: E731
                            np.round(rng.standard_normal((r, c)) * 10.0 ** rng.integers(-3, 4), 6))
    re = vals() if kind != 2 else None
    im = vals() if kind != 0 else None
    p = Planes(r, c, re is not None, im is not None)
    if re is not None:
        p.re[:] = re
    if im is not None:
        p.im[:] = im
    s = formats.to_string(c, r, re, im)
    assert s == ref.to_string(p)
    data = formats.save_host(c, r, re, im)
    assert data == ref.save_matrix(p)
    got = ref.load_matrix(data)
    assert got[0:2] == (c, r)
    if re is not None:
        assert np.array_equal(got[2], re)
    text = f"(columns={c}, rows={r}) {s}"
    want = ref.parse_matrix(text)
    mine = formats.parse_host(text)
    assert mine[0:2] == want[0:2]
    for a, b in zip(mine[2:], want[2:]):
        assert (a is None and b is None) or np.array_equal(a, b)
    path = tmp_path / "r.bin"
    ref.buffer_fwrite(path, data)
    assert formats.read_file(path) == data
