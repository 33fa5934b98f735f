"""Device forms of the matrix formats: oaph_matrix_parse / _to_string /
_save / _load copy a device matrix once and must agree exactly with the
host forms (pinned against the reference in tests/test_formats.py)."""
import numpy as np
import pytest

from oap_amd import formats
from oap_amd.api import DeviceMatrix, OapError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("planes", [(True, False), (True, True), (False, True)])
def test_device_round_trips(rng, planes):
    r, c = 6, 5
    re = np.round(rng.standard_normal((r, c)), 3) if planes[0] else None
    im = np.round(rng.standard_normal((r, c)), 3) if planes[1] else None
    if re is not None:
        re[2, :] = 0.0  # a run
    m = DeviceMatrix.from_numpy(re, im)
    s = formats.matrix_to_string(m)
    assert s == formats.to_string(c, r, re, im)
    data = formats.save(m)
    assert data == formats.save_host(c, r, re, im)
    back, used = formats.load_device(data)
    assert used == len(data)
    got_re, got_im = back.download()
    assert (got_re is None) == (re is None) and (got_im is None) == (im is None)
    if re is not None:
        assert np.array_equal(got_re, re)
    if im is not None:
        assert np.array_equal(got_im, im)
    p = formats.parse(f"(columns={c}, rows={r}) {s}")
    hc, hr, hre, him = formats.parse_host(f"(columns={c}, rows={r}) {s}")
    dre, dim = p.download()
    for a, b in ((dre, hre), (dim, him)):
        assert (a is None and b is None) or np.array_equal(a, b)


def test_null_and_errors():
    assert formats.matrix_to_string(None) == "nullptr"
    assert formats.save(None) == b"\x00"
    m, used = formats.load_device(b"\x00")
    assert m is None and used == 1
    with pytest.raises(OapError):
        formats.parse("nothing to parse")
