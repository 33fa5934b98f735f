import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle.pyoracle import Ref
    if not Ref.available():
        pytest.skip("oracle/_ref/libref.so not built (reference sources absent)")
    return Ref()


@pytest.fixture(scope="session")
def api():
    # GPU tests: the HIP library must load and run -- no fallback
    from oap_amd.api import CuProceduresApi
    return CuProceduresApi()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
