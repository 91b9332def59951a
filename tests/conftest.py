import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcsx.so on cuda:0)")


@pytest.fixture(scope="session")
def O():
    import oracle

    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build()
    return oracle.Oracle()


@pytest.fixture(scope="session")
def R():
    import oracle

    if not oracle.reference_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return oracle.Reference()


@pytest.fixture(scope="session")
def golden():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json")) as f:
        return json.load(f)["cases"]


def unhex(s, dtype=np.uint8):
    return np.frombuffer(bytes.fromhex(s), dtype=dtype)
