import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "d-pdlp_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def dp():
    import dpdlp
    return dpdlp


@pytest.fixture(scope="session")
def orc():
    import oracle_py
    return oracle_py


@pytest.fixture(params=["auto", "renumbered"])
def row_order(request, monkeypatch):
    """Runs a single-LP test with the upload's automatic row order and with
    rows renumbered by decreasing length (forced for any shape)."""
    if request.param == "renumbered":
        monkeypatch.setenv("DPDLP_ROW_ORDER", "1")
    return request.param
