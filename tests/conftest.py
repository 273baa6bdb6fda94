import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def pkg():
    """The product package (directory name contains a dash)."""
    return importlib.import_module("tchem-atm_amd")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def cb05_oracle(oracle_mod):
    return oracle_mod.Oracle(golden("cb05cl_ae5.yaml"))


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch
