import ctypes
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def _load_pkg():
    spec = importlib.util.spec_from_file_location(
        "ipcc_amd", os.path.join(ROOT, "ipcc2021-final_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def w():
    """The ctypes binding of libwilson.so (fails loudly when it is not built)."""
    return _load_pkg().wilson_hip


@pytest.fixture(scope="session")
def oracle():
    from oracle_py import Oracle
    return Oracle()
