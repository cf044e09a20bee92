import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "lulesh-task-based-hpx_amd")
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def load_pylulesh():
    spec = importlib.util.spec_from_file_location("pylulesh", os.path.join(PKG_DIR, "pylulesh.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pylulesh"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pyl():
    return load_pylulesh()
