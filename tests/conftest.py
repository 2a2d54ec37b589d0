import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stormbitmaps-internal_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_storm_hip():
    spec = importlib.util.spec_from_file_location("storm_hip", os.path.join(PKG, "storm_hip.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def sh():
    return load_storm_hip()


@pytest.fixture(scope="session")
def ctx(sh):
    c = sh.Context(0)
    yield c
    c.close()
