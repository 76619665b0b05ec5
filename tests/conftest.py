"""Shared fixtures: the product package (directory name with a hyphen) and the CPU oracle."""
import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def pk():
    return importlib.import_module("imagej-elphel_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle
