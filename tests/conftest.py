import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "picoquant.jl_amd"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpq_hip on cuda:0)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def _announce(item):
    # Name the case before each GPU test so a crash can be attributed.
    if item.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu-test] {item.nodeid}\n")
        sys.stdout.flush()


def pytest_runtest_setup(item):
    _announce(item)
