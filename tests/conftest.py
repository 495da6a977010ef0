import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def pytest_runtest_setup(item):
    if "gpu" in item.keywords:
        print(f"\n[gpu-test] {item.nodeid}", flush=True)


@pytest.fixture(scope="session")
def golden_dir():
    return Path(__file__).resolve().parent / "golden"
