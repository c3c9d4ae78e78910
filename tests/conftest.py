import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    # torch (device tensors in some GPU tests) initialises the HIP runtime
    # before the render core's library does, so both share one runtime that
    # sees the GPU, whichever test file runs first
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return ge.load_package()


@pytest.fixture(scope="session")
def wl():
    return ge.load_workloads()


@pytest.fixture(scope="session")
def ref():
    return ge.load_oracle()


@pytest.fixture(scope="session")
def golden():
    import json
    return json.load(open(ROOT / "tests" / "golden" / "reference_golden.json"))


@pytest.fixture(scope="session")
def ior_tables():
    import json
    return json.load(open(ROOT / "tests" / "golden" / "ior_tables.json"))
