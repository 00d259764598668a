import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "attribute-aware-rbfs_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the library through its C ABI")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    # the oracle is plain C and cheap to (re)build; the reference harness is
    # built only where the reference tree exists
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)
    yield


@pytest.fixture(scope="session")
def gpu_ctx():
    import aarbf
    ctx = aarbf.Context(0)
    yield ctx
    ctx.close()
