import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, GOLDEN)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the libm2 C ABI on the device)")


def load_fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated" % name)
    with open(path) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    """The oracle (test infrastructure only), built from oracle/ if needed."""
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    from oracle_ctypes import Oracle
    return Oracle(so)


@pytest.fixture(scope="session")
def gpu_ctx():
    import mersenne2gpu_amd as m2
    if m2.device_count() == 0:
        pytest.fail("no GPU visible to libm2 (gpu-marked tests need the device)")
    ctx = m2.Context(0)
    yield ctx
    ctx.close()
