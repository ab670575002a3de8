import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # Load what is built; build only when nothing is there.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "hspz_amd", "libhspz.so")):
        from hspz_amd import build
        build.build()


@pytest.fixture(scope="session")
def gpu_available():
    from hspz_amd import engine
    if engine.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return True
