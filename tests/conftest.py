import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests")


@pytest.fixture(scope="session")
def built():
    """Builds the oracle and the package libraries once if they are missing."""
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    for lib in ("libkfgen.so", "libkfmi.so", "kf_run"):
        if not os.path.exists(os.path.join(ROOT, "gpukalmanfitter_amd", lib)):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gpukalmanfitter_amd", "csrc")],
                           check=True)
    return True
