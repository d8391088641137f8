import os
import sys

import pytest

# The oracle's OpenMP threads spin while they wait: several pytest-xdist
# workers, each with a full team on the same cores, slow its small parallel
# loops down by orders of magnitude.  Each worker gets one thread instead (set
# before the oracle library, and with it libgomp, is loaded).
if "PYTEST_XDIST_WORKER" in os.environ:
    os.environ.setdefault("OMP_NUM_THREADS", "1")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "stiff-gipc_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libsgipc.so kernels)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle
