import os
import sys

import pytest

# PyTorch (used by the kernel-level tests for device memory) links its bundled HIP runtime
# by the unversioned name libamdhip64.so; the engine library asks for libamdhip64.so.7.
# Importing torch first lets the engine bind to torch's already-loaded runtime (same
# SONAME) rather than load a second HIP runtime into the process.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "vajra-opt_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def engine():
    import vajra_pdhg as vp
    e = vp.Engine(0)
    yield e
    e.close()
