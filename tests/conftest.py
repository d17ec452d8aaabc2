import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


# Load the HIP library before any test module can import torch: torch ships
# its own libamdhip64.so.7, and once that is loaded the dynamic linker would
# satisfy the library's dependency with torch's runtime instead of ROCm's.
try:
    from sources_amd import abi as _abi
    if os.path.exists(_abi.LIB_PATH):
        _abi.load_library()
except OSError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def hiplib():
    """The HIP library; GPU tests fail loudly if it is not built."""
    from sources_amd import abi
    return abi.load_library()
