import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_collection_modifyitems(config, items):
    # -m "not gpu" deselects GPU tests; without a GPU they would fail loudly
    # in the fixture below instead of silently passing.
    pass


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded through its C-ABI.  Fails loudly when the GPU
    or the built extension is missing."""
    from gb_amd import native
    return native.lib_checked_gpu()
