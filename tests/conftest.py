import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


# torch (when installed) is imported before libmgfe.so is loaded, so that the
# process holds one HIP runtime whichever tests are selected: torch's ROCm
# build brings its own libamdhip64, which libmgfe.so then shares; loading it
# after libmgfe.so has started the system runtime leaves torch without devices.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """One libmgfe context on cuda:0 for the whole GPU session."""
    from morganapublic_amd import _lib
    c = _lib.Context(0)
    yield c
    c.close()
