import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "n-body-code_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


# Load the backend's HIP runtime before anything imports torch: torch bundles
# its own libamdhip64.so.7 (same soname), and whichever loads first serves
# the process.  Tests use torch on CPU tensors only.
try:
    import nbody_amd as _nbody_amd
    _nbody_amd.lib()
except Exception:  # not built: the tests that need it fail loudly themselves
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def body_relerr(a, b):
    """Per-body |a - b| / |b| over 3-vector blocks; a, b are (3, n)."""
    num = np.sqrt(((np.asarray(a) - np.asarray(b)) ** 2).sum(0))
    den = np.sqrt((np.asarray(b) ** 2).sum(0))
    return num / den


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    """The compiled reference (oracle/_ref), when it has been built here."""
    from oracle import Reference
    try:
        return Reference("fp64"), Reference("f32")
    except RuntimeError as e:
        pytest.skip(str(e))


@pytest.fixture(scope="session")
def nb():
    import nbody_amd
    nbody_amd.lib()  # fails loudly when the HIP library is not built
    return nbody_amd
