import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "lattice-symmetries_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and liblsamd.so")


@pytest.fixture(scope="session")
def golden():
    with np.load(os.path.join(GOLDEN, "golden_ref_kernels.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def engine():
    """liblsamd.so with a usable device; fails (never skips) when either is
    missing, so a GPU run cannot pass on a fallback."""
    import lsamd

    lib = lsamd.load()
    if lib.ls_amd_device_available() != 0:
        raise RuntimeError("GPU engine unavailable: " + lsamd._abi.last_error())
    return lib
