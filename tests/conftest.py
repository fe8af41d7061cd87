import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def corr():
    """The package's C-ABI binding (raises if the library is not built)."""
    pkg = importlib.import_module("21cma_abf_correlator_amd")
    pkg.corr.load()
    return pkg.corr


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def rel_rms(got, ref):
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    num = float(np.sum(np.abs(got.astype(np.complex128) - ref) ** 2))
    den = float(np.sum(np.abs(ref) ** 2))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return (num / den) ** 0.5
