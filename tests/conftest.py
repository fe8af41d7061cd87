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


def to_tiles(a, fcap, tile):
    """(N, T, nf, ...) spectra -> the engine's spectrum order (N, nf/tile, fcap,
    tile, ...) (include/corrx.h corrx_spectra; tile = nf is the plain order),
    frames past T zero, flattened per station."""
    import numpy as np
    N, T, nf = a.shape[:3]
    out = np.zeros((N, nf // tile, fcap, tile) + a.shape[3:], dtype=a.dtype)
    out[:, :, :T] = a.reshape((N, T, nf // tile, tile) + a.shape[3:]).swapaxes(1, 2)
    return out.reshape(N, -1)


def from_tiles(a, nf, fcap, T, tile):
    """Inverse of to_tiles for the first T frames: (N, ...) -> (N, T, nf, ...)."""
    N = a.shape[0]
    t = a.reshape((N, nf // tile, fcap, tile) + a.shape[2:])[:, :, :T]
    return t.swapaxes(1, 2).reshape((N, T, nf) + a.shape[2:])


def fp32_spectra(hip, eng, N, T, nf):
    """An FP32-form engine's complex64 spectrum buffer on the host, (N, T, nf),
    in whichever order the engine writes (its spectra tile)."""
    import numpy as np
    ptr, stride = eng.spectra()
    raw = np.empty((N, stride), dtype=np.complex64)
    hip.synchronize()
    hip.check(hip.rt().hipMemcpy(raw.ctypes.data, ptr, raw.nbytes, hip.D2H), "hipMemcpy")
    return from_tiles(raw, nf, stride // nf, T, eng.xtile()[1])


def rel_rms(got, ref):
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    num = float(np.sum(np.abs(got.astype(np.complex128) - ref) ** 2))
    den = float(np.sum(np.abs(ref) ** 2))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return (num / den) ** 0.5
