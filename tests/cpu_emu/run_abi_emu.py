"""Drive the CPU build of the library (tests/cpu_emu/_build/libcorr_cuda_emu.so,
AddressSanitizer) through the golden and ABI-contract sequences. Run as a child
process with libasan preloaded (see tests/test_cpu_emu.py)."""

import ctypes
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

corr = importlib.import_module("21cma_abf_correlator_amd.corr")
from abi_driver import ALL_VALID, HOST_ZEROED, MASK_MARKED, drive, valid_matrix  # noqa: E402
from oracle import oracle  # noqa: E402


def rel_rms(a, b):
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def golden(L, N, nf, n_cum, fps, rows):
    rng = np.random.default_rng(20260914)
    x = rng.integers(-4, 5, size=(rows, N, n_cum, 2 * nf)).astype(np.int8)
    x[0, 0, 0, 0] = -128
    x[0, 0, 0, 1] = 127
    cfg = corr.make_config(N, nf, fps, min(fps, 128), block_slots=3, rows_per_fetch=rows)
    got = drive(corr, cfg, x, rows, n_cum, ALL_VALID, lib=L)
    ref = np.stack([oracle.fx_f64(x[r], frames_div=n_cum) for r in range(rows)])
    e = rel_rms(got, ref)
    z = drive(corr, cfg, x, rows, n_cum, HOST_ZEROED, lib=L)
    m = drive(corr, cfg, x, rows, n_cum, MASK_MARKED, lib=L)
    v = valid_matrix(rows, N, n_cum)
    em = rel_rms(m, np.stack([oracle.fx_f64(x[r], frames_div=n_cum, valid=v[r]) for r in range(rows)]))
    same = z.tobytes() == m.tobytes()
    print(f"golden N={N} nf={nf}: rel RMS {e:.3e}, masked {em:.3e}, zeroed==masked {same}", flush=True)
    return e < 1e-4 and em < 1e-4 and same


def main():
    L = corr.declare(ctypes.CDLL(os.path.join(HERE, "_build", "libcorr_cuda_emu.so")))
    ok = golden(L, 4, 2048, 90, 32, 2)
    ok = golden(L, 5, 4096, 37, 16, 3) and ok
    print("OK" if ok else "FAILED")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
