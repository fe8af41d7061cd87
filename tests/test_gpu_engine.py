"""Engine-level parity on device buffers (include/corrx.h).

F-engine: every supported FFT length against the FP64 oracle FFT, bound 1e-5
relative RMS per transform; masked frames come out exactly zero.
FX: the CPU-runnable configuration (4 inputs, n_freq = 1024, 64 spectra; the ABI
refuses n_freq % 2048 != 0, so it runs through the engine) against the FP64
oracle and the oracle's FP32 restatement.

Device memory comes from the HIP runtime the library links (package module
hip.py); no other GPU library is loaded in the process.
"""

import importlib

import numpy as np
import pytest

from conftest import fp32_spectra, rel_rms, to_tiles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(corr):
    return importlib.import_module("21cma_abf_correlator_amd.hip")


def _spectra(hip, eng, N, T, nf):
    """The engine's complex64 spectrum buffer on the host: (N, T, nf)."""
    return fp32_spectra(hip, eng, N, T, nf)


# powers of two, the multiples of 2048 with an odd factor (3, 5, 3, 7), and
# the staged lengths above 16384 (9 x 2048, 2 x 16384, 11 x 4096, 33 x 2048,
# 8 x 16384, 255 x 4096)
@pytest.mark.parametrize("nf", [512, 1024, 2048, 4096, 8192, 16384, 6144, 10240, 12288, 14336,
                                18432, 32768, 45056, 67584, 131072, 1044480])
def test_fengine_vs_fp64(corr, oracle, hip, nf):
    N, T = 3, 4
    L = 2 * nf
    print(f"fengine nf={nf} N={N} T={T}", flush=True)
    x = oracle.gen_int8(11, N, T, L, 0.3, sigma=16.0, max_delay=4, tone_channel=nf // 3,
                        tone_amp=8.0)
    x[0, 0, :8] = [127, -128, 127, -128, 0, 1, -1, 5]
    eng = corr.Engine(N, nf, T)
    d_in = hip.Buffer.from_array(x)
    mask = np.full((N, 1), 0b1111, dtype=np.uint8)
    mask[1, 0] = 0b1101  # station 1 frame 1 missing
    d_mask = hip.Buffer.from_array(mask)
    eng.fengine(d_in.ptr, T * L, T, d_mask.ptr, 1, 0)
    Y = _spectra(hip, eng, N, T, nf)
    valid = np.ones((N, T), dtype=np.uint8)
    valid[1, 1] = 0
    Y64 = oracle.spectra_f64(x, valid=valid)
    assert np.all(Y[1, 1] == 0)
    worst = max(rel_rms(Y[i, t], Y64[i, t]) for i in range(N) for t in range(T) if valid[i, t])
    print(f"  worst per-transform relative RMS = {worst:.3e}")
    assert worst < 1e-5
    eng.close()


def test_c1_fx_through_engine(corr, oracle, hip):
    # 4 inputs, n_freq = 1024 (L = 2048), 64 spectra, rho 0.3, delays in [0, 8]
    N, nf, T, chunk = 4, 1024, 64, 32
    L = 2 * nf
    B = N * (N + 1) // 2
    print(f"C1 FX N={N} nf={nf} T={T}", flush=True)
    x = oracle.gen_int8(1, N, T, L, 0.3, sigma=16.0, max_delay=8)
    eng = corr.Engine(N, nf, chunk)
    d_in = hip.Buffer.from_array(x)
    acc = hip.Buffer(B * nf * 8)
    acc.zero()
    for f0 in range(0, T, chunk):
        eng.fengine(d_in.ptr + f0 * L, T * L, chunk)
        eng.xengine(acc.ptr, chunk)
    corr.scale(acc.ptr, B * nf, T)
    hip.synchronize()
    got = acc.to_numpy(np.complex64, (B, nf))
    ref = oracle.fx_f64(x)
    err = rel_rms(got, ref)
    print(f"  rel RMS vs FP64 = {err:.3e}")
    assert err < 1e-5
    ref32, _ = oracle.fx_f32(x, chunk_frames=chunk)
    assert rel_rms(got, ref32) < 1e-5
    eng.close()


@pytest.mark.parametrize("nf", [2048, 16384, 6144, 32768])
def test_fp32_spectrum_orders_agree(corr, oracle, hip, nf):
    """The FP32 F-engine writes the same complex64 bins in every spectrum order
    the X-engine can read: the plain [station][frame][channel] order and tiles
    of 16 and 32 channels (forms::spec_index), byte for byte, with a missing
    frame; in-LDS, odd-factor and staged lengths."""
    N, T = 3, 5
    L = 2 * nf
    x = oracle.gen_int8(12, N, T, L, 0.3, sigma=16.0, max_delay=4)
    d_in = hip.Buffer.from_array(x)
    d_mask = hip.Buffer.from_array(np.array([[0xFF], [0xFD], [0xFF]], dtype=np.uint8))
    got = {}
    for tile in (nf, 16, 32):
        print(f"FP32 F-engine nf={nf} spectrum tile {tile}", flush=True)
        eng = corr.Engine(N, nf, T)
        eng.set_xtile(16 if tile == nf else tile, tile)
        eng.fengine(d_in.ptr, T * L, T, d_mask.ptr, 1, 0)
        got[tile] = fp32_spectra(hip, eng, N, T, nf)
        eng.close()
    assert np.all(got[nf][1, 1] == 0)
    for tile in (16, 32):
        assert np.array_equal(got[tile].view(np.uint32), got[nf].view(np.uint32)), tile
