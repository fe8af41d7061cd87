"""Engine-level parity on device buffers (include/corrx.h).

F-engine: every supported FFT length against the FP64 oracle FFT, bound 1e-5
relative RMS per transform; masked frames come out exactly zero.
FX: the CPU-runnable configuration (4 inputs, n_freq = 1024, 64 spectra; the ABI
refuses n_freq % 2048 != 0, so it runs through the engine) against the FP64
oracle and the committed fixture.
"""

import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _spectra(corr, eng, N, T, nf):
    """Copy the engine's complex64 spectrum buffer to the host: (N, T, nf)."""
    import ctypes
    ptr, stride = eng.spectra()
    out = np.empty((N, stride), dtype=np.complex64)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
    return out[:, : T * nf].reshape(N, T, nf)


@pytest.mark.parametrize("nf", [512, 1024, 2048, 4096, 8192, 16384])
def test_fengine_vs_fp64(corr, oracle, nf):
    N, T = 3, 4
    L = 2 * nf
    print(f"fengine nf={nf}", flush=True)
    x = oracle.gen_int8(11, N, T, L, 0.3, sigma=16.0, max_delay=4, tone_channel=nf // 3,
                        tone_amp=8.0)
    x[0, 0, :8] = [127, -128, 127, -128, 0, 1, -1, 5]
    eng = corr.Engine(N, nf, T)
    d_in = torch.from_numpy(x).cuda()
    mask = np.zeros((N, 1), dtype=np.uint8)
    mask[:, 0] = 0b1111
    mask[1, 0] = 0b1101  # station 1 frame 1 missing
    d_mask = torch.from_numpy(mask).cuda()
    eng.fengine(d_in.data_ptr(), T * L, T, d_mask.data_ptr(), 1, 0, _stream())
    Y = _spectra(corr, eng, N, T, nf)
    valid = np.ones((N, T), dtype=np.uint8)
    valid[1, 1] = 0
    Y64 = oracle.spectra_f64(x, valid=valid)
    assert np.all(Y[1, 1] == 0)
    worst = 0.0
    for i in range(N):
        for t in range(T):
            if not valid[i, t]:
                continue
            e = rel_rms(Y[i, t], Y64[i, t])
            worst = max(worst, e)
    print(f"  worst per-transform relative RMS = {worst:.3e}")
    assert worst < 1e-5
    eng.close()


def test_c1_fx_through_engine(corr, oracle):
    # 4 inputs, n_freq = 1024 (L = 2048), 64 spectra, rho 0.3, delays in [0, 8]
    N, nf, T = 4, 1024, 64
    L = 2 * nf
    x = oracle.gen_int8(1, N, T, L, 0.3, sigma=16.0, max_delay=8)
    eng = corr.Engine(N, nf, 32)
    B = N * (N + 1) // 2
    acc = torch.zeros((B, nf), dtype=torch.complex64, device="cuda")
    d_in = torch.from_numpy(x).cuda()
    for f0 in range(0, T, 32):
        ptr = d_in.data_ptr() + f0 * L
        eng.fengine(ptr, T * L, 32, stream=_stream())
        eng.xengine(acc.data_ptr(), 32, stream=_stream())
    corr.scale(acc.data_ptr(), B * nf, T, _stream())
    got = acc.cpu().numpy()
    ref = oracle.fx_f64(x)
    err = rel_rms(got, ref)
    print(f"C1 rel RMS vs FP64 = {err:.3e}")
    assert err < 1e-5
    ref32, _ = oracle.fx_f32(x, chunk_frames=32)
    assert rel_rms(got, ref32) < 1e-5
    eng.close()
