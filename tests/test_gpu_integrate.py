"""corrx_integrate on the GPU: the whole-integration call (F-engine on the
caller's stream, X-engine on the engine's own stream over two alternating
spectrum buffers) adds exactly what corrx_fengine + corrx_xengine add over the
same runs of chunk_frames frames, in every form: the integrators are compared
byte for byte. Cases: ragged frame counts, fewer frames than one chunk, a
missing frame past the first chunk (mask bit offsets), a non-default stream with
the dump enqueued behind the call, and an integrator that already holds values.
"""

import importlib

import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(corr):
    return importlib.import_module("21cma_abf_correlator_amd.hip")


# (form, N, nf, T, chunk)
CASES = [(0, 6, 2048, 100, 32), (1, 6, 2048, 100, 32), (2, 6, 2048, 100, 32),
         (0, 3, 512, 5, 8), (1, 3, 512, 5, 8), (2, 40, 512, 77, 16),
         (1, 24, 16384, 80, 32), (2, 24, 16384, 80, 32), (0, 24, 16384, 48, 16),
         (1, 8, 10240, 70, 32), (0, 5, 12288, 40, 16), (1, 4, 20480, 40, 16), (0, 3, 32768, 70, 32)]


@pytest.mark.parametrize("form,N,nf,T,chunk", CASES)
def test_integrate_equals_chunk_loop(corr, oracle, hip, form, N, nf, T, chunk):
    print(f"integrate form {form} N={N} nf={nf} T={T} chunk={chunk}", flush=True)
    x = oracle.gen_int8(11 + form, N, T, 2 * nf, 0.2, sigma=16.0, max_delay=4)
    mask_pitch = (T + 7) // 8
    mask = np.full((N, mask_pitch), 0xFF, dtype=np.uint8)
    miss = min(T - 1, chunk + 2)              # a frame of the second chunk when there is one
    mask[N - 1, miss >> 3] &= ~np.uint8(1 << (miss & 7))
    valid = np.ones((N, T), dtype=np.uint8)
    valid[N - 1, miss] = 0
    eng = corr.Engine(N, nf, chunk, form)
    d_in = hip.Buffer.from_array(x)
    d_mask = hip.Buffer.from_array(mask)
    pitch = T * 2 * nf
    if form != corr.FORM_FP32:
        eng.calibrate(d_in.ptr, pitch, T, d_mask.ptr, mask_pitch, 0)
    B = N * (N + 1) // 2
    nbytes = B * nf * eng.integrator_bytes()
    loop, piped = hip.Buffer(nbytes), hip.Buffer(nbytes)
    loop.zero()
    piped.zero()
    # reference: the per-chunk loop on the default stream, twice over
    for _ in range(2):
        for f0 in range(0, T, chunk):
            n = min(chunk, T - f0)
            eng.fengine(d_in.ptr + f0 * 2 * nf, pitch, n, d_mask.ptr, mask_pitch, f0)
            eng.xengine(loop.ptr, n)
    hip.synchronize()
    # the pipeline on its own stream, twice over the same integrator, then the
    # dump enqueued behind it on that stream
    s = hip.Stream()
    eng.integrate(d_in.ptr, pitch, T, piped.ptr, d_mask.ptr, mask_pitch, 0, stream=s.handle)
    eng.integrate(d_in.ptr, pitch, T, piped.ptr, d_mask.ptr, mask_pitch, 0, stream=s.handle)
    eng.integrate(d_in.ptr, pitch, 0, piped.ptr, stream=s.handle)   # no frames: nothing
    out = hip.Buffer(B * nf * 8)
    eng.dump(piped.ptr, 2 * T, out.ptr, stream=s.handle)
    s.synchronize()
    a = loop.to_numpy(np.uint8, (nbytes,))
    b = piped.to_numpy(np.uint8, (nbytes,))
    same = np.array_equal(a, b)
    V = out.to_numpy(np.complex64, (B, nf))
    e64 = rel_rms(V, oracle.fx_f64(x, valid=valid))
    print(f"  integrator identical {same}; dump vs FP64 {e64:.3e}", flush=True)
    s.destroy()
    eng.close()
    assert same
    assert e64 < {0: 1e-5, 1: 1e-3, 2: 0.1}[form]


def test_integrate_rejections(corr, hip):
    eng = corr.Engine(2, 512, 8, corr.FORM_FP16)
    d_in = hip.Buffer(2 * 16 * 1024)
    acc = hip.Buffer(3 * 512 * eng.integrator_bytes())
    pitch = 16 * 1024
    with pytest.raises(corr.CorrError):
        eng.integrate(d_in.ptr, pitch, -1, acc.ptr)          # negative frame count
    with pytest.raises(corr.CorrError):
        eng.integrate(d_in.ptr, pitch, 17, acc.ptr)          # pitch holds 16 frames
    with pytest.raises(corr.CorrError):
        eng.integrate(d_in.ptr, pitch, 16, None)              # no integrator
    with pytest.raises(corr.CorrError):
        eng.integrate(None, pitch, 16, acc.ptr)               # no input
    with pytest.raises(corr.CorrError):
        eng.integrate(d_in.ptr, pitch, 16, acc.ptr, d_in.ptr, 1, 0)   # 8 mask bits for 16 frames
    eng.close()
