"""Parity on the benchmark set's own workloads, each against the FP64 oracle,
at reduced spectrum counts (the geometry, generator and precision form are the
workload's own; the integration is shorter so that the FP64 oracle finishes in
seconds). Model: the reference's golden harness (corr-cuda/tests/golden.c:41-48,
313-403), which drives the public ABI and compares with FP64.

  C2  24 inputs, n_freq 8192, FP32 form through the drop-in ABI, correlated
      Gaussian (sigma 16, rho 0.3, delays 0..32) plus a CW tone at channel 3000
      of amplitude 8 on every input: relative RMS < 1e-4, and channel 3000 is
      the peak of every autocorrelation (which also pins the oracle's tone).
  C4  40 inputs, n_freq 16384, rho 0.05, delays 0..64: the int8 form's
      integrator bit-exact against the oracle's X-engine on the engine's own
      spectra, the FP16 form within 1e-3 of FP64, and the tuner's 16-channel
      workgroups (32 channels hold at most 24 inputs).
  C5  the three sub-arrays (20 in nf 8192; 24 in nf 16384; 4 in nf 16384; rho
      0.3 / 0.1 / 0.5, distinct seeds), FP16 form, one corrx_batch_run per dump
      replayed as a hipGraph: every replay byte-equal to the engines run one
      after another, each sub-array within 1e-3 of FP64.
Plus the FP16 form's behaviour past the half range: gains calibrated on noise
alone, then a tone bin 61 dB above the noise's per-bin RMS (the form's headroom
is 48-54 dB); the dump stays finite and the other channels keep their accuracy.
"""

import importlib

import numpy as np
import pytest

from abi_driver import ALL_VALID, drive
from conftest import rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(corr):
    return importlib.import_module("21cma_abf_correlator_amd.hip")


def test_c2_dropin_cw_tone(corr, oracle):
    N, nf, T, fps = 24, 8192, 128, 64
    L = 2 * nf
    print(f"C2: drop-in ABI, {N} inputs, n_freq {nf}, {T} spectra, tone at channel 3000", flush=True)
    x = oracle.gen_int8(2002, N, T, L, 0.3, sigma=16.0, max_delay=32, tone_channel=3000, tone_amp=8.0)
    cfg = corr.make_config(N, nf, fps, fps, block_slots=3, rows_per_fetch=1)
    got = drive(corr, cfg, x[None], 1, T, ALL_VALID)[0]
    ref = oracle.fx_f64(x)
    err = rel_rms(got, ref)
    print(f"  relative RMS vs FP64 oracle {err:.3e}", flush=True)
    assert err < 1e-4
    pairs = corr.baseline_pairs(N)
    autos = [b for b, (j, k) in enumerate(pairs) if j == k]
    peaks = np.argmax(got[autos].real, axis=1)
    assert np.all(peaks == 3000), peaks
    # the tone's power in its bin: (A L/2)^2 per frame on top of the noise's
    # sigma^2 L (a cosine of amplitude A at an exact bin has |Y| = A L / 2)
    want = (8.0 * L / 2) ** 2 + 16.0 ** 2 * L
    assert np.all(np.abs(got[autos, 3000].real / want - 1.0) < 0.1)
    assert np.abs(ref[autos, 3000].real / want - 1.0).max() < 0.1


def _spectra(hip, eng, N, T, nf, form):
    """The engine's spectra after one fengine call of T frames, complex64 (N, T, nf)."""
    ptr, stride = eng.spectra()
    eb = 4 if form == 1 else 2
    raw = np.empty((N, stride * eb), dtype=np.uint8)
    hip.synchronize()
    hip.check(hip.rt().hipMemcpy(raw.ctypes.data, ptr, raw.nbytes, hip.D2H), "hipMemcpy")
    dt = np.float16 if form == 1 else np.int8
    tile = eng.xtile()[1]
    fcap = stride // nf
    h = raw.view(dt).astype(np.float32).reshape(N, nf // tile, fcap, tile, 2)[:, :, :T]
    h = h.swapaxes(1, 2).reshape(N, T, nf, 2)
    return (h[..., 0] + 1j * h[..., 1]).astype(np.complex64)


@pytest.fixture(scope="module")
def c4_input(oracle):
    N, nf, T = 40, 16384, 64
    x = oracle.gen_int8(4004, N, T, 2 * nf, 0.05, sigma=16.0, max_delay=64)
    return x, oracle.fx_f64(x)


def test_c4_int8_exact(corr, oracle, hip, c4_input):
    x, _ = c4_input
    N, T, L = x.shape
    nf = L // 2
    B = N * (N + 1) // 2
    print(f"C4: int8 form, {N} inputs, n_freq {nf}, {T} spectra", flush=True)
    eng = corr.Engine(N, nf, T, corr.FORM_INT8)
    assert eng.xtile()[0] == 16
    d_in = hip.Buffer.from_array(x)
    eng.calibrate(d_in.ptr, T * L, T)
    eng.fengine(d_in.ptr, T * L, T)
    Yq = _spectra(hip, eng, N, T, nf, corr.FORM_INT8)
    acc = hip.Buffer(B * nf * 8)
    acc.zero()
    eng.xengine(acc.ptr, T)
    hip.synchronize()
    raw = acc.to_numpy(np.int32, (B, nf, 2))
    eng.close()
    assert np.abs(Yq.real).max() <= 7 and np.abs(Yq.imag).max() <= 7
    exact = oracle.x_f64_from_spectra(Yq, frames_div=1)
    bad = int(np.count_nonzero(raw[..., 0] != exact.real) + np.count_nonzero(raw[..., 1] != exact.imag))
    print(f"  integrator values differing from the exact sum: {bad}", flush=True)
    assert bad == 0


def test_c4_fp16_vs_fp64(corr, oracle, hip, c4_input):
    x, ref = c4_input
    N, T, L = x.shape
    nf = L // 2
    B = N * (N + 1) // 2
    print(f"C4: FP16 form, {N} inputs, n_freq {nf}, {T} spectra", flush=True)
    eng = corr.Engine(N, nf, 32, corr.FORM_FP16)
    assert eng.xtile()[0] == 16
    d_in = hip.Buffer.from_array(x)
    eng.calibrate(d_in.ptr, T * L, T)
    acc = hip.Buffer(B * nf * eng.integrator_bytes())
    acc.zero()
    eng.integrate(d_in.ptr, T * L, T, acc.ptr)
    out = hip.Buffer(B * nf * 8)
    eng.dump(acc.ptr, T, out.ptr)
    hip.synchronize()
    V = out.to_numpy(np.complex64, (B, nf))
    eng.close()
    err = rel_rms(V, ref)
    print(f"  relative RMS vs FP64 {err:.3e}", flush=True)
    assert err < 1e-3


# (inputs, n_freq, spectra, rho, seed, frames per launch): the three C5
# sub-arrays with their integrations shortened 8x
C5_SUBARRAYS = [(20, 8192, 64, 0.3, 501, 32), (24, 16384, 256, 0.1, 502, 128), (4, 16384, 1024, 0.5, 503, 256)]


def test_c5_batch_graph(corr, oracle, hip):
    subs = []
    for N, nf, T, rho, seed, chunk in C5_SUBARRAYS:
        print(f"C5 sub-array {N} inputs, n_freq {nf}, {T} spectra", flush=True)
        L = 2 * nf
        B = N * (N + 1) // 2
        x = oracle.gen_int8(seed, N, T, L, rho, sigma=16.0, max_delay=32)
        eng = corr.Engine(N, nf, chunk, corr.FORM_FP16)
        d_in = hip.Buffer.from_array(x)
        eng.calibrate(d_in.ptr, T * L, min(T, 64))
        acc = hip.Buffer(B * nf * eng.integrator_bytes())
        out = hip.Buffer(B * nf * 8)
        subs.append(dict(N=N, nf=nf, T=T, L=L, B=B, x=x, eng=eng, d_in=d_in, acc=acc, out=out))
    hip.synchronize()

    def result(s):
        return s["out"].to_numpy(np.uint8, (s["B"] * s["nf"] * 8,)).copy()

    ref = []
    for s in subs:  # one after the other on the null stream
        s["acc"].zero()
        s["eng"].integrate(s["d_in"].ptr, s["T"] * s["L"], s["T"], s["acc"].ptr)
        s["eng"].dump(s["acc"].ptr, s["T"], s["out"].ptr)
        hip.synchronize()
        ref.append(result(s))
    batch = corr.Batch([s["eng"] for s in subs])
    stream = hip.Stream()
    items = [corr.Batch.item(s["d_in"].ptr, s["T"] * s["L"], s["T"], s["acc"].ptr, d_out=s["out"].ptr,
                             dump_frames=s["T"]) for s in subs]
    for rep in range(3):
        print(f"  batch graph replay {rep}", flush=True)
        for s in subs:
            s["out"].zero()
        batch.run(items, graph=True, stream=stream.handle)
        stream.synchronize()
        for s, r in zip(subs, ref):
            assert np.array_equal(result(s), r), f"replay {rep}: {s['N']} inputs differ from the sequential run"
    for s, r in zip(subs, ref):
        V = r.view(np.complex64).reshape(s["B"], s["nf"])
        err = rel_rms(V, oracle.fx_f64(s["x"]))
        print(f"  {s['N']} inputs n_freq {s['nf']}: relative RMS vs FP64 {err:.3e}", flush=True)
        assert err < 1e-3
    batch.close()
    stream.destroy()
    for s in subs:
        s["eng"].close()


@pytest.mark.parametrize("form", [1, 2])
def test_tone_past_the_calibrated_range(corr, oracle, hip, form):
    """Gains from noise alone (sigma 4: per-bin RMS 181, FP16 gain 1), then 64
    frames with a tone of amplitude 100 at channel 300: |Y| = 204800 there, past
    the largest half (65504). The FP16 form saturates that bin (the X-engine
    reads it as +-65504), the int8 form clamps it to +-7 by design; either way
    every visibility is finite and the channels away from the tone keep the
    form's accuracy."""
    N, nf, T = 4, 2048, 64
    L = 2 * nf
    B = N * (N + 1) // 2
    noise = oracle.gen_int8(77, N, T, L, 0.3, sigma=4.0, max_delay=4)
    tone = oracle.gen_int8(78, N, T, L, 0.3, sigma=4.0, max_delay=4, tone_channel=300, tone_amp=100.0)
    eng = corr.Engine(N, nf, 32, form)
    d_noise = hip.Buffer.from_array(noise)
    d_tone = hip.Buffer.from_array(tone)
    eng.calibrate(d_noise.ptr, T * L, T)
    acc = hip.Buffer(B * nf * eng.integrator_bytes())
    acc.zero()
    eng.integrate(d_tone.ptr, T * L, T, acc.ptr)
    out = hip.Buffer(B * nf * 8)
    eng.dump(acc.ptr, T, out.ptr)
    hip.synchronize()
    V = out.to_numpy(np.complex64, (B, nf))
    eng.close()
    assert np.all(np.isfinite(V.real)) and np.all(np.isfinite(V.imag))
    ref = oracle.fx_f64(tone)
    away = np.r_[0:290, 311:nf]
    e_away = rel_rms(V[:, away], ref[:, away])
    autos = [b for b, (j, k) in enumerate(corr.baseline_pairs(N)) if j == k]
    ratio = V[autos, 300].real / ref[autos, 300].real
    print(f"  form {form}: away from the tone {e_away:.3e}; tone bin dumped/true {ratio}", flush=True)
    assert e_away < (1e-3 if form == corr.FORM_FP16 else 0.1)
    assert np.all(ratio > 0.0) and np.all(ratio < 1.0)   # saturated below the true power
    if form == corr.FORM_FP16:
        # the real part (the tone) saturates at 65504 / g, g = 1; the imaginary
        # part is noise, far below: the bin's power is 65504^2 to a few percent
        assert np.allclose(V[autos, 300].real, 65504.0 ** 2, rtol=0.05)


def test_int8_dump_refuses_wrapping_lengths(corr, hip):
    """The int8 form's int32 integrator holds at most 2^31/98 frames; a dump
    of a longer nominal integration is refused rather than read wrapped."""
    eng = corr.Engine(2, 512, 8, corr.FORM_INT8)
    acc = hip.Buffer(3 * 512 * 8)
    acc.zero()
    out = hip.Buffer(3 * 512 * 8)
    eng.dump(acc.ptr, 2 ** 31 // 98, out.ptr)
    with pytest.raises(corr.CorrError):
        eng.dump(acc.ptr, 2 ** 31 // 98 + 1, out.ptr)
    hip.synchronize()
    eng.close()
