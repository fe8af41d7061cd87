"""Precision forms on the GPU (include/corrx.h, CORRX_FORM_FP16 / CORRX_FORM_INT8).

MFMA X-engine alone, on spectra written straight into the engine's spectrum
buffer: int8 bit-exact against an int64 sum of the same products, FP16 within
FP32 rounding of an FP64 sum; the integrator's start values are kept (+=), for
1..48 inputs (every row-tile count the kernel has), ragged frame counts, both
channel-tile mappings and both workgroup widths (16 and 32 channels; the tuner's
choice and each width forced).

Whole form pipeline (calibrate, F-engine per chunk with a missing frame,
X-engine, FP64 dump) against the FP64 oracle: on the engine's own requantized
spectra within 1e-6 (int8 integrator exact), and against unquantized FP64
visibilities, FP16 within the 1e-3 relative RMS the precision form is held to.
"""

import importlib

import numpy as np
import pytest

from conftest import fp32_spectra, from_tiles, rel_rms, to_tiles

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(corr):
    return importlib.import_module("21cma_abf_correlator_amd.hip")


def _upload_spectra(hip, eng, host_words):
    """host_words: (N, stride) array of the form's element (uint32 / uint16)."""
    ptr, stride = eng.spectra()
    assert host_words.shape[1] == stride
    a = np.ascontiguousarray(host_words)
    hip.check(hip.rt().hipMemcpy(ptr, a.ctypes.data, a.nbytes, hip.H2D), "hipMemcpy")


def _to_tiles(a, fcap, tile):
    """(N, T, nf, ...) spectra -> the engine's tile-major order
    (N, nf/tile, fcap, tile, ...), frames past T zero, flattened per station."""
    N, T, nf = a.shape[:3]
    out = np.zeros((N, nf // tile, fcap, tile) + a.shape[3:], dtype=a.dtype)
    out[:, :, :T] = a.reshape((N, T, nf // tile, tile) + a.shape[3:]).swapaxes(1, 2)
    return out.reshape(N, -1)


def _from_tiles(a, nf, fcap, T, tile):
    """Inverse of _to_tiles for the first T frames: (N, ...) -> (N, T, nf, ...)."""
    N = a.shape[0]
    t = a.reshape((N, nf // tile, fcap, tile) + a.shape[2:])[:, :, :T]
    return t.swapaxes(1, 2).reshape((N, T, nf) + a.shape[2:])


def _download_spectra(hip, eng, N, T, nf, form):
    ptr, stride = eng.spectra()
    eb = 4 if form == 1 else 2
    raw = np.empty((N, stride * eb), dtype=np.uint8)
    hip.synchronize()
    hip.check(hip.rt().hipMemcpy(raw.ctypes.data, ptr, raw.nbytes, hip.D2H), "hipMemcpy")
    dt = np.float16 if form == 1 else np.int8
    h = _from_tiles(raw.view(dt).astype(np.float32).reshape(N, stride, 2), nf, stride // nf, T, eng.xtile()[1])
    return (h[..., 0] + 1j * h[..., 1]).astype(np.complex64)


def _x_int64(q, N):
    """q: (N, T, nf, 2) int; integer visibilities (B, nf, 2) in j<=k order."""
    out = []
    for j in range(N):
        for k in range(j, N):
            xr, xi = q[j, ..., 0].astype(np.int64), q[j, ..., 1].astype(np.int64)
            yr, yi = q[k, ..., 0].astype(np.int64), q[k, ..., 1].astype(np.int64)
            out.append(np.stack([(xr * yr + xi * yi).sum(0), (xi * yr - xr * yi).sum(0)], -1))
    return np.stack(out)


# (N, nf, T, channels per workgroup, spectrum tile): 0 = the tuner's choice,
# tile -1 = n_freq (the plain order)
INT8_CASES = [(1, 512, 3, 0, 0), (4, 512, 70, 0, 0), (9, 1024, 130, 0, 0), (24, 512, 64, 0, 0),
              (24, 16384, 37, 0, 0), (30, 512, 20, 0, 0), (40, 512, 65, 0, 0), (48, 512, 7, 0, 0),
              (1, 512, 3, 32, 32), (5, 1024, 70, 32, -1), (16, 512, 64, 32, 32), (24, 16384, 37, 32, 32),
              (24, 16384, 37, 16, -1), (24, 512, 130, 16, 16), (40, 1024, 65, 16, -1),
              (8, 6144, 33, 32, -1), (24, 14336, 20, 16, 16), (40, 10240, 9, 16, -1)]
FP16_CASES = [(1, 512, 1, 0, 0), (5, 512, 33, 0, 0), (8, 1024, 64, 0, 0), (24, 512, 96, 0, 0),
              (24, 16384, 40, 0, 0), (40, 512, 31, 0, 0), (48, 512, 50, 0, 0),
              (3, 512, 33, 32, 32), (24, 16384, 40, 32, 32), (24, 16384, 40, 32, -1), (24, 512, 96, 16, 16),
              (48, 512, 50, 16, -1), (8, 6144, 33, 32, 32), (24, 12288, 20, 32, -1), (40, 14336, 9, 16, 16)]


def _engine(corr, N, nf, T, form, ch, tile, step=0):
    eng = corr.Engine(N, nf, T, form)
    c, t = eng.xtile()
    assert c in ((16, 32) if N <= 24 else (16,)) and t in (c, nf)
    fb = {corr.FORM_FP32: 16, corr.FORM_FP16: 32, corr.FORM_INT8: 64}[form]
    assert eng.xstep() in (fb, 2 * fb, 4 * fb)
    if ch:
        eng.set_xtile(ch, nf if tile < 0 else tile)
        assert eng.xtile() == (ch, nf if tile < 0 else tile)
    if step:
        eng.set_xstep(step)
        assert eng.xstep() == step
    return eng


@pytest.mark.parametrize("form", [0, 1])
def test_xtile_rejections(corr, hip, form):
    eng = corr.Engine(40, 512, 8, form)
    assert eng.xtile()[0] == 16
    with pytest.raises(corr.CorrError):
        eng.set_xtile(32, 32)  # 32 channels hold at most 24 inputs
    with pytest.raises(corr.CorrError):
        eng.set_xtile(24, 24)
    with pytest.raises(corr.CorrError):
        eng.set_xtile(16, 32)  # the spectrum tile is the width or n_freq
    eng.close()


@pytest.mark.parametrize("N,nf,T,ch,tile", INT8_CASES)
def test_xengine_int8_exact(corr, hip, N, nf, T, ch, tile):
    _xengine_int8(corr, hip, N, nf, T, ch, tile)


def _xengine_int8(corr, hip, N, nf, T, ch, tile, step=0):
    print(f"int8 MFMA X-engine N={N} nf={nf} T={T} ch={ch} tile={tile} step={step}", flush=True)
    rng = np.random.default_rng(N * 1000 + T)
    eng = _engine(corr, N, nf, T, corr.FORM_INT8, ch, tile, step)
    _, stride = eng.spectra()
    q = rng.integers(-7, 8, size=(N, T, nf, 2)).astype(np.int8)
    q[0, 0, : min(nf, 4)] = [[7, -7], [-7, 7], [7, 7], [-7, -7]][: min(nf, 4)]
    words = _to_tiles(q.view(np.uint16).reshape(N, T, nf), stride // nf, eng.xtile()[1])
    _upload_spectra(hip, eng, words)
    B = N * (N + 1) // 2
    start = rng.integers(-1000, 1000, size=(B, nf, 2)).astype(np.int32)
    acc = hip.Buffer.from_array(start)
    eng.xengine(acc.ptr, T)
    hip.synchronize()
    got = acc.to_numpy(np.int32, (B, nf, 2)).astype(np.int64)
    want = _x_int64(q, N) + start
    bad = np.argwhere(got != want)
    if len(bad):
        b, c, p = bad[0]
        print(f"  first mismatch baseline {b} channel {c} part {p}: {got[b, c, p]} != {want[b, c, p]}"
              f" ({len(bad)} of {got.size})")
    assert len(bad) == 0
    eng.close()


@pytest.mark.parametrize("N,nf,T,ch,tile", FP16_CASES)
def test_xengine_fp16(corr, hip, N, nf, T, ch, tile):
    _xengine_fp16(corr, hip, N, nf, T, ch, tile)


def _xengine_fp16(corr, hip, N, nf, T, ch, tile, step=0):
    print(f"FP16 MFMA X-engine N={N} nf={nf} T={T} ch={ch} tile={tile} step={step}", flush=True)
    rng = np.random.default_rng(N * 77 + T)
    eng = _engine(corr, N, nf, T, corr.FORM_FP16, ch, tile, step)
    _, stride = eng.spectra()
    h = rng.normal(0.0, 200.0, size=(N, T, nf, 2)).astype(np.float16)
    words = _to_tiles(h.view(np.uint32).reshape(N, T, nf), stride // nf, eng.xtile()[1])
    _upload_spectra(hip, eng, words)
    B = N * (N + 1) // 2
    start = np.zeros((B, nf, 4), dtype=np.float32)
    start[..., 0], start[..., 1], start[..., 2], start[..., 3] = 1000.0, -500.0, 0.25, -0.125
    acc = hip.Buffer.from_array(start)
    eng.xengine(acc.ptr, T)
    hip.synchronize()
    got = acc.to_numpy(np.float32, (B, nf, 4)).astype(np.float64)
    hd = h.astype(np.float64)
    Y = hd[..., 0] + 1j * hd[..., 1]
    worst = 0.0
    b = 0
    for j in range(N):
        for k in range(j, N):
            v = (Y[j] * np.conj(Y[k])).sum(0) + (1000.25 - 500.125j)
            mag = (np.abs(Y[j]) * np.abs(Y[k])).sum(0) + 1.0
            g = (got[b, :, 0] + got[b, :, 2]) + 1j * (got[b, :, 1] + got[b, :, 3])
            worst = max(worst, float(np.max(np.abs(g - v) / mag)))
            b += 1
    print(f"  worst |error| / sum |products| = {worst:.3e}")
    assert worst < 2e-6
    eng.close()


# (N, nf, T, channels per workgroup, spectrum tile): 0 = the tuner's choice,
# tile -1 = n_freq (the plain order). Every row-tile count of the matrix-core
# FP32 X-engine, both widths, both spectrum orders, ragged frame counts, and 49
# inputs, where the FP32 form keeps the vector X-engine
FP32_CASES = [(1, 512, 3, 0, 0), (5, 512, 37, 0, 0), (8, 1024, 64, 0, 0), (13, 512, 17, 32, -1),
              (24, 512, 100, 0, 0), (24, 16384, 40, 32, -1), (24, 16384, 40, 32, 32), (24, 16384, 40, 16, 16),
              (30, 512, 33, 16, -1), (40, 512, 31, 0, 0), (44, 2048, 20, 16, 16), (48, 512, 50, 0, 0),
              (49, 512, 9, 0, 0), (8, 6144, 33, 32, 32), (24, 2048, 37, -1, -1)]


@pytest.mark.parametrize("N,nf,T,ch,tile", FP32_CASES)
def test_xengine_fp32(corr, hip, N, nf, T, ch, tile):
    """The FP32 form's X-engine alone on complex64 spectra written straight into
    the engine's spectrum buffer, accumulated onto start values, against an
    FP64 sum: every product within FP32 rounding of the sum of |products|."""
    _xengine_fp32(corr, hip, N, nf, T, ch, tile)


def _xengine_fp32(corr, hip, N, nf, T, ch, tile, step=0):
    print(f"FP32 X-engine N={N} nf={nf} T={T} ch={ch} tile={tile} step={step}", flush=True)
    rng = np.random.default_rng(N * 91 + T)
    if N > 48:
        eng = corr.Engine(N, nf, T, corr.FORM_FP32)
        assert eng.xtile() == (0, nf), "more than 48 inputs take the vector X-engine"
        with pytest.raises(corr.CorrError):
            eng.set_xtile(16, nf)
    elif ch < 0:  # the vector X-engine chosen
        eng = corr.Engine(N, nf, T, corr.FORM_FP32)
        eng.set_xtile(0, nf)
        assert eng.xtile() == (0, nf)
    else:
        eng = _engine(corr, N, nf, T, corr.FORM_FP32, ch, tile, step)
    ptr, stride = eng.spectra()
    Y = (rng.normal(0.0, 3000.0, size=(N, T, nf)) + 1j * rng.normal(0.0, 3000.0, size=(N, T, nf)))
    Y = Y.astype(np.complex64)
    y = np.ascontiguousarray(to_tiles(Y, stride // nf, eng.xtile()[1]))
    hip.check(hip.rt().hipMemcpy(ptr, y.ctypes.data, y.nbytes, hip.H2D), "hipMemcpy")
    B = N * (N + 1) // 2
    start = np.zeros((B, nf), dtype=np.complex64) + (1000.0 - 500.0j)
    acc = hip.Buffer.from_array(start)
    eng.xengine(acc.ptr, T)
    hip.synchronize()
    got = acc.to_numpy(np.complex64, (B, nf)).astype(np.complex128)
    eng.close()
    Yd = Y.astype(np.complex128)
    A = np.abs(Yd)
    j, k = np.triu_indices(N)
    worst = 0.0
    for b0 in range(0, B, 64):
        jj, kk = j[b0:b0 + 64], k[b0:b0 + 64]
        v = np.einsum("btc,btc->bc", Yd[jj], np.conj(Yd[kk])) + (1000.0 - 500.0j)
        mag = np.einsum("btc,btc->bc", A[jj], A[kk]) + 1.0
        worst = max(worst, float(np.max(np.abs(got[b0:b0 + 64] - v) / mag)))
    print(f"  worst |error| / sum |products| = {worst:.3e}")
    assert worst < 2e-6


# (form, N, nf, T, channels per workgroup, spectrum tile, frames per staged
# step): two and four K extents of the form's MFMA staged at once (form 0 FP32
# 16 frames per extent, 1 FP16 32, 2 int8 64), at the largest input counts
# they admit (inputs x extents <= 64), both widths and both spectrum orders,
# with frame counts below one step and ragged past it
STEP_CASES = [(f, *c[:-1], c[-1] * fb) for f, fb in ((0, 16), (1, 32), (2, 64)) for c in (
    (4, 512, 0, 32, 32, 4), (8, 1024, 0, 32, -1, 2), (8, 512, 0, 32, 32, 4), (16, 1024, 0, 16, -1, 4),
    (16, 512, 0, 32, 32, 2), (24, 512, 0, 16, 16, 2), (32, 2048, 0, 16, -1, 2), (13, 6144, 0, 16, 16, 4))]
STEP_CASES = [(f, N, nf, {0: 37, 1: 70, 2: 130}[f] + 3 * i, ch, tile, step)
              for i, (f, N, nf, _, ch, tile, step) in enumerate(STEP_CASES)]
STEP_CASES += [(0, 3, 512, 5, 16, 16, 64), (1, 3, 512, 7, 16, -1, 128), (2, 3, 512, 9, 16, 16, 256)]


@pytest.mark.parametrize("form,N,nf,T,ch,tile,step", STEP_CASES)
def test_xengine_time_k(corr, hip, form, N, nf, T, ch, tile, step):
    """Each forced step depth against the same checks as the one-extent step:
    exact integers, FP16 and FP32 products within FP32 rounding."""
    {0: _xengine_fp32, 1: _xengine_fp16, 2: _xengine_int8}[form](corr, hip, N, nf, T, ch, tile, step)


@pytest.mark.parametrize("form", [0, 1, 2])
def test_xstep_rejections(corr, hip, form):
    fb = (16, 32, 64)[form]
    eng = corr.Engine(40, 512, 8, form)
    for bad in (0, -fb, fb // 2, 3 * fb, 8 * fb):
        with pytest.raises(corr.CorrError):
            eng.set_xstep(bad)
    with pytest.raises(corr.CorrError):
        eng.set_xstep(2 * fb)  # 40 inputs x 2 extents pass 64
    eng.set_xstep(fb)
    eng.close()
    eng = corr.Engine(16, 512, 8, form)
    eng.set_xtile(16, 16)
    eng.set_xstep(4 * fb)
    eng.set_xtile(32, 32)  # four extents at 32 channels pass the LDS: back to one
    assert eng.xtile() == (32, 32) and eng.xstep() == fb
    with pytest.raises(corr.CorrError):
        eng.set_xstep(4 * fb)
    eng.set_xstep(2 * fb)
    assert eng.xstep() == 2 * fb
    eng.close()
    if form == 0:
        eng = corr.Engine(8, 512, 8, form)
        eng.set_xtile(0, 512)  # the vector X-engine has no step depth
        assert eng.xstep() == 0
        with pytest.raises(corr.CorrError):
            eng.set_xstep(fb)
        eng.close()


def _pipeline(corr, oracle, hip, form, N, nf, T, chunk, cal_frames, rho, seed, max_delay):
    x = oracle.gen_int8(seed, N, T, 2 * nf, rho, sigma=16.0, max_delay=max_delay)
    mask_pitch = (T + 7) // 8
    mask = np.full((N, mask_pitch), 0xFF, dtype=np.uint8)
    mask[1, 0] &= ~np.uint8(1 << 3)  # station 1, frame 3 missing
    valid = np.ones((N, T), dtype=np.uint8)
    valid[1, 3] = 0
    eng = corr.Engine(N, nf, chunk, form)
    d_in = hip.Buffer.from_array(x)
    d_mask = hip.Buffer.from_array(mask)
    pitch = T * 2 * nf
    eng.calibrate(d_in.ptr, pitch, cal_frames, d_mask.ptr, mask_pitch, 0)
    hip.synchronize()
    gains = np.empty((N, nf), dtype=np.float32)
    hip.check(hip.rt().hipMemcpy(gains.ctypes.data, eng.gains(), gains.nbytes, hip.D2H), "hipMemcpy")
    B = N * (N + 1) // 2
    acc = hip.Buffer(B * nf * eng.integrator_bytes())
    acc.zero()
    Yq = np.zeros((N, T, nf), dtype=np.complex64)
    for f0 in range(0, T, chunk):
        n = min(chunk, T - f0)
        eng.fengine(d_in.ptr + f0 * 2 * nf, pitch, n, d_mask.ptr, mask_pitch, f0)
        Yq[:, f0:f0 + n] = _download_spectra(hip, eng, N, n, nf, form)
        eng.xengine(acc.ptr, n)
    out = hip.Buffer(B * nf * 8)
    eng.dump(acc.ptr, T, out.ptr)
    hip.synchronize()
    V = out.to_numpy(np.complex64, (B, nf))
    if form == corr.FORM_INT8:
        raw = acc.to_numpy(np.int32, (B, nf, 2))
    else:
        raw = acc.to_numpy(np.float32, (B, nf, 4))
    eng.close()
    return x, valid, gains, Yq, V, raw


@pytest.mark.parametrize("form,nf", [(1, 2048), (2, 2048), (1, 6144), (2, 10240), (1, 14336), (2, 12288),
                                     (1, 20480), (2, 32768)])
def test_form_pipeline(corr, oracle, hip, form, nf):
    N, T, chunk = 6, 100, 32
    print(f"form {form} pipeline N={N} nf={nf} T={T} chunk={chunk}", flush=True)
    x, valid, gains, Yq, V, raw = _pipeline(corr, oracle, hip, form, N, nf, T, chunk, 60, 0.3, 5, 8)
    assert np.all(Yq[1, 3] == 0)
    assert np.all(gains > 0) and np.all(np.isfinite(gains))
    ref_q = oracle.x_f64_from_spectra(Yq, gain=gains, frames_div=T)
    eq = rel_rms(V, ref_q)
    ref = oracle.fx_f64(x, valid=valid)
    e64 = rel_rms(V, ref)
    print(f"  vs oracle on the engine's spectra {eq:.3e}; vs unquantized FP64 {e64:.3e}")
    assert eq < 1e-6
    if form == corr.FORM_INT8:
        exact = oracle.x_f64_from_spectra(Yq, frames_div=1)
        assert np.array_equal(raw[..., 0], exact.real) and np.array_equal(raw[..., 1], exact.imag)
        assert e64 < 0.1
    else:
        assert e64 < 1e-3


@pytest.mark.parametrize("form", [1, 2])
def test_calibrated_gains(corr, oracle, hip, form):
    """Gains against the rule applied to the oracle's FP64 spectra of the same
    frames: FP16 gains are powers of two (equal, or one octave off where the
    RMS sits on a power-of-two boundary), int8 gains within 1e-4."""
    N, nf, T = 4, 1024, 40
    x = oracle.gen_int8(9, N, T, 2 * nf, 0.2, sigma=16.0, max_delay=4, tone_channel=300, tone_amp=8.0)
    eng = corr.Engine(N, nf, 16, form)
    d_in = hip.Buffer.from_array(x)
    eng.calibrate(d_in.ptr, T * 2 * nf, T)
    hip.synchronize()
    g = np.empty((N, nf), dtype=np.float32)
    hip.check(hip.rt().hipMemcpy(g.ctypes.data, eng.gains(), g.nbytes, hip.D2H), "hipMemcpy")
    eng.close()
    Y = oracle.spectra_f64(x)
    sigma = np.sqrt((np.abs(Y) ** 2).mean(axis=1) / 2.0)
    if form == corr.FORM_FP16:
        m, e = np.frexp(sigma)
        want = np.ldexp(1.0, 8 - e)
        ratio = g / want
        assert np.all(np.isin(ratio, [0.5, 1.0, 2.0]))
        assert np.mean(ratio == 1.0) > 0.99
        assert np.all(np.log2(g) == np.round(np.log2(g)))
    else:
        want = 1.0 / (0.35 * sigma)
        assert np.max(np.abs(g / want - 1.0)) < 1e-4


def test_c3_shape_fp16_error(corr, oracle, hip):
    """The precision-form geometry (24 inputs, n_freq 16384, rho 0.1, delays in
    [0, 32]) on 64 spectra: FP16 MFMA form within 1e-3 of FP64; the FP32 form's
    error reported beside it."""
    N, nf, T = 24, 16384, 64
    print(f"C3 shape FP16 vs FP64 N={N} nf={nf} T={T}", flush=True)
    x, valid, gains, Yq, V, raw = _pipeline(corr, oracle, hip, 1, N, nf, T, 32, 64, 0.1, 3, 32)
    ref = oracle.fx_f64(x, valid=valid)
    e16 = rel_rms(V, ref)
    B = N * (N + 1) // 2
    eng = corr.Engine(N, nf, 32)
    d_in = hip.Buffer.from_array(x)
    mask = np.full((N, T // 8), 0xFF, dtype=np.uint8)
    mask[1, 0] &= ~np.uint8(1 << 3)
    d_mask = hip.Buffer.from_array(mask)
    acc = hip.Buffer(B * nf * 8)
    acc.zero()
    for f0 in range(0, T, 32):
        eng.fengine(d_in.ptr + f0 * 2 * nf, T * 2 * nf, 32, d_mask.ptr, T // 8, f0)
        eng.xengine(acc.ptr, 32)
    eng.dump(acc.ptr, T, acc.ptr)
    hip.synchronize()
    e32 = rel_rms(acc.to_numpy(np.complex64, (B, nf)), ref)
    eng.close()
    print(f"  relative RMS vs FP64: FP16 form {e16:.3e}, FP32 form {e32:.3e}")
    assert e16 < 1e-3
    assert e32 < 1e-5


@pytest.mark.parametrize("form,nf", [(1, 2048), (1, 16384), (2, 2048), (2, 16384), (1, 12288), (2, 14336),
                                     (1, 18432), (2, 24576), (1, 32768), (1, 49152), (2, 65536)])
def test_fengine_form_rounding(corr, oracle, hip, form, nf):
    """The form's F-engine output is the FP32 form's spectrum y requantized on
    the host: FP16 = y*g rounded to nearest even (numpy float16), int8 =
    clamp(rint(y*g), +-7), for at least 99.9% of values and every value within
    one FP16 ulp / one integer step (the two kernels share the FFT, so the rare
    difference can only come from a 1-ulp FP32 difference at a rounding tie).
    Pins the rounding mode of the device's FP16 conversion. Above 16384 the
    forms' spectra come from the one-kernel staged transform at every length
    here, the FP32 form's from the two-kernel one at 18432, 24576 and 49152
    and from the one-kernel one at 32768 and 65536, so each pair of staged
    paths is compared."""
    N, T = 4, 8
    print(f"F-engine form {form} rounding N={N} nf={nf}", flush=True)
    x = oracle.gen_int8(21 + form, N, T, 2 * nf, 0.3, sigma=16.0, max_delay=4)
    d_in = hip.Buffer.from_array(x)
    pitch = T * 2 * nf
    e32 = corr.Engine(N, nf, T, corr.FORM_FP32)
    e32.fengine(d_in.ptr, pitch, T)
    y = fp32_spectra(hip, e32, N, T, nf)
    e32.close()
    eng = corr.Engine(N, nf, T, form)
    eng.calibrate(d_in.ptr, pitch, T)
    eng.fengine(d_in.ptr, pitch, T)
    hip.synchronize()
    g = np.empty((N, nf), dtype=np.float32)
    hip.check(hip.rt().hipMemcpy(g.ctypes.data, eng.gains(), g.nbytes, hip.D2H), "hipMemcpy")
    got = _download_spectra(hip, eng, N, T, nf, form)
    eng.close()
    ys = y * g[:, None, :]                       # exact: FP16 gains are powers of two
    comp = np.stack([ys.real, ys.imag], -1).astype(np.float32)
    have = np.stack([got.real, got.imag], -1)
    if form == corr.FORM_FP16:
        want = comp.astype(np.float16)
        hb = have.astype(np.float16).view(np.int16).astype(np.int32)
        wb = want.view(np.int16).astype(np.int32)
        same = np.mean(hb == wb)
        worst = int(np.max(np.abs(hb - wb)))
    else:
        want = np.clip(np.rint(comp), -7, 7)
        same = np.mean(have == want)
        worst = float(np.max(np.abs(have - want)))
    print(f"  equal {same:.6f}, largest difference {worst}", flush=True)
    assert same >= 0.999 and worst <= 1


def test_xtile_pinned_by_environment(corr, hip, monkeypatch):
    """CORRX_XTILE=<channels>,<spectra tile> (0 = n_freq) pins the tile choice
    at creation where it fits the shape; where it does not, the tuner runs."""
    monkeypatch.setenv("CORRX_XTILE", "16,0")
    eng = corr.Engine(4, 512, 8, corr.FORM_FP16)
    assert eng.xtile() == (16, 512)
    eng.close()
    monkeypatch.setenv("CORRX_XTILE", "32,32")
    eng = corr.Engine(4, 1024, 8, corr.FORM_INT8)
    assert eng.xtile() == (32, 32)
    eng.close()
    eng = corr.Engine(40, 512, 8, corr.FORM_FP16)     # 32 channels do not fit 40 inputs
    ch, tl = eng.xtile()
    assert ch == 16 and tl in (16, 512)
    eng.close()
