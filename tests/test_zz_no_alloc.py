"""corrx_correlate_noalloc through the C ABI: one FP32-form integration with no
allocation (spectra, twiddles and validity in the library's static scratch
array), against the existing entry points that use an engine's buffers
(corrx_integrate + corrx_dump on an FP32 engine with the same run length and
spectrum tile): byte-identical visibilities, on two calls in a row, the second
with a smaller input (fewer inputs, fewer frames, other masks), and within
1e-5 of the FP64 oracle. The run length K follows the entry point's documented
rule: the frames of one call, or as many as its 256 MiB scratch holds."""

import ctypes
import importlib

import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu

SCRATCH = 256 << 20
ALIGN = 256


def _align(n):
    return (n + ALIGN - 1) // ALIGN * ALIGN


def _run_length(N, nf, frames):
    """Frames per run of corrx_correlate_noalloc (include/corrx.h)."""
    tw = _align(8 * nf)  # power-of-two n_freq: n_freq twiddles
    wl = _align(8 * nf)
    room = SCRATCH - tw - wl - 2 * ALIGN
    fit = room // (N * (8 * nf + 4))
    return max(1, frames) if frames < fit else fit


@pytest.fixture(scope="module")
def hip(corr):
    return importlib.import_module("21cma_abf_correlator_amd.hip")


def _noalloc(corr, hip, x, mask, dump_frames):
    lib = corr.load()
    N, T, L = x.shape
    nf = L // 2
    B = N * (N + 1) // 2
    d_in = hip.Buffer.from_array(x)
    out = hip.Buffer(B * nf * 8)
    # garbage: the entry point must clear what it accumulates into
    hip.check(hip.rt().hipMemset(out.ptr, 0x7F, out.nbytes), "hipMemset")
    d_mask = hip.Buffer.from_array(mask) if mask is not None else None
    s = lib.corrx_correlate_noalloc(N, nf, ctypes.c_void_p(d_in.ptr), T * L,
                                    ctypes.c_void_p(d_mask.ptr if d_mask else None),
                                    mask.shape[1] if mask is not None else 0, 0, T, dump_frames,
                                    ctypes.c_void_p(out.ptr), None)
    corr.check(s, "corrx_correlate_noalloc")
    hip.synchronize()
    return out.to_numpy(np.complex64, (B, nf)).copy()


def _engine(corr, hip, x, mask, dump_frames):
    N, T, L = x.shape
    nf = L // 2
    B = N * (N + 1) // 2
    K = _run_length(N, nf, T)
    eng = corr.Engine(N, nf, K, corr.FORM_FP32)
    eng.set_xtile(16, 16)
    d_in = hip.Buffer.from_array(x)
    acc = hip.Buffer(B * nf * 8)
    acc.zero()
    d_mask = hip.Buffer.from_array(mask) if mask is not None else None
    eng.integrate(d_in.ptr, T * L, T, acc.ptr, d_mask.ptr if d_mask else None,
                  mask.shape[1] if mask is not None else 0, 0)
    out = hip.Buffer(B * nf * 8)
    eng.dump(acc.ptr, dump_frames, out.ptr)
    hip.synchronize()
    got = out.to_numpy(np.complex64, (B, nf)).copy()
    eng.close()
    return got


def test_noalloc_matches_engine_two_calls(corr, oracle, hip):
    # call 1: 24 inputs, n_freq 16384, 200 frames (runs of 85 in the scratch),
    # a missing frame; call 2 (smaller): 6 inputs, n_freq 16384, 37 frames, no mask
    cases = [(24, 16384, 200, 0.1, True), (6, 16384, 37, 0.3, False)]
    for i, (N, nf, T, rho, masked) in enumerate(cases):
        print(f"no-alloc call {i + 1}: {N} inputs, n_freq {nf}, {T} frames, run length "
              f"{_run_length(N, nf, T)}", flush=True)
        x = oracle.gen_int8(900 + i, N, T, 2 * nf, rho, sigma=16.0, max_delay=16)
        mask = None
        valid = None
        if masked:
            mask = np.full((N, (T + 7) // 8), 0xFF, dtype=np.uint8)
            mask[2, 12] &= ~np.uint8(1 << 5)   # station 2, frame 101 missing
            valid = np.ones((N, T), dtype=np.uint8)
            valid[2, 101] = 0
        got = _noalloc(corr, hip, x, mask, T + 3)
        want = _engine(corr, hip, x, mask, T + 3)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"call {i + 1} differs"
        ref = oracle.fx_f64(x, frames_div=T + 3, valid=valid)
        err = rel_rms(got, ref)
        print(f"  byte-identical to the engine path; relative RMS vs FP64 {err:.3e}", flush=True)
        assert err < 1e-5


def test_noalloc_rejections(corr, hip):
    lib = corr.load()
    buf = hip.Buffer(1 << 20)
    p = ctypes.c_void_p(buf.ptr)
    # 49 inputs (beyond the matrix-core X-engine), n_freq above the in-LDS sizes,
    # a dump of 0 frames, a short station pitch
    assert lib.corrx_correlate_noalloc(49, 512, p, 1024, None, 0, 0, 1, 1, p, None) == corr.ERR_ARG
    assert lib.corrx_correlate_noalloc(2, 32768, p, 65536, None, 0, 0, 1, 1, p, None) == corr.ERR_NOT_IMPL
    assert lib.corrx_correlate_noalloc(2, 512, p, 1024, None, 0, 0, 1, 0, p, None) == corr.ERR_ARG
    assert lib.corrx_correlate_noalloc(2, 512, p, 1023, None, 0, 0, 1, 1, p, None) == corr.ERR_ARG
    assert len(corr.last_error()) > 0
