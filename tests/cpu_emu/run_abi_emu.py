"""Drive the CPU build of the library (tests/cpu_emu/_build/libcorr_cuda_emu.so,
AddressSanitizer) through the golden and ABI-contract sequences. Run as a child
process with libasan preloaded (see tests/test_cpu_emu.py)."""

import ctypes
import importlib
import os
import sys
import time

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


def packed(L, N, nf, n_cum, fps, rows):
    """4-bit staging (corrx_create_packed) against the 8-bit context on the same
    samples in [-8, 7]: byte-identical rows, all valid and mask-marked, and half
    the sample bytes copied."""
    rng = np.random.default_rng(7)
    x = rng.integers(-8, 8, size=(rows, N, n_cum, 2 * nf)).astype(np.int8)
    cfg = corr.make_config(N, nf, fps, min(fps, 128), block_slots=3, rows_per_fetch=rows)
    s8, s4 = [], []
    g8 = drive(corr, cfg, x, rows, n_cum, ALL_VALID, lib=L, stats=s8)
    g4 = drive(corr, cfg, x, rows, n_cum, ALL_VALID, lib=L, stats=s4, sample_bits=4)
    m8 = drive(corr, cfg, x, rows, n_cum, MASK_MARKED, lib=L)
    m4 = drive(corr, cfg, x, rows, n_cum, MASK_MARKED, lib=L, sample_bits=4)
    same = g4.tobytes() == g8.tobytes() and m4.tobytes() == m8.tobytes()
    masks = s8[0].transfers * N * ((fps + 7) // 8)
    half = 2 * (s4[0].bytes_h2d - masks) == s8[0].bytes_h2d - masks
    e = rel_rms(g4, np.stack([oracle.fx_f64(x[r], frames_div=n_cum) for r in range(rows)]))
    print(f"packed N={N} nf={nf}: 4-bit == 8-bit {same}, half the sample bytes {half}, rel RMS {e:.3e}",
          flush=True)
    return same and half and e < 1e-4


def ptr(a):
    return a.ctypes.data


def skip(L):
    """Chunks in which no station has a present frame are skipped (no copy): the
    same bytes as transforming them as zeros, and exactly their bytes fewer copied."""
    N, nf, fps, chunk, n_cum, rows = 3, 2048, 32, 8, 96, 2

    def pattern(r, f):  # row 0: a whole slot; row 1: one whole chunk and part of another
        return 32 <= f < 64 if r == 0 else (8 <= f < 16 or 70 <= f < 76)

    rng = np.random.default_rng(5)
    x = rng.integers(-4, 5, size=(rows, N, n_cum, 2 * nf)).astype(np.int8)
    cfg = corr.make_config(N, nf, fps, chunk, block_slots=3, rows_per_fetch=rows)
    zs, ms = [], []
    z = drive(corr, cfg, x, rows, n_cum, HOST_ZEROED, lib=L, missing=pattern, stats=zs)
    m = drive(corr, cfg, x, rows, n_cum, MASK_MARKED, lib=L, missing=pattern, stats=ms)
    same = z.tobytes() == m.tobytes()
    saved = zs[0].bytes_h2d - ms[0].bytes_h2d
    want = 5 * N * (chunk * 2 * nf + (fps + 7) // 8)
    print(f"skip: zeroed == masked {same}, copied bytes saved {saved} (want {want})", flush=True)
    return same and saved == want


def forms(L, form, N, nf, T, chunk, cal_frames):
    """Engine path of a precision form in the CPU build: calibrate, then
    F-engine + X-engine per chunk with a missing frame, then the dump; against
    the FP64 oracle on the engine's own requantized spectra."""
    x = oracle.gen_int8(31 + form, N, T, 2 * nf, 0.3, max_delay=4)
    mask_pitch = (T + 7) // 8 + 1
    mask = np.full((N, mask_pitch), 0xFF, dtype=np.uint8)
    mask[1, 0] &= ~np.uint8(1 << 3)  # station 1, frame 3 missing
    valid = np.ones((N, T), dtype=bool)
    valid[1, 3] = False
    e = corr.Engine(N, nf, chunk, form, lib=L)
    pin_ok = os.environ.get("CORRX_XTILE") != "16,16" or e.xtile() == (16, 16)
    pitch = T * 2 * nf
    e.calibrate(ptr(x), pitch, cal_frames, ptr(mask), mask_pitch, 0)
    gains = np.ctypeslib.as_array(ctypes.cast(e.gains(), ctypes.POINTER(ctypes.c_float)),
                                  shape=(N * nf,)).reshape(N, nf).copy()
    B = N * (N + 1) // 2
    ib = e.integrator_bytes()
    integ = np.zeros(B * nf * ib // 4, dtype=np.uint32)
    sp, stride = e.spectra()
    eb = 4 if form == corr.FORM_FP16 else 2
    Yq = np.zeros((N, T, nf), dtype=np.complex64)
    for f0 in range(0, T, chunk):
        n = min(chunk, T - f0)
        e.fengine(ptr(x) + f0 * 2 * nf, pitch, n, ptr(mask), mask_pitch, f0)
        raw = np.ctypeslib.as_array(ctypes.cast(sp, ctypes.POINTER(ctypes.c_uint8)),
                                    shape=(N * stride * eb,)).reshape(N, stride * eb)
        # tile-major order: [station][nf/tile][chunk frames][tile channels]
        dt = np.float16 if form == corr.FORM_FP16 else np.int8
        tl = e.xtile()[1]
        h = raw.view(dt).astype(np.float32).reshape(N, nf // tl, stride // nf, tl, 2)[:, :, :n]
        h = h.transpose(0, 2, 1, 3, 4).reshape(N, n, nf, 2)
        Yq[:, f0:f0 + n] = h[..., 0] + 1j * h[..., 1]
        e.xengine(ptr(integ), n)
    # the two-buffer pipeline over the whole integration adds the same runs in
    # the same order: the integrator must match byte for byte
    integ2 = np.zeros_like(integ)
    e.integrate(ptr(x), pitch, T, ptr(integ2), ptr(mask), mask_pitch, 0)
    same = integ2.tobytes() == integ.tobytes()
    out = np.zeros(B * nf, dtype=np.complex64)
    e.dump(ptr(integ), T, ptr(out))
    e.close()
    V = out.reshape(B, nf)
    ref_q = oracle.x_f64_from_spectra(Yq, gain=gains, frames_div=T)
    ref = oracle.fx_f64(x, valid=valid)
    eq = rel_rms(V, ref_q)
    ok = eq < 1e-6
    if form == corr.FORM_INT8:
        acc = integ.view(np.int32).reshape(B, nf, 2)
        exact = oracle.x_f64_from_spectra(Yq, frames_div=1)
        ok = ok and np.array_equal(acc[..., 0], exact.real) and np.array_equal(acc[..., 1], exact.imag)
    missing_zero = not np.any(Yq[1, 3])
    e64 = rel_rms(V, ref)
    bound = 1e-3 if form == corr.FORM_FP16 else 0.3
    ok = ok and missing_zero and e64 < bound and np.all(gains > 0) and same and pin_ok
    print(f"form {form} N={N} nf={nf} T={T} chunk={chunk}: vs oracle on engine spectra {eq:.3e}, "
          f"vs FP64 {e64:.3e}, missing frame zero {missing_zero}, integrate == loop {same}", flush=True)
    return ok


def integrate_fp32(L, N, nf, T, chunk):
    """FP32 form: corrx_integrate against the F-engine + X-engine loop (same
    bytes) and against the FP64 oracle."""
    x = oracle.gen_int8(77, N, T, 2 * nf, 0.3, max_delay=4)
    e = corr.Engine(N, nf, chunk, corr.FORM_FP32, lib=L)
    pitch = T * 2 * nf
    B = N * (N + 1) // 2
    a = np.zeros(B * nf, dtype=np.complex64)
    for f0 in range(0, T, chunk):
        n = min(chunk, T - f0)
        e.fengine(ptr(x) + f0 * 2 * nf, pitch, n)
        e.xengine(ptr(a), n)
    b = np.zeros_like(a)
    e.integrate(ptr(x), pitch, T, ptr(b))
    e.integrate(ptr(x), pitch, 0, ptr(b))  # no frames: no work
    e.close()
    same = a.tobytes() == b.tobytes()
    err = rel_rms(b.reshape(B, nf) / T, oracle.fx_f64(x))
    print(f"integrate FP32 N={N} nf={nf} T={T} chunk={chunk}: integrate == loop {same}, vs FP64 {err:.3e}",
          flush=True)
    return same and err < 1e-5


def batch(L):
    """corrx_batch_run in direct mode (two engines of different shapes and forms)
    against each engine alone: the same bytes. Graph mode needs a runtime that
    defers work, which the stand-in runtime does not: it must fail cleanly."""
    cases = [(corr.FORM_FP16, 3, 512, 8, 20), (corr.FORM_FP32, 2, 512, 16, 33)]
    subs = []
    for i, (form, N, nf, chunk, T) in enumerate(cases):
        x = oracle.gen_int8(200 + i, N, T, 2 * nf, 0.3, max_delay=4)
        e = corr.Engine(N, nf, chunk, form, lib=L)
        if form != corr.FORM_FP32:
            e.calibrate(ptr(x), T * 2 * nf, T)
        B = N * (N + 1) // 2
        nb = B * nf * e.integrator_bytes()
        subs.append(dict(x=x, e=e, T=T, nf=nf, B=B, pitch=T * 2 * nf,
                         acc=np.zeros(nb // 4, dtype=np.uint32), out=np.zeros(B * nf, dtype=np.complex64)))
    ref = []
    for s in subs:
        s["acc"][:] = 0
        s["e"].integrate(ptr(s["x"]), s["pitch"], s["T"], ptr(s["acc"]))
        s["e"].dump(ptr(s["acc"]), s["T"], ptr(s["out"]))
        ref.append((s["acc"].copy(), s["out"].copy()))
    b = corr.Batch([s["e"] for s in subs], lib=L)
    items = [corr.Batch.item(ptr(s["x"]), s["pitch"], s["T"], ptr(s["acc"]), d_out=ptr(s["out"]),
                             dump_frames=s["T"]) for s in subs]
    for s in subs:
        s["acc"][:] = 0xFFFFFFFF    # the run must zero the integrator itself
        s["out"][:] = 0
    b.run(items, graph=False)
    same = all(np.array_equal(s["acc"], r[0]) and np.array_equal(s["out"], r[1]) for s, r in zip(subs, ref))
    graph_refused = False
    try:
        b.run(items, graph=True)
    except corr.CorrError:
        graph_refused = True
    b.close()
    for s in subs:
        s["e"].close()
    print(f"batch direct == engines alone {same}; graph refused by the stand-in runtime {graph_refused}",
          flush=True)
    return same and graph_refused


def timed(fn, *args):
    t0 = time.perf_counter()
    r = fn(*args)
    print(f"  ({fn.__name__}: {time.perf_counter() - t0:.1f} s)", flush=True)
    return r


# (CORRX_XTILE, case, arguments after the library). The stand-in runtime's
# timer reads zero, so tuning the X-engine tile would only cost emulation time
# here: the tile is pinned (the GPU tests tune). The matrix-core X-engine's
# 512-lane workgroups exchange operands through a workgroup barrier per product
# in this emulation, which makes it slow: one drop-in run covers the FP32 form
# on it, the others pin the FP32 form's vector X-engine ("0,0"); the small
# precision-form shapes take the matrix-core X-engine of every form ("16,16").
CASES = [
    ("16,16", golden, (3, 2048, 20, 20, 1)),
    ("0,0", golden, (4, 2048, 90, 32, 2)),
    ("0,0", golden, (5, 4096, 37, 16, 3)),
    ("0,0", golden, (3, 6144, 20, 16, 1)),   # n_freq = 3 x 2048
    ("0,0", golden, (2, 18432, 6, 4, 1)),    # 9 x 2048, staged through HBM
    ("0,0", packed, (3, 2048, 40, 16, 2)),
    ("0,0", skip, ()),
    ("16,16", forms, (corr.FORM_FP16, 5, 512, 70, 32, 20)),
    ("16,16", forms, (corr.FORM_INT8, 6, 512, 70, 32, 70)),
    ("16,16", integrate_fp32, (3, 512, 45, 8)),
    ("16,16", batch, ()),
]


def run_case(i):
    xtile, fn, args = CASES[i]
    os.environ["CORRX_XTILE"] = xtile
    L = corr.declare(ctypes.CDLL(os.path.join(HERE, "_build", "libcorr_cuda_emu.so")))
    return timed(fn, L, *args)


def main():
    """Every case in a process of its own (the cases share nothing), at most
    eight at a time; each prints its own lines, shown in case order."""
    if len(sys.argv) == 3 and sys.argv[1] == "--case":
        return 0 if run_case(int(sys.argv[2])) else 1
    import subprocess
    from concurrent.futures import ThreadPoolExecutor

    def child(i):
        return subprocess.run([sys.executable, os.path.abspath(__file__), "--case", str(i)],
                              capture_output=True, text=True, timeout=900)

    workers = max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        results = list(pool.map(child, range(len(CASES))))
    ok = True
    for i, r in enumerate(results):
        print(f"case {i}: {CASES[i][1].__name__}{CASES[i][2]} CORRX_XTILE={CASES[i][0]}", flush=True)
        print(r.stdout, end="", flush=True)
        if r.returncode != 0:
            print(r.stderr[-3000:], flush=True)
            ok = False
    print("OK" if ok else "FAILED")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
