#!/usr/bin/env python3
"""Throughput of the MI355X FX correlator hot path on one GPU.

Workload (one "step" = one integration, inputs already resident in HBM):
24 int8 inputs, n_freq = 16384 (32768-sample real FFTs), 2048 spectra per
integration: F-engine (mask + unpack + R2C FFT + the form's requantization) and
X-engine over every chunk of frames (2048 frames per launch, the fastest
measured for every form), then the dump of the integration into complex64
visibilities. The step is one corrx_integrate
call and the dump; the plain per-chunk loop is timed beside it and must give
the same bytes. Synthetic correlated-Gaussian int8 data (sigma 16,
rho 0.1, delays in [0, 32]): 64 distinct frames per input generated with numpy,
tiled over the integration in HBM.

Forms, on the same input: FP16 (the headline: FP16 MFMA X-engine, compensated
FP32 integration, FP64 dump rescale; gains calibrated once, before timing), the
plain FP32 form (the drop-in numerics) and the int8 form. Each form's
visibilities are compared with FP64 visibilities computed here with numpy (the
64-frame pool tiled 32 times integrates to exactly the pool's own average).

Prints one JSON line: Msamples/s per input (2*n_freq*T / step time), the
roofline of the dominant kernel (HIP events on the launch stream; beside it,
as roofline.profile, the same from the committed rocprofv3 kernel trace of this
workload), the other forms' figures, and the CPU baseline (the oracle's FP32 restatement of the
reference path, timed on a bounded sample on this host).
"""

import argparse
import importlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Msamples/s/input at 24 in, nf16384; %HBM/MFMA roofline; MFMA vs FP32 form"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP16_MFMA_PEAK_TFLOPS = 2500.0
INT8_MFMA_PEAK_TOPS = 5000.0
FORMS = {"fp32": 0, "fp16": 1, "int8": 2}
SPEC_BYTES = {"fp32": 8, "fp16": 4, "int8": 2}      # per complex spectrum value
INTEG_BYTES = {"fp32": 8, "fp16": 16, "int8": 8}    # per integrator element
REQUANT = {"fp32": "none: complex64 spectra",
           "fp16": "IEEE half per component, power-of-two gain per (input, channel), saturated at +-65504",
           "int8": "4-bit: each component rint(y*g) clamped to +-7 (15 levels), held as int8"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n-station", type=int, default=24)
    p.add_argument("--n-freq", type=int, default=16384)
    p.add_argument("--frames", type=int, default=2048, help="spectra per integration")
    p.add_argument("--chunk", type=int, default=0,
                   help="frames per engine launch (0: 2048, at most --frames: the fastest measured "
                        "for every form)")
    p.add_argument("--rho", type=float, default=0.1)
    p.add_argument("--max-delay", type=int, default=32)
    p.add_argument("--forms", default="fp16,fp32,int8",
                   help="forms to time; the first one is the headline value")
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 error check")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target CPU time of the bounded CPU-baseline sample (0 = skip)")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the batched three-sub-array workload (one hipGraph per dump)")
    p.add_argument("--batch-steps", type=int, default=5)
    p.add_argument("--no-abi", action="store_true",
                   help="skip the end-to-end drop-in ABI leg (pinned staging -> host visibilities)")
    return p.parse_args()


def gen_pool(N, P, L, rho, sigma, max_delay, seed=2026):
    """P frames per input of x_i[t] = clip(round(sigma*(sqrt(rho)*s[t-d_i]
    + sqrt(1-rho)*n_i[t])), -127, 127), s and n_i standard normal (numpy)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n_total = P * L
    s = rng.standard_normal(n_total + max_delay, dtype=np.float32)
    d = rng.integers(0, max_delay + 1, N)
    a, b = math.sqrt(rho), math.sqrt(1.0 - rho)
    x = np.empty((N, P, L), dtype=np.int8)
    for i in range(N):
        n = rng.standard_normal(n_total, dtype=np.float32)
        v = sigma * (a * s[max_delay - d[i]: max_delay - d[i] + n_total] + b * n)
        x[i] = np.clip(np.rint(v), -127, 127).astype(np.int8).reshape(P, L)
    return x


def cpu_baseline(N, nf, target_s, chunk):
    """Oracle FP32 path (the reference's arithmetic) on a bounded sample of the same
    geometry, all host threads: one integration's worth of frames at most,
    repeated until about `target_s` seconds of CPU work; returns the baseline dict."""
    from oracle import oracle
    L = 2 * nf
    T = 64
    x = oracle.gen_int8(99, N, T, L, 0.1, sigma=16.0, max_delay=32)
    oracle.fx_f32(x, chunk_frames=min(chunk, T))  # warm-up
    reps, dt = 0, 0.0
    t0 = time.perf_counter()
    while dt < target_s:
        _, threads = oracle.fx_f32(x, chunk_frames=min(chunk, T))
        reps += 1
        dt = time.perf_counter() - t0
    value = 2 * nf * T * reps / dt / 1e6
    return {"value": round(value, 3), "unit": "Msamples/s/input", "cores": threads,
            "kind": "port",
            "sample": f"{reps} x ({N} inputs x {T} frames of {L} samples, n_freq {nf}), FP32 "
                      f"F+X in the reference's arithmetic, {dt:.1f} s on {threads} threads"}


def fp64_visibilities(pool):
    """FP64 visibilities (B, nf) of the frames in `pool` (N, P, L), j <= k
    order, averaged over the P frames: numpy rfft in float64, Nyquist dropped."""
    import numpy as np
    N, P, L = pool.shape
    nf = L // 2
    Y = np.fft.rfft(pool.astype(np.float64), axis=2)[:, :, :nf]     # (N, P, nf)
    Yc = np.ascontiguousarray(Y.transpose(2, 0, 1))                    # (nf, N, P)
    V = Yc @ np.conj(Yc.transpose(0, 2, 1))                            # (nf, N, N)
    j, k = np.triu_indices(N)
    return (V[:, j, k].T / P).astype(np.complex128)


def rel_rms(got, ref, rows=None):
    import numpy as np
    if rows is not None:
        got, ref = got[rows], ref[rows]
    return float(np.sqrt(np.sum(np.abs(got - ref) ** 2) / np.sum(np.abs(ref) ** 2)))


def per_baseline_error(got, ref, N):
    """Relative RMS error of every visibility product (baseline) over its
    channels against FP64: the worst baseline, its (j, k), and the median."""
    import numpy as np
    num = np.sum(np.abs(got.astype(np.complex128) - ref) ** 2, axis=1)
    den = np.sum(np.abs(ref) ** 2, axis=1)
    e = np.sqrt(num / np.where(den > 0, den, 1.0))
    b = int(np.argmax(e))
    j, k = np.triu_indices(N)
    return {"worst": float(e[b]), "worst_baseline": b, "worst_pair": [int(j[b]), int(k[b])],
            "median": float(np.median(e)), "baselines": int(e.size)}


class ProfilerRegion:
    """roctxProfilerResume(0) / roctxProfilerPause(0) around the per-kernel
    timing pass, so that `rocprofv3 --marker-trace --selected-regions` records
    exactly the launches whose HIP-event durations bench.py reports, and not
    the engine's tile tuner, calibration or warm-up. Without a profiler attached
    the calls do nothing; without the roctx library this is a no-op."""

    def __init__(self):
        import ctypes
        self._lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
            try:
                self._lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self._lib is not None:
            for fn in (self._lib.roctxProfilerResume, self._lib.roctxProfilerPause):
                fn.argtypes = [ctypes.c_uint64]
                fn.restype = ctypes.c_int

    def __enter__(self):
        if self._lib is not None:
            self._lib.roctxProfilerResume(0)
        return self

    def __exit__(self, *exc):
        if self._lib is not None:
            self._lib.roctxProfilerPause(0)
        return False


def resident_input(hip, pool_host, T):
    """Device buffer [N][T][L] int8 holding the (N, P, L) pool tiled T/P times."""
    N, P, L = pool_host.shape
    pool = hip.Buffer.from_array(pool_host)
    x = hip.Buffer(N * T * L)
    for i in range(N):
        for c in range(T // P):
            hip.copy_d2d(x.ptr + (i * T + c * P) * L, pool.ptr + i * P * L, P * L)
    pool.free()
    return x


# The batched workload: three independent sub-arrays of uneven shapes, one dump
# of each per run. (inputs, n_freq, spectra per integration, rho, seed, frames
# per launch)
BATCH_SUBARRAYS = [(20, 8192, 512, 0.3, 501, 512), (24, 16384, 2048, 0.1, 502, 2048),
                   (4, 16384, 8192, 0.5, 503, 2048)]


def batch_c5(corr, hip, args, fp64):
    """Three sub-arrays (FP16 form) in one corrx_batch_run per dump, replayed as
    a hipGraph, against the same work run one sub-array after the other on one
    stream. Reports time per batch run, total sample rate, the real-time factor
    (800 Msamples/s per input) and each sub-array's relative RMS vs FP64."""
    import numpy as np
    stream = hip.Stream()
    sp = stream.handle
    subs = []
    for N, nf, T, rho, seed, chunk in BATCH_SUBARRAYS:
        L = 2 * nf
        B = N * (N + 1) // 2
        P = min(64, T)
        pool = gen_pool(N, P, L, rho, 16.0, args.max_delay, seed=seed)
        x = resident_input(hip, pool, T)
        eng = corr.Engine(N, nf, chunk, corr.FORM_FP16)
        eng.calibrate(x.ptr, T * L, P, stream=sp)
        acc = hip.Buffer(B * nf * eng.integrator_bytes())
        out = hip.Buffer(B * nf * 8)
        subs.append(dict(N=N, nf=nf, T=T, L=L, B=B, pool=pool, x=x, eng=eng, acc=acc, out=out,
                         tile=eng.xtile(), step=eng.xstep(), chunk=chunk))
    stream.synchronize()
    batch = corr.Batch([s["eng"] for s in subs])
    items = [corr.Batch.item(s["x"].ptr, s["T"] * s["L"], s["T"], s["acc"].ptr, d_out=s["out"].ptr,
                             dump_frames=s["T"]) for s in subs]

    def sequential():
        for s in subs:
            hip.check(hip.rt().hipMemsetAsync(s["acc"].ptr, 0, s["acc"].nbytes, sp), "hipMemsetAsync")
            s["eng"].integrate(s["x"].ptr, s["T"] * s["L"], s["T"], s["acc"].ptr, stream=sp)
            s["eng"].dump(s["acc"].ptr, s["T"], s["out"].ptr, stream=sp)

    def timed(fn, steps):
        for _ in range(2):
            fn()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        stream.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    ms_seq = timed(sequential, args.batch_steps)
    seq_out = [s["out"].to_numpy(np.uint32, (s["B"] * s["nf"] * 2,)).copy() for s in subs]
    ms_direct = timed(lambda: batch.run(items, graph=False, stream=sp), args.batch_steps)
    ms_graph = timed(lambda: batch.run(items, graph=True, stream=sp), args.batch_steps)
    same = all(np.array_equal(s["out"].to_numpy(np.uint32, (s["B"] * s["nf"] * 2,)), o)
               for s, o in zip(subs, seq_out))
    samples = sum(s["N"] * s["T"] * s["L"] for s in subs)
    inputs = sum(s["N"] for s in subs)
    rate = samples / (ms_graph * 1e-3)
    rep = []
    for s in subs:
        d = {"n_station": s["N"], "n_freq": s["nf"], "spectra": s["T"], "frames_per_launch": s["chunk"],
             "channels_per_workgroup": s["tile"][0], "spectra_tile": s["tile"][1],
             "frames_per_staged_step": s["step"],
             "integration_ms_realtime": round(s["T"] * s["L"] / 800e6 * 1e3, 3)}
        if fp64:
            V = s["out"].to_numpy(np.complex64, (s["B"], s["nf"]))
            ref = fp64_visibilities(s["pool"])
            d["rel_rms_vs_fp64"] = rel_rms(V, ref)
            d["rel_rms_vs_fp64_per_baseline"] = per_baseline_error(V, ref, s["N"])
        rep.append(d)
    batch.close()
    for s in subs:
        s["eng"].close()
        for k in ("x", "acc", "out"):
            s[k].free()
    stream.destroy()
    return {"workload": "3 sub-arrays (20 in nf 8192 512 spectra; 24 in nf 16384 2048; 4 in nf 16384 8192), "
                        "FP16 form, one dump of each per run",
            "ms_per_run_graph": round(ms_graph, 4), "ms_per_run_direct": round(ms_direct, 4),
            "ms_sequential": round(ms_seq, 4), "speedup_vs_sequential": round(ms_seq / ms_graph, 3),
            "graph_equals_sequential": bool(same),
            "Gsamples_per_s_total": round(rate / 1e9, 3),
            "realtime_factor": round(rate / (inputs * 800e6), 2),
            "subarrays": rep}


def abi_end_to_end(corr, args, pool_host, fp64, sample_bits=8):
    """The drop-in path end to end: frames in pinned staging (as the receiver
    leaves them) -> corr_row_accumulate (H2D + F + X, FP32 form) per slot ->
    corr_row_end -> one corr_rows_fetch of the row stack to host memory.
    Geometry of the reference's 24-input nf 16384 YAML: 256 frames per slot,
    11 slots, rows of 2048 frames. The staging is written once before timing;
    each released slot gets its mask bits back (the library clears them on
    release, the receiver sets them). PCIe-inclusive: not the headline value.
    sample_bits 4: a corrx_create_packed context, the same frames requantized
    to 4 bits (clip(rint(x/4), -8, 7)) and packed two per byte."""
    import numpy as np
    N, P, L = pool_host.shape
    nf = L // 2
    if nf % 2048 != 0:
        return {"skipped": "the drop-in ABI takes n_freq in multiples of 2048 (its packet rule)"}
    fps, slots, n_cum, rows = 256, 11, 2048, 4
    cfg = corr.make_config(N, nf, fps, fps, block_slots=slots, rows_per_fetch=rows)
    if sample_bits == 4:
        pool_host = np.clip(np.rint(pool_host / 4.0), -8, 7).astype(np.int8)
        staged = corr.pack4(pool_host)
    else:
        staged = pool_host
    with corr.Correlator(cfg, sample_bits=sample_bits) as c:
        B = c.n_baseline
        for g in range(N):
            for sl in range(slots):
                c.slot_frames(g, sl)[:] = staged[g, (sl * fps + np.arange(fps)) % P]
        masks = [[c.slot_mask(g, sl) for sl in range(slots)] for g in range(N)]
        per_row = n_cum // fps
        out = np.empty(c.fetch_bytes() // 8, dtype=np.complex64)

        def integration(seq0):
            seq = seq0
            for _ in range(rows):
                for _ in range(per_row):
                    sl = seq % slots
                    if seq >= slots:
                        c.row_slot_free(seq)          # release the slot's previous use
                    for g in range(N):
                        masks[g][sl][:] = 0xFF         # all frames present
                    c.row_accumulate(seq, fps)
                    seq += 1
                c.row_end(n_cum)
            c.rows_wait()
            got, _ = c.rows_fetch(out)
            assert got == rows
            return seq

        seq = integration(0)                          # warm-up
        t0 = time.perf_counter()
        seq = integration(seq)
        dt = time.perf_counter() - t0
        frames = rows * n_cum
        res = {"workload": f"drop-in ABI, FP32 form, {N} in, nf {nf}, {fps} frames per slot, {slots} slots, "
                           f"{rows} rows of {n_cum} frames, {sample_bits}-bit samples in pinned staging "
                           f"-> visibilities in host memory",
               "ms_per_row": round(dt / rows * 1e3, 3),
               "Msamples_per_s_per_input": round(2 * nf * frames / dt / 1e6, 1),
               "h2d_GB_per_s": round(N * frames * L * sample_bits / 8 / dt / 1e9, 2),
               "realtime_factor": round(2 * nf * frames / dt / 800e6, 2)}
        if fp64:
            V = out.reshape(rows, B, nf)
            ref = fp64_visibilities(pool_host)
            res["rel_rms_vs_fp64"] = max(rel_rms(V[r], ref) for r in range(rows))
        return res


def measured_traffic(N, nf, chunk, form):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size);
    empty when there is none for this geometry."""
    path = os.path.join(ROOT, "profiles", f"fx_{form}_{N}in_nf{nf}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return {}
    if t.get("workload") != f"{N} inputs, n_freq {nf}, {chunk} frames per launch":
        return {}
    return t


def traced_kernels(N, nf, chunk):
    """Average duration per launch (µs) of each kernel in the committed
    rocprofv3 kernel trace of this workload (the default bench.py command run
    under `rocprofv3 --kernel-trace --marker-trace --selected-regions`);
    empty when there is none for this geometry."""
    path = os.path.join(ROOT, "profiles", f"fx_{N}in_nf{nf}_trace.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return {}, None
    if t.get("workload") != f"{N} inputs, n_freq {nf}, {chunk} frames per launch":
        return {}, None
    return t.get("avg_us", {}), os.path.relpath(path, ROOT)


def valu_issue(N, nf, chunk):
    """VALU issue figures of the F-engines (instructions per CU-cycle and per
    wave and frame, clock) from the committed rocprofv3 counter summary of this
    workload; empty when there is none for this geometry."""
    path = os.path.join(ROOT, "profiles", f"fx_{N}in_nf{nf}_util.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return {}, None
    if t.get("workload") != f"{N} inputs, n_freq {nf}, {chunk} frames per launch":
        return {}, None
    return t, os.path.relpath(path, ROOT)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the multi-device split is out of scope)", file=sys.stderr)
        sys.exit(2)

    pkg = importlib.import_module("21cma_abf_correlator_amd")
    corr = pkg.corr
    corr.load()
    hip = importlib.import_module("21cma_abf_correlator_amd.hip")
    import numpy as np

    N, nf, T = args.n_station, args.n_freq, args.frames
    L = 2 * nf
    B = N * (N + 1) // 2
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    assert forms and all(f in FORMS for f in forms), args.forms

    def chunk_of(name):
        c = args.chunk or min(T, 2048)
        assert T % c == 0, (T, c)
        return c

    # synthetic input: a pool of distinct frames per input, tiled over the
    # integration in HBM (the throughput does not depend on the sample values)
    P = min(64, T)
    assert T % P == 0
    pool_host = gen_pool(N, P, L, args.rho, 16.0, args.max_delay)
    x = resident_input(hip, pool_host, T)
    stream = hip.Stream()
    sp = stream.handle
    out = hip.Buffer(B * nf * 8)
    region = ProfilerRegion()

    def run_form(name):
        form = FORMS[name]
        chunk = chunk_of(name)
        n_chunks = T // chunk
        eng = corr.Engine(N, nf, chunk, form)
        xtile = eng.xtile()
        xstep = eng.xstep()
        acc = hip.Buffer(B * nf * eng.integrator_bytes())
        cal_ms = None
        if form != corr.FORM_FP32:
            t0 = time.perf_counter()
            eng.calibrate(x.ptr, T * L, P, stream=sp)   # gains from the distinct frames
            stream.synchronize()
            cal_ms = (time.perf_counter() - t0) * 1e3

        def step():
            # the whole integration in one call: the X-engine of each chunk runs
            # on the engine's second stream beside the next chunk's F-engine
            hip.check(hip.rt().hipMemsetAsync(acc.ptr, 0, acc.nbytes, sp), "hipMemsetAsync")
            eng.integrate(x.ptr, T * L, T, acc.ptr, stream=sp)
            eng.dump(acc.ptr, T, out.ptr, stream=sp)

        def step_serial():
            # the same work as F-engine, X-engine per chunk on one stream
            hip.check(hip.rt().hipMemsetAsync(acc.ptr, 0, acc.nbytes, sp), "hipMemsetAsync")
            for c in range(n_chunks):
                eng.fengine(x.ptr + c * chunk * L, T * L, chunk, stream=sp)
                eng.xengine(acc.ptr, chunk, stream=sp)
            eng.dump(acc.ptr, T, out.ptr, stream=sp)

        def timed(fn, steps):
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            stream.synchronize()
            return (time.perf_counter() - t0) / steps

        for _ in range(args.warmup):
            step_serial()
            step()
        # the serial loop is timed before and after the timed steps and the two
        # averaged: the clock drifts as the GPU stays loaded, and timing it on
        # one side only would credit or charge the overlap with that drift
        n_serial = max(1, min(args.steps, 3))
        serial_before = timed(step_serial, n_serial)
        vis_serial = out.to_numpy(np.complex64, (B, nf))
        wall = timed(step, args.steps) * args.steps
        vis = out.to_numpy(np.complex64, (B, nf))
        serial_s = 0.5 * (serial_before + timed(step_serial, n_serial))
        assert np.array_equal(vis.view(np.uint32), vis_serial.view(np.uint32)), \
            "overlapped and serial integrations differ"
        # Per-kernel durations from HIP events on the launch stream, each kernel
        # timed over back-to-back launches (the chunks of one integration), so
        # the event records add no gap between launches of the same kernel.
        # This pass alone is the profiler region: the trace of a profiled run
        # holds exactly the launches timed here.
        ev = {"f": [], "x": [], "d": []}

        def span(key, n, launch):
            a, b = hip.Event(), hip.Event()
            a.record(stream)
            for c in range(n):
                launch(c)
            b.record(stream)
            ev[key].append((a, b, n))

        with region:
            for _ in range(max(1, min(args.steps, 3))):
                span("f", n_chunks, lambda c: eng.fengine(x.ptr + c * chunk * L, T * L, chunk, stream=sp))
                span("x", n_chunks, lambda c: eng.xengine(acc.ptr, chunk, stream=sp))
                span("d", 1, lambda c: eng.dump(acc.ptr, T, out.ptr, stream=sp))
            stream.synchronize()
        ms = {k: sum(a.elapsed_ms(b) for a, b, _ in v) / sum(n for _, _, n in v) for k, v in ev.items()}
        eng.close()
        acc.free()
        f_bytes = N * chunk * L + N * chunk * nf * SPEC_BYTES[name]   # int8 in + spectra out
        x_flops = 8 * B * nf * chunk                                  # useful, lower triangle
        x_bytes = SPEC_BYTES[name] * N * nf * chunk + 2 * INTEG_BYTES[name] * B * nf
        f_gbs = f_bytes / (ms["f"] * 1e-3) / 1e9
        x_tflops = x_flops / (ms["x"] * 1e-3) / 1e12
        x_gbs = x_bytes / (ms["x"] * 1e-3) / 1e9
        peak = {"fp32": FP32_PEAK_TFLOPS, "fp16": FP16_MFMA_PEAK_TFLOPS,
                "int8": INT8_MFMA_PEAK_TOPS}[name]
        kern = {
            f"fengine_{name}": {"avg_ms": round(ms["f"], 4), "algorithmic_bytes": f_bytes,
                                "GB/s": round(f_gbs, 1),
                                "frac_hbm": round(f_gbs / HBM_PEAK_GBS, 4)},
            f"xengine_{name}": {"avg_ms": round(ms["x"], 4), "algorithmic_flops": x_flops,
                                "TFLOP/s": round(x_tflops, 2), "peak": peak,
                                "frac_peak": round(x_tflops / peak, 4),
                                "algorithmic_bytes": x_bytes, "GB/s": round(x_gbs, 1),
                                "frac_hbm": round(x_gbs / HBM_PEAK_GBS, 4)},
            f"dump_{name}": {"avg_ms": round(ms["d"], 4)},
        }
        kern[f"fengine_{name}"]["frames_per_launch"] = chunk
        kern[f"xengine_{name}"]["frames_per_launch"] = chunk
        if xtile is not None:
            kern[f"xengine_{name}"]["channels_per_workgroup"] = xtile[0]
            kern[f"xengine_{name}"]["spectra_tile"] = xtile[1]
            kern[f"xengine_{name}"]["frames_per_staged_step"] = xstep
        return {"chunk": chunk, "ms_per_step": wall / args.steps * 1e3,
                "ms_per_step_serial": serial_s * 1e3,
                "value": 2 * nf * T / (wall / args.steps) / 1e6,
                "ms": ms, "kernels": kern, "vis": vis, "calibrate_ms": cal_ms,
                "f_gbs": f_gbs, "f_bytes": f_bytes, "x_tflops": x_tflops, "peak": peak,
                "x_gbs": x_gbs}

    res = {name: run_form(name) for name in forms}
    head = forms[0]
    h = res[head]

    ref = None if args.no_fp64 else fp64_visibilities(pool_host)
    import numpy as np
    jj, kk = np.triu_indices(N)
    cross = np.nonzero(jj != kk)[0]
    per_form = {}
    for name, r in res.items():
        d = {"value": round(r["value"], 2), "ms_per_step": round(r["ms_per_step"], 4),
             "ms_per_step_serial": round(r["ms_per_step_serial"], 4),
             "overlap_speedup": round(r["ms_per_step_serial"] / r["ms_per_step"], 3),
             "kernels": r["kernels"]}
        if r["calibrate_ms"] is not None:
            d["calibrate_ms"] = round(r["calibrate_ms"], 2)
        if ref is not None:
            d["rel_rms_vs_fp64"] = rel_rms(r["vis"], ref)
            if len(cross):
                d["rel_rms_vs_fp64_cross"] = rel_rms(r["vis"], ref, cross)
            d["rel_rms_vs_fp64_per_baseline"] = per_baseline_error(r["vis"], ref, N)
        d["requantization"] = REQUANT[name]
        if name != "fp32" and "fp32" in res:
            d["rel_rms_vs_fp32_form"] = rel_rms(r["vis"], res["fp32"]["vis"].astype(np.complex128))
            # per frame: the forms may launch different frame counts
            f32 = res["fp32"]
            d["xengine_speedup_vs_fp32"] = round((f32["ms"]["x"] / f32["chunk"]) / (r["ms"]["x"] / r["chunk"]), 3)
            d["step_speedup_vs_fp32"] = round(res["fp32"]["ms_per_step"] / r["ms_per_step"], 3)
        per_form[name] = d

    traffic = measured_traffic(N, nf, h["chunk"], head)
    if h["ms"]["f"] >= h["ms"]["x"]:
        roof = {"bound": "hbm", "achieved": round(h["f_gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(h["f_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(f"fengine_{head}"), "kernel": f"fengine_{head}"}
    elif head == "fp32":
        roof = {"bound": "mfma", "achieved": round(h["x_tflops"], 2), "peak": h["peak"],
                "unit": "TFLOP/s", "frac": round(h["x_tflops"] / h["peak"], 4),
                "traffic": traffic.get(f"xengine_{head}"), "kernel": f"xengine_{head}"}
    else:
        # an MFMA form's X-engine reads 2N (FP16) or 4N (int8) flop per byte, far
        # below the MFMA ridge: HBM bounds it
        roof = {"bound": "hbm", "achieved": round(h["x_gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(h["x_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(f"xengine_{head}"), "kernel": f"xengine_{head}"}
    # the same roofline from the committed kernel trace's average duration of
    # that kernel, beside the one from this run's HIP events
    traced, source = traced_kernels(N, nf, h["chunk"])
    if roof["kernel"] in traced:
        us = traced[roof["kernel"]]
        kk = h["kernels"][roof["kernel"]]
        if roof["unit"] == "GB/s":
            ach = kk["algorithmic_bytes"] / (us * 1e-6) / 1e9
        else:
            ach = kk["algorithmic_flops"] / (us * 1e-6) / 1e12
        roof["profile"] = {"avg_us": us, "achieved": round(ach, 2 if roof["unit"] == "TFLOP/s" else 1),
                           "frac": round(ach / roof["peak"], 4), "source": source,
                           "event_over_trace": round(kk["avg_ms"] * 1e3 / us, 4)}
    # an F-engine is bound by VALU issue, not by HBM: about one VALU instruction
    # per CU and cycle is what a CU issues (a probe of independent FMA chains)
    valu, vsource = valu_issue(N, nf, h["chunk"])
    if roof["kernel"] in valu:
        roof["valu_issue"] = dict(valu[roof["kernel"]], peak_per_cu_cycle=1.0, source=vsource)
    value = h["value"]
    ms_per_step = h["ms_per_step"]
    kernels = h["kernels"]

    batch = None if args.no_batch else batch_c5(corr, hip, args, not args.no_fp64)
    abi = None if args.no_abi else abi_end_to_end(corr, args, pool_host, not args.no_fp64)
    abi4 = None if args.no_abi else abi_end_to_end(corr, args, pool_host, not args.no_fp64, sample_bits=4)

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(N, nf, args.cpu_seconds, h["chunk"])

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Msamples/s/input",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp16": "fp16", "int8": "int8"}[head],
        "data": "synthetic",
        "config": {"workload": f"FX integration, {N} int8 inputs, n_freq {nf}, {T} spectra, "
                               f"{head} form, inputs resident in HBM",
                   "n_station": N, "n_freq": nf, "spectra": T, "chunk_frames": h["chunk"],
                   "form": head},
        "roofline": roof,
        "kernels": kernels,
        "forms": per_form,
        "realtime_Msps_per_input": 800.0,
        "cpu_baseline": cpu,
        "batch": batch,
        "abi_end_to_end": abi,
        "abi_end_to_end_4bit": abi4,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
