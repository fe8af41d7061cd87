#!/usr/bin/env python3
"""Throughput of the MI355X FX correlator hot path on one GPU.

Workload (one "step" = one integration, inputs already resident in HBM):
24 int8 inputs, n_freq = 16384 (32768-sample real FFTs), 2048 spectra per
integration: F-engine (mask + unpack + R2C FFT + the form's requantization) and
X-engine over every chunk of frames, then the dump of the integration into
complex64 visibilities. Synthetic correlated-Gaussian int8 data (sigma 16,
rho 0.1, delays in [0, 32]): 64 distinct frames per input generated with numpy,
tiled over the integration in HBM.

Forms, on the same input: FP16 (the headline: FP16 MFMA X-engine, compensated
FP32 integration, FP64 dump rescale; gains calibrated once, before timing), the
plain FP32 form (the drop-in numerics) and the int8 form. Each form's
visibilities are compared with FP64 visibilities computed here with numpy (the
64-frame pool tiled 32 times integrates to exactly the pool's own average).

Prints one JSON line: Msamples/s per input (2*n_freq*T / step time), the
roofline of the dominant kernel (HIP events on the launch stream), the other
forms' figures, and the CPU baseline (the oracle's FP32 restatement of the
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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n-station", type=int, default=24)
    p.add_argument("--n-freq", type=int, default=16384)
    p.add_argument("--frames", type=int, default=2048, help="spectra per integration")
    p.add_argument("--chunk", type=int, default=512, help="frames per engine launch")
    p.add_argument("--rho", type=float, default=0.1)
    p.add_argument("--max-delay", type=int, default=32)
    p.add_argument("--forms", default="fp16,fp32,int8",
                   help="forms to time; the first one is the headline value")
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 error check")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target CPU time of the bounded CPU-baseline sample (0 = skip)")
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

    N, nf, T, chunk = args.n_station, args.n_freq, args.frames, args.chunk
    L = 2 * nf
    B = N * (N + 1) // 2
    assert T % chunk == 0
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    assert forms and all(f in FORMS for f in forms), args.forms
    # synthetic input: a pool of distinct frames per input, tiled over the
    # integration in HBM (the throughput does not depend on the sample values)
    P = min(64, T)
    assert T % P == 0
    pool_host = gen_pool(N, P, L, args.rho, 16.0, args.max_delay)
    pool = hip.Buffer.from_array(pool_host)
    x = hip.Buffer(N * T * L)
    for i in range(N):
        for c in range(T // P):
            hip.copy_d2d(x.ptr + (i * T + c * P) * L, pool.ptr + i * P * L, P * L)
    pool.free()
    stream = hip.Stream()
    sp = stream.handle
    n_chunks = T // chunk
    out = hip.Buffer(B * nf * 8)

    def run_form(name):
        form = FORMS[name]
        eng = corr.Engine(N, nf, chunk, form)
        xtile = eng.xtile() if form != corr.FORM_FP32 else None
        acc = hip.Buffer(B * nf * eng.integrator_bytes())
        cal_ms = None
        if form != corr.FORM_FP32:
            t0 = time.perf_counter()
            eng.calibrate(x.ptr, T * L, P, stream=sp)   # gains from the distinct frames
            stream.synchronize()
            cal_ms = (time.perf_counter() - t0) * 1e3

        def step(events=None):
            hip.check(hip.rt().hipMemsetAsync(acc.ptr, 0, acc.nbytes, sp), "hipMemsetAsync")
            for c in range(n_chunks):
                ptr = x.ptr + c * chunk * L
                if events is not None:
                    events["f"].append((hip.Event(), hip.Event()))
                    events["f"][-1][0].record(stream)
                eng.fengine(ptr, T * L, chunk, stream=sp)
                if events is not None:
                    events["f"][-1][1].record(stream)
                    events["x"].append((hip.Event(), hip.Event()))
                    events["x"][-1][0].record(stream)
                eng.xengine(acc.ptr, chunk, stream=sp)
                if events is not None:
                    events["x"][-1][1].record(stream)
            if events is not None:
                events["d"].append((hip.Event(), hip.Event()))
                events["d"][-1][0].record(stream)
            eng.dump(acc.ptr, T, out.ptr, stream=sp)
            if events is not None:
                events["d"][-1][1].record(stream)

        for _ in range(args.warmup):
            step()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        stream.synchronize()
        wall = time.perf_counter() - t0
        import numpy as np
        vis = out.to_numpy(np.complex64, (B, nf))
        # per-kernel durations from HIP events on the launch stream (separate
        # pass, so the timed loop above carries no event overhead)
        ev = {"f": [], "x": [], "d": []}
        for _ in range(max(1, min(args.steps, 3))):
            step(ev)
        stream.synchronize()
        ms = {k: sum(a.elapsed_ms(b) for a, b in v) / len(v) for k, v in ev.items()}
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
        if xtile is not None:
            kern[f"xengine_{name}"]["channels_per_workgroup"] = xtile[0]
            kern[f"xengine_{name}"]["spectra_tile"] = xtile[1]
        return {"ms_per_step": wall / args.steps * 1e3,
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
             "kernels": r["kernels"]}
        if r["calibrate_ms"] is not None:
            d["calibrate_ms"] = round(r["calibrate_ms"], 2)
        if ref is not None:
            d["rel_rms_vs_fp64"] = rel_rms(r["vis"], ref)
            if len(cross):
                d["rel_rms_vs_fp64_cross"] = rel_rms(r["vis"], ref, cross)
        if name != "fp32" and "fp32" in res:
            d["rel_rms_vs_fp32_form"] = rel_rms(r["vis"], res["fp32"]["vis"].astype(np.complex128))
            d["xengine_speedup_vs_fp32"] = round(res["fp32"]["ms"]["x"] / r["ms"]["x"], 3)
            d["step_speedup_vs_fp32"] = round(res["fp32"]["ms_per_step"] / r["ms_per_step"], 3)
        per_form[name] = d

    traffic = measured_traffic(N, nf, chunk, head)
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
    value = h["value"]
    ms_per_step = h["ms_per_step"]
    kernels = h["kernels"]

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(N, nf, args.cpu_seconds, chunk)

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
                   "n_station": N, "n_freq": nf, "spectra": T, "chunk_frames": chunk,
                   "form": head},
        "roofline": roof,
        "kernels": kernels,
        "forms": per_form,
        "realtime_Msps_per_input": 800.0,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
