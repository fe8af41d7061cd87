#!/usr/bin/env python3
"""Throughput of the MI355X FX correlator hot path on one GPU.

Workload (one "step" = one integration, inputs already resident in HBM):
24 int8 inputs, n_freq = 16384 (32768-sample real FFTs), 2048 spectra per
integration: F-engine (mask + unpack + R2C FFT) and X-engine over every chunk of
frames, then the 1/frames row scale. Synthetic correlated-Gaussian int8 data
(sigma 16, rho 0.1, delays in [0, 32]): 64 distinct frames per input generated
with numpy, tiled over the integration in HBM.

Prints one JSON line: Msamples/s per input (2*n_freq*T / step time), the
roofline of the dominant kernel (HIP events on the launch stream), and the
CPU baseline (the oracle's FP32 restatement of the reference path, timed on a
bounded sample on this host).
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


def measured_traffic(N, nf, chunk):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size);
    empty when there is none for this geometry."""
    path = os.path.join(ROOT, "profiles", f"fx_fp32_{N}in_nf{nf}_traffic.json")
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
    # synthetic input: a pool of distinct frames per input, tiled over the
    # integration in HBM (the throughput does not depend on the sample values)
    P = min(64, T)
    pool = hip.Buffer.from_array(gen_pool(N, P, L, args.rho, 16.0, args.max_delay))
    x = hip.Buffer(N * T * L)
    for i in range(N):
        for c in range(T // P):
            hip.copy_d2d(x.ptr + (i * T + c * P) * L, pool.ptr + i * P * L, P * L)
    pool.free()
    eng = corr.Engine(N, nf, chunk, corr.FORM_FP32)
    acc = hip.Buffer(B * nf * 8)
    stream = hip.Stream()
    sp = stream.handle
    n_chunks = T // chunk

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
        corr.scale(acc.ptr, B * nf, T, sp)

    for _ in range(args.warmup):
        step()
    stream.synchronize()

    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    stream.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = 2 * nf * T / (wall / args.steps) / 1e6

    # per-kernel durations from HIP events on the launch stream (separate pass,
    # so the timed loop above carries no event overhead)
    ev = {"f": [], "x": []}
    for _ in range(max(1, min(args.steps, 3))):
        step(ev)
    stream.synchronize()
    f_ms = sum(a.elapsed_ms(b) for a, b in ev["f"]) / len(ev["f"])
    x_ms = sum(a.elapsed_ms(b) for a, b in ev["x"]) / len(ev["x"])

    f_bytes = N * chunk * L * 1 + N * chunk * nf * 8          # int8 in + complex64 out
    x_flops = 8 * B * nf * chunk                               # useful, lower triangle
    x_bytes = 8 * N * nf * chunk + 16 * B * nf                 # spectra in + row RMW
    f_gbs = f_bytes / (f_ms * 1e-3) / 1e9
    x_tflops = x_flops / (x_ms * 1e-3) / 1e12
    x_gbs = x_bytes / (x_ms * 1e-3) / 1e9
    kernels = {
        "fengine_fp32": {"avg_ms": round(f_ms, 4), "algorithmic_bytes": f_bytes,
                         "GB/s": round(f_gbs, 1), "frac_hbm": round(f_gbs / HBM_PEAK_GBS, 4)},
        "xengine_fp32": {"avg_ms": round(x_ms, 4), "algorithmic_flops": x_flops,
                         "TFLOP/s": round(x_tflops, 2),
                         "frac_fp32": round(x_tflops / FP32_PEAK_TFLOPS, 4),
                         "GB/s": round(x_gbs, 1)},
    }
    traffic = measured_traffic(N, nf, chunk)
    if f_ms >= x_ms:
        roof = {"bound": "hbm", "achieved": round(f_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(f_gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get("fengine_fp32"),
                "kernel": "fengine_fp32"}
    else:
        roof = {"bound": "mfma", "achieved": round(x_tflops, 2), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(x_tflops / FP32_PEAK_TFLOPS, 4),
                "traffic": traffic.get("xengine_fp32"), "kernel": "xengine_fp32"}

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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"FX integration, {N} int8 inputs, n_freq {nf}, {T} spectra, "
                               f"FP32 form, inputs resident in HBM",
                   "n_station": N, "n_freq": nf, "spectra": T, "chunk_frames": chunk,
                   "form": "fp32"},
        "roofline": roof,
        "kernels": kernels,
        "realtime_Msps_per_input": 800.0,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
