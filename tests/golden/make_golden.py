#!/usr/bin/env python3
"""Write the committed fixtures of tests/golden/ (run from the repo root).

Inputs come from the seeded generator (oracle/corr_oracle.c orc_gen_int8), whose
output is pinned here by SHA-256 per configuration; expected visibilities come
from numpy's FP64 rfft and a plain numpy accumulation -- an independent
library, restating the reference's golden recipe (corr-cuda/tests/gen_golden.py:
66-78: rfft in float64, Nyquist dropped, sum_t Y_j conj(Y_k), divided by n_cum).

Nothing from the reference is imported or run.
"""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

# (name, seed, n_station, L, rho, max_delay, tone_channel, tone_amp, frames_hashed)
CONFIGS = [
    ("c1", 1, 4, 2048, 0.3, 8, 0, 0.0, 64),
    ("c2", 2, 24, 16384, 0.3, 32, 3000, 8.0, 2),
    ("c3", 3, 24, 32768, 0.1, 32, 0, 0.0, 2),
    ("c4", 4, 40, 32768, 0.05, 64, 0, 0.0, 2),
    ("c5a", 51, 20, 16384, 0.3, 32, 0, 0.0, 2),
    ("c5b", 52, 24, 32768, 0.1, 32, 0, 0.0, 2),
    ("c5c", 53, 4, 32768, 0.5, 32, 0, 0.0, 2),
]


def gen(cfg, frames=None):
    name, seed, N, L, rho, dmax, tch, tamp, fh = cfg
    return oracle.gen_int8(seed, N, frames or fh, L, rho, sigma=16.0, max_delay=dmax,
                           tone_channel=tch, tone_amp=tamp)


def numpy_fx(x, frames_div):
    """gen_golden.py's recipe in numpy FP64: (N, T, L) int8 -> (B, L/2) complex128."""
    N, T, L = x.shape
    Y = np.fft.rfft(x.astype(np.float64), axis=2)[:, :, : L // 2]
    pairs = [(j, k) for j in range(N) for k in range(j, N)]
    return np.stack([(Y[j] * np.conj(Y[k])).sum(axis=0) for j, k in pairs]) / frames_div


def main():
    fixtures = {"generator_sha256": {}, "delays": {}}
    for cfg in CONFIGS:
        x = gen(cfg)
        fixtures["generator_sha256"][cfg[0]] = hashlib.sha256(x.tobytes()).hexdigest()
        fixtures["delays"][cfg[0]] = oracle.delays(cfg[1], cfg[2], cfg[5]).tolist()
    x = gen(CONFIGS[0])  # C1 in full: 4 inputs, L = 2048, 64 spectra
    v = numpy_fx(x, 64)
    np.save(os.path.join(HERE, "c1_expect.npy"), v.astype(np.complex64))
    # the golden_partial_slot recipe (gen_golden.py:57-62): samples in [-4, 4] with
    # the two extremes planted, 4 stations, n_freq 2048; one row of 40 frames
    rng = np.random.default_rng(20260914)
    g = rng.integers(-4, 5, size=(4, 40, 4096)).astype(np.int8)
    g[0, 0, 0] = -128
    g[0, 0, 1] = 127
    np.save(os.path.join(HERE, "g4_input.npy"), g)
    np.save(os.path.join(HERE, "g4_expect.npy"), numpy_fx(g, 40).astype(np.complex64))
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(fixtures, f, indent=1, sort_keys=True)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
