"""Pin the CPU oracle before it is trusted as the checker.

Each function of oracle/corr_oracle.c is checked against something it does not
share code with: brute-force DFT, numpy's FFT (an independent library), closed
forms (impulse, pure tone, delay phase slope), invariants (Hermitian
autocorrelations, Parseval), the reference's own tolerances, and committed
fixtures written by tests/golden/make_golden.py.
"""

import hashlib
import json
import os

import numpy as np
import pytest

from conftest import rel_rms

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# powers of two, and odd parts 3, 5, 7, 9, 15 (n_freq = 2048 * odd is a valid
# reference configuration)
@pytest.mark.parametrize("L", [8, 16, 64, 128, 6, 24, 30, 40, 56, 72, 120, 3 * 64])
def test_fft_matches_bruteforce_dft(oracle, L):
    rng = np.random.default_rng(L)
    x = rng.integers(-128, 128, L).astype(np.float64)
    a = oracle.fft_r2c_f64(x)
    b = oracle.dft_r2c_f64(x)
    assert np.abs(a - b).max() <= 1e-11 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("L", [2048, 4096, 8192, 16384, 32768, 2 * 6144, 2 * 10240, 2 * 12288, 2 * 14336,
                               2 * 18432])
def test_fft_matches_numpy(oracle, L):
    rng = np.random.default_rng(L)
    x = rng.integers(-128, 128, L).astype(np.float64)
    ref = np.fft.rfft(x)[: L // 2]  # Nyquist dropped
    assert rel_rms(oracle.fft_r2c_f64(x), ref) < 1e-13


def test_fft_closed_forms(oracle):
    L = 4096
    # impulse at t=0 -> flat spectrum of ones
    x = np.zeros(L)
    x[0] = 1.0
    assert np.allclose(oracle.fft_r2c_f64(x), 1.0, atol=1e-12)
    # impulse at t=1 -> exp(-2 pi i k / L)
    x = np.zeros(L)
    x[1] = 1.0
    k = np.arange(L // 2)
    assert np.allclose(oracle.fft_r2c_f64(x), np.exp(-2j * np.pi * k / L), atol=1e-12)
    # pure cosine at bin c with amplitude A -> A*L/2 at bin c, ~0 elsewhere
    c, A = 300, 8.0
    x = A * np.cos(2 * np.pi * c * np.arange(L) / L)
    y = oracle.fft_r2c_f64(x)
    assert abs(y[c] - A * L / 2) < 1e-8
    y[c] = 0
    assert np.abs(y).max() < 1e-8
    # Parseval with the Nyquist bin added back
    rng = np.random.default_rng(0)
    x = rng.standard_normal(L)
    y = oracle.fft_r2c_f64(x)
    nyq = np.sum(x * (-1.0) ** np.arange(L))
    energy = (np.abs(y[0]) ** 2 + 2 * np.sum(np.abs(y[1:]) ** 2) + nyq ** 2) / L
    assert abs(energy - np.sum(x * x)) < 1e-9 * np.sum(x * x)


def test_pairs_kat(oracle):
    # corr-cuda/tests/abi_smoke.c:86-93
    assert oracle.pairs(4).ravel().tolist() == [0, 0, 0, 1, 0, 2, 0, 3, 1, 1, 1, 2, 1, 3, 2, 2,
                                                 2, 3, 3, 3]


def test_fx_f64_against_committed_fixtures(oracle):
    fx = json.load(open(os.path.join(GOLDEN, "fixtures.json")))
    x = oracle.gen_int8(1, 4, 64, 2048, 0.3, sigma=16.0, max_delay=8)
    assert hashlib.sha256(x.tobytes()).hexdigest() == fx["generator_sha256"]["c1"]
    want = np.load(os.path.join(GOLDEN, "c1_expect.npy"))
    assert rel_rms(oracle.fx_f64(x), want) < 1e-6  # fixture stored as complex64
    g = np.load(os.path.join(GOLDEN, "g4_input.npy"))
    want = np.load(os.path.join(GOLDEN, "g4_expect.npy"))
    assert rel_rms(oracle.fx_f64(g, frames_div=40), want) < 1e-6


def test_generator_pinned_and_plausible(oracle):
    fx = json.load(open(os.path.join(GOLDEN, "fixtures.json")))
    cfgs = {
        "c1": (1, 4, 2048, 0.3, 8, 0, 0.0, 64), "c2": (2, 24, 16384, 0.3, 32, 3000, 8.0, 2),
        "c3": (3, 24, 32768, 0.1, 32, 0, 0.0, 2), "c4": (4, 40, 32768, 0.05, 64, 0, 0.0, 2),
        "c5a": (51, 20, 16384, 0.3, 32, 0, 0.0, 2), "c5b": (52, 24, 32768, 0.1, 32, 0, 0.0, 2),
        "c5c": (53, 4, 32768, 0.5, 32, 0, 0.0, 2),
    }
    for name, (seed, N, L, rho, dmax, tch, tamp, fr) in cfgs.items():
        x = oracle.gen_int8(seed, N, fr, L, rho, sigma=16.0, max_delay=dmax, tone_channel=tch,
                            tone_amp=tamp)
        assert hashlib.sha256(x.tobytes()).hexdigest() == fx["generator_sha256"][name], name
        d = oracle.delays(seed, N, dmax)
        assert d.tolist() == fx["delays"][name]
        assert d.min() >= 0 and d.max() <= dmax
    # moments: sigma 16, and correlation rho at lag d_j - d_i
    x = oracle.gen_int8(7, 2, 64, 4096, 0.3, sigma=16.0, max_delay=8)
    d = oracle.delays(7, 2, 8)
    a = x[0].ravel().astype(np.float64)
    b = x[1].ravel().astype(np.float64)
    assert abs(a.std() - 16.0) < 0.2
    lag = int(d[1] - d[0])
    if lag >= 0:
        r = np.corrcoef(a[: a.size - lag], b[lag:])[0, 1]
    else:
        r = np.corrcoef(a[-lag:], b[: b.size + lag])[0, 1]
    assert abs(r - 0.3) < 0.02


def test_generator_tone_closed_form(oracle):
    """The CW tone of the C2 workload, pinned by its spectrum rather than by a
    hash the oracle wrote: with no noise, A cos(2 pi c t / L) rounded to int8 is
    one line at bin c of height A L / 2 (numpy's FFT, within the rounding's
    share), identical on every input and in every frame, and nothing elsewhere
    above the rounding noise; with noise, the tone adds (A L / 2)^2 to the mean
    power of bin c, on top of the noise's sigma^2 L."""
    L, c, A = 2048, 37, 100.0
    x = oracle.gen_int8(5, 3, 4, L, 0.3, sigma=0.0, max_delay=4, tone_channel=c, tone_amp=A)
    assert np.array_equal(x[0], x[2]) and np.array_equal(x[0, 0], x[0, 3])
    Y = np.fft.rfft(x[0, 0].astype(np.float64))
    assert np.argmax(np.abs(Y)) == c
    assert abs(abs(Y[c]) / (A * L / 2) - 1.0) < 1e-3
    others = np.abs(np.delete(Y, c))
    assert others.max() < 1e-3 * abs(Y[c])
    # the C2 tone in noise, at the C2 geometry's channel and amplitude
    L, c, A = 16384, 3000, 8.0
    x = oracle.gen_int8(2, 2, 16, L, 0.3, sigma=16.0, max_delay=32, tone_channel=c, tone_amp=A)
    P = (np.abs(np.fft.rfft(x.astype(np.float64), axis=2)) ** 2).mean(axis=1)   # (inputs, bins)
    assert np.all(np.argmax(P, axis=1) == c)
    want = (A * L / 2) ** 2 + 16.0 ** 2 * L
    assert np.all(np.abs(P[:, c] / want - 1.0) < 0.05)
    assert abs(np.median(P[:, 100:2900]) / (16.0 ** 2 * L) - 1.0) < 0.1


def test_delay_phase_slope(oracle):
    # tools/phase_check.py:4-15 closed form: x1[t] = x0[t - tau] gives
    # phase(V_01[c]) = 2 pi c tau / L. rho = 1 makes x_i = s[t - d_i] exactly.
    L, T, seed = 4096, 16, 9
    x = oracle.gen_int8(seed, 2, T, L, 1.0, sigma=16.0, max_delay=6)
    d = oracle.delays(seed, 2, 6)
    tau = int(d[1] - d[0])
    V = oracle.fx_f64(x)
    c = np.arange(1, L // 8)
    phase = np.unwrap(np.angle(V[1][c]))
    slope = np.polyfit(c, phase, 1)[0]
    assert abs(slope - 2 * np.pi * tau / L) < 2e-3 * 2 * np.pi / L + 1e-9
    # autocorrelations are real and non-negative; V_kj = conj(V_jk) by construction
    assert np.abs(V[0].imag).max() < 1e-9 * np.abs(V[0].real).max()
    assert V[0].real.min() >= 0 and V[2].real.min() >= 0


def test_mask_semantics(oracle):
    x = oracle.gen_int8(3, 3, 10, 2048, 0.3)
    valid = np.ones((3, 10), dtype=np.uint8)
    valid[1, 4] = 0
    valid[2, 0] = 0
    z = x.copy()
    z[1, 4] = 0
    z[2, 0] = 0
    assert np.array_equal(oracle.fx_f64(x, valid=valid), oracle.fx_f64(z))
    # divided by the nominal count, not the valid one (gen_golden.py:77-78)
    assert np.allclose(oracle.fx_f64(z, frames_div=20) * 2, oracle.fx_f64(z))


def test_fx_f32_reference_arithmetic(oracle):
    # FP32 path within the reference's golden tolerance (golden.c:41-48), and at
    # float32 rounding level (measured 1.6e-7 on the reference's hardware)
    x = oracle.gen_int8(4, 5, 24, 4096, 0.3)
    v64 = oracle.fx_f64(x)
    v32, threads = oracle.fx_f32(x, chunk_frames=8)
    err = rel_rms(v32, v64)
    assert err < 1e-6
    assert threads >= 1
    # per-chunk flushing changes rounding only
    v32b, _ = oracle.fx_f32(x, chunk_frames=24)
    assert rel_rms(v32b, v64) < 1e-6


def test_x_from_spectra_matches_fx(oracle):
    x = oracle.gen_int8(5, 3, 6, 2048, 0.3)
    Y = oracle.spectra_f64(x)
    v = oracle.x_f64_from_spectra(Y.astype(np.complex64))
    assert rel_rms(v, oracle.fx_f64(x)) < 1e-6
    g = np.full((3, 1024), 0.5, dtype=np.float32)
    v2 = oracle.x_f64_from_spectra((Y * 0.5).astype(np.complex64), gain=g)
    assert rel_rms(v2, oracle.fx_f64(x)) < 1e-6


def test_planted_error_is_caught(oracle):
    # a dropped conjugate or transposed operand in the X-engine would flip the
    # sign of Im(V_jk); the fixture comparison must see it
    want = np.load(os.path.join(GOLDEN, "c1_expect.npy"))
    x = oracle.gen_int8(1, 4, 64, 2048, 0.3, sigma=16.0, max_delay=8)
    v = oracle.fx_f64(x)
    assert rel_rms(np.conj(v), want) > 1e-2
