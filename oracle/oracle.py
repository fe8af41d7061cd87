"""ctypes wrapper of the C oracle (oracle/corr_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker (or the CPU path timed beside the
GPU). The shipped library never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        vp = ctypes.c_void_p
        L.orc_n_baseline.argtypes = [ctypes.c_int]
        L.orc_n_baseline.restype = ctypes.c_int
        L.orc_pairs.argtypes = [ctypes.c_int, i32p]
        L.orc_delays.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, i32p]
        L.orc_gen_int8.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, i8p]
        L.orc_fft_r2c_f64.argtypes = [f64p, ctypes.c_int, f64p]
        L.orc_dft_r2c_f64.argtypes = [f64p, ctypes.c_int, f64p]
        L.orc_fx_f64.argtypes = [i8p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, f64p]
        L.orc_fx_f32.argtypes = [i8p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_uint64, f32p]
        L.orc_fx_f32.restype = ctypes.c_int
        L.orc_x_f64_from_spectra.argtypes = [f32p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, f64p]
        L.orc_spectra_f64.argtypes = [i8p, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p]
        _lib = L
    return _lib


def _valid_ptr(valid, n_station, n_frames):
    if valid is None:
        return None, None
    v = np.ascontiguousarray(valid, dtype=np.uint8)
    assert v.shape == (n_station, n_frames)
    return v, v.ctypes.data_as(ctypes.c_void_p)


def pairs(n_station):
    """Baseline -> (j, k) table, shape (B, 2)."""
    B = n_station * (n_station + 1) // 2
    out = np.zeros(2 * B, dtype=np.int32)
    lib().orc_pairs(n_station, out)
    return out.reshape(B, 2)


def delays(seed, n_station, max_delay):
    out = np.zeros(n_station, dtype=np.int32)
    lib().orc_delays(seed, n_station, max_delay, out)
    return out


def gen_int8(seed, n_station, n_frames, L, rho, sigma=16.0, max_delay=8, tone_channel=0,
             tone_amp=0.0, frame0=0):
    """Correlated-Gaussian int8 samples, shape (n_station, n_frames, L)."""
    out = np.empty((n_station, n_frames, L), dtype=np.int8)
    lib().orc_gen_int8(seed, n_station, frame0, n_frames, L, rho, sigma, max_delay, tone_channel,
                       tone_amp, out)
    return out


def fft_r2c_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(x.size, dtype=np.float64)  # L/2 complex
    lib().orc_fft_r2c_f64(x, x.size, out)
    return out.view(np.complex128)


def dft_r2c_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(x.size, dtype=np.float64)
    lib().orc_dft_r2c_f64(x, x.size, out)
    return out.view(np.complex128)


def fx_f64(x, frames_div=None, valid=None):
    """FP64 correlator: x int8 (N, T, L) -> V complex128 (B, L/2) divided by frames_div."""
    x = np.ascontiguousarray(x, dtype=np.int8)
    N, T, L = x.shape
    if frames_div is None:
        frames_div = T
    B = N * (N + 1) // 2
    V = np.empty(2 * B * (L // 2), dtype=np.float64)
    keep, vp = _valid_ptr(valid, N, T)
    lib().orc_fx_f64(x, vp, N, T, L, float(frames_div), V)
    return V.view(np.complex128).reshape(B, L // 2)


def fx_f32(x, chunk_frames=256, frames_div=None, valid=None):
    """The reference's FP32 arithmetic on the CPU: returns (V complex64 (B, L/2), threads)."""
    x = np.ascontiguousarray(x, dtype=np.int8)
    N, T, L = x.shape
    if frames_div is None:
        frames_div = T
    B = N * (N + 1) // 2
    V = np.empty(2 * B * (L // 2), dtype=np.float32)
    keep, vp = _valid_ptr(valid, N, T)
    threads = lib().orc_fx_f32(x, vp, N, T, L, int(chunk_frames), int(frames_div), V)
    if threads < 0:
        raise ValueError(f"fx_f32: n_freq {L // 2} is not a power of two (its radix-2 FFT)")
    return V.view(np.complex64).reshape(B, L // 2), threads


def spectra_f64(x, valid=None):
    """FP64 spectra, complex128 (N, T, L/2), Nyquist dropped, unnormalised."""
    x = np.ascontiguousarray(x, dtype=np.int8)
    N, T, L = x.shape
    Y = np.empty(2 * N * T * (L // 2), dtype=np.float64)
    keep, vp = _valid_ptr(valid, N, T)
    lib().orc_spectra_f64(x, vp, N, T, L, Y)
    return Y.view(np.complex128).reshape(N, T, L // 2)


def x_f64_from_spectra(Y, gain=None, frames_div=None):
    """FP64 X-engine on given spectra (complex64 (N, T, nf)); gain float32 (N, nf) or None."""
    Yf = np.ascontiguousarray(Y, dtype=np.complex64)
    N, T, nf = Yf.shape
    if frames_div is None:
        frames_div = T
    B = N * (N + 1) // 2
    V = np.empty(2 * B * nf, dtype=np.float64)
    g = None
    gp = None
    if gain is not None:
        g = np.ascontiguousarray(gain, dtype=np.float32)
        gp = g.ctypes.data_as(ctypes.c_void_p)
    lib().orc_x_f64_from_spectra(Yf.view(np.float32).reshape(-1), gp, N, T, nf,
                                 float(frames_div), V)
    return V.view(np.complex128).reshape(B, nf)
