"""Host-only checks of the drop-in library: it loads, exports every declared
symbol, and its pure-host paths (banner, numbering, config validation) follow
the reference. No device is touched.

Reference tests re-expressed: corr-cuda/tests/abi_smoke.c:60-141 and
corr-rs/tests/abi.rs:27-68.
"""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"\b(corrx?_[a-z_0-9]+)\s*\(", text)) - {"corr_notify_fn"}


def test_exports_every_declared_symbol(corr):
    lib = corr.load()
    declared = _declared("corr_cuda.h") | _declared("corrx.h")
    assert declared == set(corr.ABI_SYMBOLS) | set(corr.ENGINE_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", corr.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (corrx?_\w+)", out))
    missing = declared - exported
    assert not missing, f"not exported: {sorted(missing)}"
    for name in declared:
        assert getattr(lib, name) is not None


def test_static_library_built(corr):
    a = os.path.join(os.path.dirname(corr.LIB_PATH), "libcorr_cuda.a")
    assert os.path.exists(a)
    out = subprocess.run(["nm", "--defined-only", a], capture_output=True, text=True,
                         check=True).stdout
    assert " T corr_create" in out and " T corr_rows_fetch" in out


def test_struct_layout(corr):
    # x86-64 sizes of the ABI structs (reference corr_cuda.h:171-294).
    assert ctypes.sizeof(corr.corr_config) == 76
    assert corr.corr_config.struct_size.offset == 4
    assert corr.corr_config.split_mode.offset == 72
    assert ctypes.sizeof(corr.corr_device_plan) == 20
    assert ctypes.sizeof(corr.corr_device_stats) == 72
    assert ctypes.sizeof(corr.corr_staging) == 48


def test_banner_and_error_string(corr):
    b = corr.abi_banner()
    assert "abi=7" in b  # corr-rs/tests/abi.rs:28-36
    lib = corr.load()
    assert lib.corr_last_error() is not None
    s = lib.corr_create(None, None)
    assert s == corr.ERR_ARG
    assert len(corr.last_error()) > 0


def test_baseline_numbering_kat(corr):
    # abi_smoke.c:77-100
    assert [corr.n_baseline(n) for n in (0, 1, 4, 15, 24)] == [-1, 1, 10, 120, 300]
    assert corr.baseline_pairs(4) == [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3),
                                      (2, 2), (2, 3), (3, 3)]
    lib = corr.load()
    buf = (ctypes.c_int32 * 20)()
    assert lib.corr_baseline_pairs(4, buf, 19) == corr.ERR_ARG
    assert lib.corr_baseline_pairs(0, buf, 0) == corr.ERR_ARG


def test_baseline_numbering_all_sizes(corr, oracle):
    # corr-rs/tests/abi.rs:38-68: count and closed-form index for N = 1..24, and
    # the oracle's table agrees.
    for n in range(1, 45):
        pairs = corr.baseline_pairs(n)
        assert len(pairs) == n * (n + 1) // 2
        for b, (j, k) in enumerate(pairs):
            assert j <= k
            assert b == j * n - j * (j - 1) // 2 + (k - j)
        assert [tuple(p) for p in oracle.pairs(n).tolist()] == pairs


def _cfg(corr, **kw):
    base = dict(n_station=4, n_freq=4096, frames_per_slot=16, chunk_frames=8, block_slots=3,
                rows_per_fetch=2)
    base.update(kw)
    return corr.make_config(**base)


@pytest.mark.parametrize("mutate,expect", [
    (lambda c: setattr(c, "abi_version", 8), "ERR_ABI"),
    (lambda c: setattr(c, "struct_size", 72), "ERR_ABI"),
    (lambda c: setattr(c, "n_freq", 4096 + 1024), "ERR_ARG"),
    (lambda c: setattr(c, "block_slots", 1), "ERR_ARG"),
    (lambda c: setattr(c, "chunk_frames", 17), "ERR_ARG"),
    (lambda c: setattr(c, "rows_per_fetch", 0), "ERR_ARG"),
    (lambda c: setattr(c, "n_station", 0), "ERR_ARG"),
    (lambda c: setattr(c, "device_id", -1), "ERR_ARG"),
    (lambda c: setattr(c, "split_mode", 3), "ERR_ARG"),
    (lambda c: setattr(c, "n_devices", 9), "ERR_ARG"),
    (lambda c: setattr(c, "frames_per_slot", 0), "ERR_ARG"),
])
def test_config_rejection_before_device(corr, mutate, expect):
    # abi_smoke.c:102-141: every machine-independent rejection happens before
    # any device is touched, and leaves *out_ctx NULL.
    c = _cfg(corr)
    mutate(c)
    h = ctypes.c_void_p(0x1)
    s = corr.load().corr_create(ctypes.byref(c), ctypes.byref(h))
    assert s == getattr(corr, expect)
    assert not h.value
    assert corr.last_error()


def test_null_handles_rejected(corr):
    lib = corr.load()
    assert lib.corr_destroy(None) == corr.OK
    assert lib.corr_n_devices(None) == -1
    assert lib.corr_row_accumulate(None, 0, 1) == corr.ERR_ARG
    assert lib.corr_row_slot_free(None, 0) == corr.ERR_ARG
    assert lib.corr_row_end(None, 1) == corr.ERR_ARG
    assert lib.corr_rows_wait(None) == corr.ERR_ARG
    assert lib.corr_rows_ready(None) == corr.ERR_ARG
    p = ctypes.c_int32()
    assert lib.corr_rows_pending(None, ctypes.byref(p)) == corr.ERR_ARG
    assert lib.corr_rows_fetch(None, None, 0, None) == corr.ERR_ARG
    assert lib.corr_rows_fetch_ready(None) == corr.ERR_ARG
    s = corr.corr_staging()
    assert lib.corr_staging_info(None, 0, ctypes.byref(s)) == corr.ERR_ARG
    assert lib.corr_host_alloc(0, ctypes.byref(ctypes.c_void_p())) == corr.ERR_ARG
    lib.corr_host_free(None)
    e = ctypes.c_void_p(0x1)
    assert lib.corrx_engine_create(None, ctypes.byref(e)) == corr.ERR_ARG and not e.value
    assert lib.corrx_fengine(None, None, 0, None, 0, 0, 1, None) == corr.ERR_ARG
    assert lib.corrx_scale(None, 1, 1, None) == corr.ERR_ARG
    n = ctypes.c_size_t()
    assert lib.corrx_integrator_bytes(None, ctypes.byref(n)) == corr.ERR_ARG
    assert lib.corrx_dump(None, None, 1, None, None) == corr.ERR_ARG
    assert lib.corrx_calibrate(None, None, 0, None, 0, 0, 1, None) == corr.ERR_ARG
    assert lib.corrx_set_gains(None, None, None) == corr.ERR_ARG
    assert lib.corrx_gains(None, ctypes.byref(ctypes.c_void_p())) == corr.ERR_ARG
    assert lib.corrx_integrate(None, None, 0, None, 0, 0, 1, None, None) == corr.ERR_ARG
    b = ctypes.c_void_p(0x1)
    assert lib.corrx_batch_create(None, 1, ctypes.byref(b)) == corr.ERR_ARG and not b.value
    two = (ctypes.c_void_p * 2)(None, None)
    b = ctypes.c_void_p(0x1)
    assert lib.corrx_batch_create(two, 2, ctypes.byref(b)) == corr.ERR_ARG and not b.value  # NULL engines
    assert lib.corrx_batch_create(two, 0, ctypes.byref(b)) == corr.ERR_ARG
    assert lib.corrx_batch_run(None, None, 1, None) == corr.ERR_ARG
    assert lib.corrx_batch_destroy(None) == corr.OK
    assert ctypes.sizeof(corr.corrx_batch_item) == 72


def test_packed_context_rejections(corr):
    """corrx_create_packed checks the configuration as corr_create does, then the
    sample width, before any device is touched."""
    lib = corr.load()
    h = ctypes.c_void_p(0x1)
    cfg = corr.make_config(4, 2048, 64, 64)
    for bits in (0, 1, 5, 16, -4):
        assert lib.corrx_create_packed(ctypes.byref(cfg), bits, ctypes.byref(h)) == corr.ERR_ARG
        assert not h.value and "sample_bits" in corr.last_error()
    assert lib.corrx_create_packed(ctypes.byref(cfg), 4, None) == corr.ERR_ARG
    bad = corr.make_config(4, 1000, 64, 64)  # the packet rule, as corr_create
    assert lib.corrx_create_packed(ctypes.byref(bad), 4, ctypes.byref(h)) == corr.ERR_ARG


def test_pack4_layout(corr):
    """pack4 puts sample 2m in bits 0-3 and 2m+1 in bits 4-7 of byte m, two's
    complement; decoding the bytes by sign-extending each nibble gives x back."""
    x = np.array([[0, 1, -1, 7, -8, 3, 5, -6]], dtype=np.int8)
    p = corr.pack4(x).view(np.uint8)
    assert p.tolist() == [[0x10, 0x7F, 0x38, 0xA5]]
    rng = np.random.default_rng(4)
    y = rng.integers(-8, 8, size=(3, 5, 64)).astype(np.int8)
    b = corr.pack4(y).astype(np.int16)
    lo = ((b & 0xF) ^ 8) - 8
    hi = b >> 4
    back = np.empty_like(y)
    back[..., 0::2], back[..., 1::2] = lo, hi
    assert np.array_equal(back, y)
    with pytest.raises(ValueError):
        corr.pack4(np.array([8, 0], dtype=np.int8))


@pytest.mark.parametrize("shape,status", [
    ((4, 1024, 16, 7), "ERR_ARG"),        # unknown form
    ((4, 1024, 16, -1), "ERR_ARG"),
    ((49, 1024, 16, 1), "ERR_NOT_IMPL"),  # FP16 form: at most 48 inputs
    ((49, 1024, 16, 2), "ERR_NOT_IMPL"),
    ((4, 1000, 16, 1), "ERR_NOT_IMPL"),   # FFT length
    ((4, 3072, 16, 0), "ERR_NOT_IMPL"),   # odd factor, not a multiple of 2048
    ((4, 1050624, 16, 2), "ERR_NOT_IMPL"),  # 513 * 2048, above 2^20
    ((4, 20000, 16, 0), "ERR_NOT_IMPL"),  # above 16384 and not a multiple of 2048
    ((0, 1024, 16, 2), "ERR_ARG"),
    ((4, 1024, 0, 1), "ERR_ARG"),
    ((1, 16384, 262145, 0), "ERR_ARG"),   # one input's spectra past 2^32 values
])
def test_engine_shape_rejections(corr, shape, status):
    """Shapes the engine refuses before touching a device."""
    lib = corr.load()
    sh = corr.corrx_shape(*shape)
    e = ctypes.c_void_p(0x1)
    assert lib.corrx_engine_create(ctypes.byref(sh), ctypes.byref(e)) == getattr(corr, status)
    assert not e.value
    assert lib.corr_last_error()
