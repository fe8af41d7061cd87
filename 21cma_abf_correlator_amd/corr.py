"""Host-side binding of libcorr_cuda (the MI355X build) through its C ABI.

`Correlator` mirrors the reference's safe Rust wrapper (corr-rs/src/gpu.rs:32-333)
over the same 20 calls its FFI binds (corr-rs/src/ffi.rs:157-219), plus the
asynchronous halves the C tests use (corr-cuda/include/corr_cuda.h:399-520).
`Engine` binds the engine-level entry points of include/corrx.h, which take
device pointers (integers, e.g. hip.Buffer.ptr) and a HIP stream handle.

There is no CPU fallback: if the library is missing this module raises.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcorr_cuda.so")

ABI_VERSION = 7
MAX_DEVICES = 8
SPLIT_BASELINE_RANGE, SPLIT_PAIR_COVERING, SPLIT_TIME_SLICE = 0, 1, 2
FORM_FP32, FORM_FP16, FORM_INT8 = 0, 1, 2

STATUS_NAMES = {
    0: "CORR_OK", 1: "CORR_ERR_ARG", 2: "CORR_ERR_ABI", 3: "CORR_ERR_CUDA",
    4: "CORR_ERR_CUFFT", 5: "CORR_ERR_STATE", 6: "CORR_ERR_BUSY", 7: "CORR_ERR_NOMEM",
    8: "CORR_ERR_NOT_IMPL", 9: "CORR_ERR_INTERNAL",
}
OK, ERR_ARG, ERR_ABI, ERR_CUDA, ERR_CUFFT, ERR_STATE, ERR_BUSY, ERR_NOMEM, ERR_NOT_IMPL, \
    ERR_INTERNAL = range(10)


class corr_config(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_uint32),
        ("struct_size", ctypes.c_uint32),
        ("device_id", ctypes.c_int32),
        ("n_station", ctypes.c_int32),
        ("n_freq", ctypes.c_int32),
        ("frames_per_slot", ctypes.c_int32),
        ("chunk_frames", ctypes.c_int32),
        ("block_slots", ctypes.c_int32),
        ("rows_per_fetch", ctypes.c_int32),
        ("n_devices", ctypes.c_int32),
        ("devices", ctypes.c_int32 * MAX_DEVICES),
        ("split_mode", ctypes.c_int32),
    ]


class corr_device_plan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("device_id", "n_station_resident", "first_baseline", "n_baseline", "contiguous")]


class corr_device_stats(ctypes.Structure):
    _fields_ = [("device_id", ctypes.c_int32)] + [
        (n, ctypes.c_uint64) for n in ("bytes_h2d", "bytes_d2h", "transfers", "ns_accumulate",
                                       "ns_slot_free", "ns_row_end", "ns_rows_wait", "ns_fetch")]


class corr_staging(ctypes.Structure):
    _fields_ = [
        ("base", ctypes.POINTER(ctypes.c_int8)),
        ("frame_stride", ctypes.c_size_t),
        ("slot_stride", ctypes.c_size_t),
        ("total_bytes", ctypes.c_size_t),
        ("validity_offset", ctypes.c_size_t),
        ("validity_bytes_per_slot", ctypes.c_size_t),
    ]


class corrx_shape(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("n_station", "n_freq", "chunk_frames", "form")]


BATCH_ZERO = 1
BATCH_DUMP = 2


class corrx_batch_item(ctypes.Structure):
    _fields_ = [
        ("d_in", ctypes.c_void_p), ("station_pitch", ctypes.c_size_t),
        ("d_mask", ctypes.c_void_p), ("mask_pitch", ctypes.c_size_t),
        ("mask_bit0", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("frames", ctypes.c_int64), ("d_acc", ctypes.c_void_p),
        ("dump_frames", ctypes.c_uint64), ("d_out", ctypes.c_void_p),
    ]


NOTIFY_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)

# Every symbol include/corr_cuda.h and include/corrx.h declare.
ABI_SYMBOLS = (
    "corr_abi_banner", "corr_last_error", "corr_create", "corr_destroy", "corr_staging_info",
    "corr_row_accumulate", "corr_row_slot_free", "corr_row_slot_free_submit",
    "corr_row_slot_free_finish", "corr_row_end", "corr_rows_wait", "corr_rows_wait_submit",
    "corr_rows_ready", "corr_rows_pending", "corr_rows_fetch", "corr_rows_fetch_submit",
    "corr_rows_fetch_ready", "corr_n_baseline", "corr_baseline_pairs", "corr_n_devices",
    "corr_device_plan_get", "corr_device_stations", "corr_device_baselines",
    "corr_device_stats_get", "corr_host_alloc", "corr_host_free",
)
ENGINE_SYMBOLS = (
    "corrx_engine_create", "corrx_engine_destroy", "corrx_fengine", "corrx_xengine",
    "corrx_integrate", "corrx_batch_create", "corrx_batch_destroy", "corrx_batch_run", "corrx_scale", "corrx_spectra", "corrx_integrator_bytes", "corrx_dump", "corrx_calibrate",
    "corrx_set_gains", "corrx_gains", "corrx_set_xtile", "corrx_get_xtile", "corrx_create_packed",
    "corrx_correlate_noalloc", "corrx_set_xstep", "corrx_get_xstep",
)

_lib = None


def _mapped(pattern):
    """Distinct files whose path contains `pattern` mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f
                           if pattern in ln and ln.split()[-1].startswith("/")})
    except OSError:
        return []


def load():
    """Load libcorr_cuda.so and declare every signature. Raises if it is not built,
    or if the process already runs another HIP runtime (e.g. one bundled with a
    Python wheel): the library's code objects are built for the runtime it links,
    and two runtimes in one process do not share devices, memory or streams."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    before = _mapped("libamdhip64")
    if before and not _mapped(os.path.basename(LIB_PATH)):
        raise RuntimeError(
            f"another HIP runtime is already loaded in this process ({before[0]}); load "
            f"libcorr_cuda before any library that brings its own HIP runtime, and do not "
            f"share device memory or streams with it")
    L = ctypes.CDLL(LIB_PATH)
    after = _mapped("libamdhip64")
    if len(after) != 1:
        raise RuntimeError(f"expected one HIP runtime after loading libcorr_cuda, found {after}")
    _lib = declare(L)
    return _lib


def declare(L):
    """Declare the argument and result types of every exported call on `L`."""
    vp, i32, i64, u64, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, \
        ctypes.c_size_t
    st = ctypes.c_int
    sig = {
        "corr_abi_banner": ([], ctypes.c_char_p),
        "corr_last_error": ([], ctypes.c_char_p),
        "corr_create": ([ctypes.POINTER(corr_config), ctypes.POINTER(vp)], st),
        "corr_destroy": ([vp], st),
        "corr_staging_info": ([vp, i32, ctypes.POINTER(corr_staging)], st),
        "corr_row_accumulate": ([vp, i64, i32], st),
        "corr_row_slot_free": ([vp, i64], st),
        "corr_row_slot_free_submit": ([vp, i64, NOTIFY_FN, vp], st),
        "corr_row_slot_free_finish": ([vp, i64], st),
        "corr_row_end": ([vp, u64], st),
        "corr_rows_wait": ([vp], st),
        "corr_rows_wait_submit": ([vp, NOTIFY_FN, vp], st),
        "corr_rows_ready": ([vp], st),
        "corr_rows_pending": ([vp, ctypes.POINTER(i32)], st),
        "corr_rows_fetch": ([vp, vp, sz, ctypes.POINTER(i32)], st),
        "corr_rows_fetch_submit": ([vp, vp, sz, ctypes.POINTER(i32), NOTIFY_FN, vp], st),
        "corr_rows_fetch_ready": ([vp], st),
        "corr_n_baseline": ([i32], i32),
        "corr_baseline_pairs": ([i32, ctypes.POINTER(i32), i32], st),
        "corr_n_devices": ([vp], i32),
        "corr_device_plan_get": ([vp, i32, ctypes.POINTER(corr_device_plan)], st),
        "corr_device_stations": ([vp, i32, ctypes.POINTER(i32), i32], st),
        "corr_device_baselines": ([vp, i32, ctypes.POINTER(i32), i32], st),
        "corr_device_stats_get": ([vp, i32, ctypes.POINTER(corr_device_stats)], st),
        "corr_host_alloc": ([sz, ctypes.POINTER(vp)], st),
        "corr_host_free": ([vp], None),
        "corrx_engine_create": ([ctypes.POINTER(corrx_shape), ctypes.POINTER(vp)], st),
        "corrx_engine_destroy": ([vp], st),
        "corrx_fengine": ([vp, vp, sz, vp, sz, i32, i32, vp], st),
        "corrx_xengine": ([vp, vp, i32, vp], st),
        "corrx_integrate": ([vp, vp, sz, vp, sz, i32, ctypes.c_int64, vp, vp], st),
        "corrx_batch_create": ([ctypes.POINTER(vp), i32, ctypes.POINTER(vp)], st),
        "corrx_create_packed": ([ctypes.POINTER(corr_config), i32, ctypes.POINTER(vp)], st),
        "corrx_batch_destroy": ([vp], st),
        "corrx_batch_run": ([vp, ctypes.POINTER(corrx_batch_item), i32, vp], st),
        "corrx_scale": ([vp, sz, u64, vp], st),
        "corrx_spectra": ([vp, ctypes.POINTER(vp), ctypes.POINTER(sz)], st),
        "corrx_integrator_bytes": ([vp, ctypes.POINTER(sz)], st),
        "corrx_dump": ([vp, vp, u64, vp, vp], st),
        "corrx_calibrate": ([vp, vp, sz, vp, sz, i32, i32, vp], st),
        "corrx_set_gains": ([vp, vp, vp], st),
        "corrx_gains": ([vp, ctypes.POINTER(vp)], st),
        "corrx_set_xtile": ([vp, i32, i32], st),
        "corrx_get_xtile": ([vp, ctypes.POINTER(i32), ctypes.POINTER(i32)], st),
        "corrx_correlate_noalloc": ([i32, i32, vp, sz, vp, sz, i32, i64, u64, vp, vp], st),
        "corrx_set_xstep": ([vp, i32], st),
        "corrx_get_xstep": ([vp, ctypes.POINTER(i32)], st),
    }
    for name, (args, res) in sig.items():
        # the CPU build of the library (tests/cpu_emu) has no static device
        # scratch, hence no corrx_correlate_noalloc; the GPU build exports all
        fn = getattr(L, name, None)
        if fn is None and name == "corrx_correlate_noalloc":
            continue
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


class CorrError(RuntimeError):
    def __init__(self, status, what, lib=None):
        self.status = status
        msg = (lib or load()).corr_last_error().decode(errors="replace")
        super().__init__(f"{what}: {STATUS_NAMES.get(status, status)}: {msg}")


def check(status, what, lib=None):
    if status != OK:
        raise CorrError(status, what, lib)


def abi_banner():
    return load().corr_abi_banner().decode()


def last_error():
    return load().corr_last_error().decode(errors="replace")


def n_baseline(n_station):
    return load().corr_n_baseline(n_station)


def baseline_pairs(n_station):
    """[(j, k)] in the library's numbering (reference corr_cuda.cpp:1515-1534)."""
    B = n_baseline(n_station)
    if B < 0:
        raise ValueError("n_station must be >= 1")
    buf = (ctypes.c_int32 * (2 * B))()
    check(load().corr_baseline_pairs(n_station, buf, 2 * B), "corr_baseline_pairs")
    return [(buf[2 * i], buf[2 * i + 1]) for i in range(B)]


def make_config(n_station, n_freq, frames_per_slot, chunk_frames, block_slots=3,
                rows_per_fetch=1, device_id=0, devices=None, split_mode=SPLIT_BASELINE_RANGE):
    c = corr_config()
    c.abi_version = ABI_VERSION
    c.struct_size = ctypes.sizeof(corr_config)
    c.device_id = device_id
    c.n_station = n_station
    c.n_freq = n_freq
    c.frames_per_slot = frames_per_slot
    c.chunk_frames = chunk_frames
    c.block_slots = block_slots
    c.rows_per_fetch = rows_per_fetch
    c.split_mode = split_mode
    if devices:
        c.n_devices = len(devices)
        for i, d in enumerate(devices):
            c.devices[i] = d
    return c


def create_raw(cfg, lib=None, sample_bits=8):
    """corr_create (sample_bits 8) or corrx_create_packed; returns (status,
    handle or None) without raising."""
    h = ctypes.c_void_p()
    L = lib or load()
    if sample_bits == 8:
        s = L.corr_create(ctypes.byref(cfg), ctypes.byref(h))
    else:
        s = L.corrx_create_packed(ctypes.byref(cfg), sample_bits, ctypes.byref(h))
    return s, (h.value if h.value else None)


def pack4(x):
    """int8 samples in [-8, 7], last axis even -> the 4-bit staging bytes of
    corrx_create_packed: sample 2m in bits 0-3, sample 2m+1 in bits 4-7 of byte m."""
    x = np.asarray(x)
    if x.min(initial=0) < -8 or x.max(initial=0) > 7:
        raise ValueError("4-bit samples must lie in [-8, 7]")
    u = (x.astype(np.int16) & 0xF).astype(np.uint8)
    return (u[..., 0::2] | (u[..., 1::2] << 4)).astype(np.uint8).view(np.int8)


class Correlator:
    """One correlator context (reference corr-rs/src/gpu.rs:32-333). `lib` is the
    loaded library to drive (default: this package's libcorr_cuda.so);
    `sample_bits` 4 makes a context with 4-bit packed staging (frames of
    n_freq bytes, see pack4)."""

    def __init__(self, cfg, lib=None, sample_bits=8):
        self.cfg = cfg
        self.sample_bits = sample_bits
        self._lib = lib or load()
        s, h = create_raw(cfg, self._lib, sample_bits)
        check(s, "corr_create" if sample_bits == 8 else "corrx_create_packed", self._lib)
        self.h = h
        self._staging = [self.staging(i) for i in range(cfg.n_station)]

    def _check(self, status, what):
        check(status, what, self._lib)

    def close(self):
        if self.h is not None:
            s = self._lib.corr_destroy(self.h)
            self.h = None
            self._check(s, "corr_destroy")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- geometry -----------------------------------------------------------
    @property
    def n_station(self):
        return self.cfg.n_station

    @property
    def n_freq(self):
        return self.cfg.n_freq

    @property
    def n_baseline(self):
        return n_baseline(self.cfg.n_station)

    def row_bytes(self):
        return self.n_baseline * self.cfg.n_freq * 8

    def fetch_bytes(self):
        return self.row_bytes() * self.cfg.rows_per_fetch

    def staging(self, station):
        out = corr_staging()
        self._check(self._lib.corr_staging_info(self.h, station, ctypes.byref(out)), "corr_staging_info")
        return out

    def station_view(self, station):
        """The station's whole staging extent as a writable int8 numpy array."""
        s = self._staging[station]
        return np.ctypeslib.as_array(s.base, shape=(s.total_bytes,))

    def slot_frames(self, station, slot):
        """(frames_per_slot, frame_stride) int8 view of one staging slot:
        2*n_freq samples per frame, or n_freq bytes of packed 4-bit samples."""
        s = self._staging[station]
        v = self.station_view(station)
        a = v[slot * s.slot_stride: (slot + 1) * s.slot_stride]
        return a.reshape(self.cfg.frames_per_slot, s.frame_stride)

    def slot_mask(self, station, slot):
        s = self._staging[station]
        v = self.station_view(station).view(np.uint8)
        o = s.validity_offset + slot * s.validity_bytes_per_slot
        return v[o: o + s.validity_bytes_per_slot]

    # --- row lifecycle ------------------------------------------------------
    def row_accumulate(self, slot_seq, frames):
        self._check(self._lib.corr_row_accumulate(self.h, slot_seq, frames), "corr_row_accumulate")

    def row_slot_free(self, slot_seq):
        self._check(self._lib.corr_row_slot_free(self.h, slot_seq), "corr_row_slot_free")

    def row_end(self, frames):
        self._check(self._lib.corr_row_end(self.h, frames), "corr_row_end")

    def rows_wait(self):
        self._check(self._lib.corr_rows_wait(self.h), "corr_rows_wait")

    def rows_pending(self):
        p = ctypes.c_int32(-1)
        self._check(self._lib.corr_rows_pending(self.h, ctypes.byref(p)), "corr_rows_pending")
        return p.value

    def rows_fetch(self, out=None):
        """Blocking fetch into `out` (complex64 array of fetch_bytes()); returns (rows, out)."""
        if out is None:
            out = np.empty(self.fetch_bytes() // 8, dtype=np.complex64)
        rows = ctypes.c_int32(-1)
        self._check(self._lib.corr_rows_fetch(self.h, out.ctypes.data, out.nbytes, ctypes.byref(rows)),
              "corr_rows_fetch")
        return rows.value, out

    def device_stats(self, device_index=0):
        s = corr_device_stats()
        self._check(self._lib.corr_device_stats_get(self.h, device_index, ctypes.byref(s)),
              "corr_device_stats_get")
        return s

    def placement(self):
        n = self._lib.corr_n_devices(self.h)
        plans = []
        for d in range(n):
            p = corr_device_plan()
            self._check(self._lib.corr_device_plan_get(self.h, d, ctypes.byref(p)), "corr_device_plan_get")
            plans.append(p)
        return plans


class Engine:
    """Engine-level stages on device buffers (include/corrx.h)."""

    def __init__(self, n_station, n_freq, chunk_frames, form=FORM_FP32, lib=None):
        self._lib = lib if lib is not None else load()
        sh = corrx_shape(n_station, n_freq, chunk_frames, form)
        h = ctypes.c_void_p()
        check(self._lib.corrx_engine_create(ctypes.byref(sh), ctypes.byref(h)),
              "corrx_engine_create", self._lib)
        self.h = h.value
        self.n_station, self.n_freq, self.chunk_frames, self.form = \
            n_station, n_freq, chunk_frames, form

    def close(self):
        if self.h is not None:
            self._lib.corrx_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fengine(self, d_in, station_pitch, frames, d_mask=None, mask_pitch=0, mask_bit0=0,
                stream=None):
        check(self._lib.corrx_fengine(self.h, d_in, station_pitch, d_mask, mask_pitch, mask_bit0,
                                      frames, stream), "corrx_fengine", self._lib)

    def xengine(self, d_acc, frames, stream=None):
        check(self._lib.corrx_xengine(self.h, d_acc, frames, stream), "corrx_xengine", self._lib)

    def integrate(self, d_in, station_pitch, frames, d_acc, d_mask=None, mask_pitch=0, mask_bit0=0,
                  stream=None):
        """F- and X-engine over `frames` frames (any count) into d_acc, the
        X-engine of each run of chunk_frames overlapping the next F-engine."""
        check(self._lib.corrx_integrate(self.h, d_in, station_pitch, d_mask, mask_pitch, mask_bit0,
                                        frames, d_acc, stream), "corrx_integrate", self._lib)

    def spectra(self):
        p = ctypes.c_void_p()
        s = ctypes.c_size_t()
        check(self._lib.corrx_spectra(self.h, ctypes.byref(p), ctypes.byref(s)), "corrx_spectra", self._lib)
        return p.value, s.value

    def integrator_bytes(self):
        """Bytes of one integrator element: 8 (FP32), 16 (FP16), 8 (INT8)."""
        n = ctypes.c_size_t()
        check(self._lib.corrx_integrator_bytes(self.h, ctypes.byref(n)), "corrx_integrator_bytes", self._lib)
        return n.value

    def dump(self, d_acc, frames, d_out, stream=None):
        check(self._lib.corrx_dump(self.h, d_acc, frames, d_out, stream), "corrx_dump", self._lib)

    def calibrate(self, d_in, station_pitch, frames, d_mask=None, mask_pitch=0, mask_bit0=0,
                  stream=None):
        check(self._lib.corrx_calibrate(self.h, d_in, station_pitch, d_mask, mask_pitch,
                                        mask_bit0, frames, stream), "corrx_calibrate", self._lib)

    def set_gains(self, d_gains, stream=None):
        check(self._lib.corrx_set_gains(self.h, d_gains, stream), "corrx_set_gains", self._lib)

    def gains(self):
        p = ctypes.c_void_p()
        check(self._lib.corrx_gains(self.h, ctypes.byref(p)), "corrx_gains", self._lib)
        return p.value

    def xtile(self):
        """(channels per MFMA X-engine workgroup, channel tile of the spectrum order)."""
        c = ctypes.c_int32()
        t = ctypes.c_int32()
        check(self._lib.corrx_get_xtile(self.h, ctypes.byref(c), ctypes.byref(t)), "corrx_get_xtile", self._lib)
        return c.value, t.value

    def set_xtile(self, channels, spectra_tile):
        check(self._lib.corrx_set_xtile(self.h, channels, spectra_tile), "corrx_set_xtile", self._lib)

    def xstep(self):
        """Frames staged per step of the matrix-core X-engine (time-K; 0 for
        the FP32 vector X-engine)."""
        n = ctypes.c_int32()
        check(self._lib.corrx_get_xstep(self.h, ctypes.byref(n)), "corrx_get_xstep", self._lib)
        return n.value

    def set_xstep(self, frames_per_step):
        check(self._lib.corrx_set_xstep(self.h, frames_per_step), "corrx_set_xstep", self._lib)


class Batch:
    """One dump of several engines (independent sub-arrays) at once
    (corrx_batch_*): per engine, zero + integrate + dump on a branch stream of
    its own; optionally captured into a hipGraph and replayed."""

    def __init__(self, engines, lib=None):
        self._lib = lib if lib is not None else load()
        self.engines = list(engines)
        arr = (ctypes.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = ctypes.c_void_p()
        check(self._lib.corrx_batch_create(arr, len(self.engines), ctypes.byref(h)), "corrx_batch_create", self._lib)
        self.h = h.value

    def close(self):
        if self.h is not None:
            self._lib.corrx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def item(d_in, station_pitch, frames, d_acc, d_out=None, dump_frames=0, d_mask=None,
             mask_pitch=0, mask_bit0=0, zero=True):
        flags = (BATCH_ZERO if zero else 0) | (BATCH_DUMP if d_out is not None else 0)
        return corrx_batch_item(d_in, station_pitch, d_mask, mask_pitch, mask_bit0, flags, frames,
                                d_acc, dump_frames, d_out)

    def run(self, items, graph=True, stream=None):
        arr = (corrx_batch_item * len(items))(*items)
        check(self._lib.corrx_batch_run(self.h, arr, 1 if graph else 0, stream), "corrx_batch_run", self._lib)


def scale(d_acc, n_complex, frames, stream=None):
    check(load().corrx_scale(d_acc, n_complex, frames, stream), "corrx_scale")
