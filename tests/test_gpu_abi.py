"""The drop-in ABI's contract on a device, re-expressing the reference's
corr-cuda/tests/abi_smoke.c (staging layout, mask lifecycle, row lifecycle
errors, zero in -> zero out, reuse), corr-cuda/tests/async_notify.c (the
non-blocking halves and their callback counting) and corr-rs/tests/abi.rs
(repeated create/destroy)."""

import ctypes
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_cfg(corr, **kw):
    # abi_smoke.c:47-59
    base = dict(n_station=4, n_freq=4096, frames_per_slot=16, chunk_frames=8, block_slots=3,
                rows_per_fetch=2)
    base.update(kw)
    return corr.make_config(**base)


def mark_all_valid(c, slot):
    for g in range(c.n_station):
        c.slot_mask(g, slot)[:] = 0xFF


def status(corr, fn, *args):
    return getattr(corr.load(), fn)(*args)


def test_device_dependent_rejections(corr):
    lib = corr.load()
    for kw, want in [
        (dict(device_id=64), corr.ERR_ARG),                    # abi_smoke.c:136-140
        (dict(devices=[0, 64]), corr.ERR_ARG),
        (dict(devices=[0, 0]), corr.ERR_ARG),                  # a card driven twice
        (dict(n_station=45), corr.ERR_ARG),                    # 1035 > 1024 baselines
        (dict(n_freq=1050624), corr.ERR_NOT_IMPL),             # 513*2048: above 2^20
    ]:
        cfg = small_cfg(corr, **kw)
        h = ctypes.c_void_p(0x1)
        s = lib.corr_create(ctypes.byref(cfg), ctypes.byref(h))
        assert s == want, (kw, corr.STATUS_NAMES[s], corr.last_error())
        assert not h.value
        assert corr.last_error()


@pytest.mark.parametrize("sample_bits", [8, 4])
def test_staging_layout_and_mask_lifecycle(corr, sample_bits):
    """abi_smoke.c:157-226 strides; with 4-bit staging (corrx_create_packed) a
    frame takes n_freq bytes and everything else keeps its meaning."""
    cfg = small_cfg(corr)
    with corr.Correlator(cfg, sample_bits=sample_bits) as c:
        frame_stride = 2 * cfg.n_freq * sample_bits // 8
        slot_stride = frame_stride * cfg.frames_per_slot
        samples = slot_stride * cfg.block_slots
        mps = (cfg.frames_per_slot + 7) // 8
        station_stride = samples + mps * cfg.block_slots
        base0 = ctypes.addressof(c.staging(0).base.contents)
        for g in range(cfg.n_station):
            s = c.staging(g)
            assert (s.frame_stride, s.slot_stride, s.validity_bytes_per_slot) == \
                (frame_stride, slot_stride, mps)
            assert s.validity_offset == samples and s.total_bytes == station_stride
            assert ctypes.addressof(s.base.contents) == base0 + g * station_stride
        out = corr.corr_staging()
        assert status(corr, "corr_staging_info", c.h, cfg.n_station, ctypes.byref(out)) == \
            corr.ERR_ARG
        v = c.station_view(0)
        off = 2 * slot_stride + 7 * frame_stride + 4096
        v[off], v[off + 1] = -128, 127
        assert v[off] == -128 and v[off + 1] == 127
        v[off] = v[off + 1] = 0
        for g in range(cfg.n_station):
            assert not c.station_view(g).any(), "staging not zero initialised"
        # mask: starts empty, is set by the writer, cleared by slot_free
        m = c.slot_mask(0, 0)
        m[:] = 0xAA
        c.row_accumulate(0, cfg.frames_per_slot)
        c.row_slot_free(0)
        for g in range(cfg.n_station):
            assert not c.station_view(g)[samples:].any(), "slot_free must clear the mask"


def test_row_lifecycle(corr):
    cfg = small_cfg(corr)
    lib = corr.load()
    with corr.Correlator(cfg) as c:
        h = c.h
        out = np.empty(c.fetch_bytes() // 8, dtype=np.complex64)
        rows = ctypes.c_int32(-1)
        assert lib.corr_row_slot_free(h, 1) == corr.ERR_STATE
        assert lib.corr_row_accumulate(h, -1, 1) == corr.ERR_ARG
        assert lib.corr_rows_fetch(h, out.ctypes.data, out.nbytes, None) == corr.ERR_ARG
        assert lib.corr_rows_fetch(h, out.ctypes.data, 8, ctypes.byref(rows)) == corr.ERR_ARG
        assert c.rows_pending() == 0
        assert c.rows_fetch(out)[0] == 0
        assert lib.corr_row_accumulate(h, 0, 0) == corr.ERR_ARG
        assert lib.corr_row_accumulate(h, 0, cfg.frames_per_slot + 1) == corr.ERR_ARG
        assert lib.corr_row_end(h, 0) == corr.ERR_ARG
        # zero in -> zero out, with every frame marked present
        for seq in range(cfg.block_slots):
            slot = seq % cfg.block_slots
            mark_all_valid(c, slot)
            c.row_accumulate(seq, cfg.frames_per_slot)
            assert lib.corr_row_accumulate(h, seq, cfg.frames_per_slot) == corr.ERR_BUSY
            c.row_slot_free(seq)
            assert lib.corr_row_slot_free(h, seq) == corr.ERR_STATE
            if slot % 2 == 1:
                c.row_end(cfg.frames_per_slot)
        assert c.rows_pending() == 1
        c.row_end(cfg.frames_per_slot)
        assert lib.corr_row_end(h, cfg.frames_per_slot) == corr.ERR_BUSY
        c.rows_wait()
        n, out = c.rows_fetch(out)
        assert n == 2
        assert not np.any(out.view(np.float32)), "zero input produced nonzero output"
        # fetching clears the rows, so the pipeline cycles
        for i in range(cfg.block_slots):
            seq = cfg.block_slots + i
            mark_all_valid(c, seq % cfg.block_slots)
            c.row_accumulate(seq, cfg.frames_per_slot)
            c.row_slot_free(seq)
        c.row_end(cfg.frames_per_slot)
        c.rows_wait()
        assert c.rows_fetch(out)[0] == 1
        # what the split was: one device holding everything
        assert lib.corr_n_devices(h) == 1
        plan = c.placement()[0]
        assert (plan.n_station_resident, plan.first_baseline, plan.n_baseline, plan.contiguous) \
            == (4, 0, 10, 1)
        st = (ctypes.c_int32 * 4)()
        assert lib.corr_device_stations(h, 0, st, 4) == corr.OK and list(st) == [0, 1, 2, 3]
        assert lib.corr_device_stations(h, 0, st, 3) == corr.ERR_ARG
        bl = (ctypes.c_int32 * 10)()
        assert lib.corr_device_baselines(h, 0, bl, 10) == corr.OK and list(bl) == list(range(10))
        assert lib.corr_device_plan_get(h, 1, ctypes.byref(corr.corr_device_plan())) == \
            corr.ERR_ARG
        stats = c.device_stats()
        assert stats.transfers == 2 * 2 * cfg.block_slots  # 6 slots of 2 chunks each
        assert stats.bytes_d2h == 3 * c.row_bytes()
        assert stats.bytes_h2d > 0 and stats.ns_accumulate > 0


def test_host_alloc_and_repeated_create(corr):
    lib = corr.load()
    p = ctypes.c_void_p()
    assert lib.corr_host_alloc(1 << 20, ctypes.byref(p)) == corr.OK and p.value
    ctypes.memset(p.value, 0x5A, 1 << 20)
    lib.corr_host_free(p.value)
    for _ in range(5):  # corr-rs/tests/abi.rs:254-261
        with corr.Correlator(small_cfg(corr)):
            pass


class Counter:
    """Counts callbacks fired from the runtime's thread."""

    def __init__(self, corr):
        self.n = 0
        self.lock = threading.Lock()

        def cb(user):
            with self.lock:
                self.n += 1
        self.fn = corr.NOTIFY_FN(cb)

    def wait(self, want, seconds, poll=None):
        t_end = time.time() + seconds
        while time.time() < t_end:
            if self.n >= want:
                return True
            if poll:
                poll()
            time.sleep(0.0005)
        return self.n >= want


def async_cfg(corr):
    # async_notify.c:109-129, with a 32 MiB slot instead of 8 MiB: on this host link
    # an 8 MiB copy finishes in ~0.16 ms, too close to the cost of queueing the
    # callback to show that the submit does not wait for the copy
    return corr.make_config(n_station=2, n_freq=8192, frames_per_slot=1024, chunk_frames=256,
                            block_slots=3, rows_per_fetch=1)


def test_slot_free_submit_finish(corr):
    lib = corr.load()
    cfg = async_cfg(corr)
    with corr.Correlator(cfg) as c:
        cnt = Counter(corr)
        # warm up, then time the blocking release
        mark_all_valid(c, 0)
        c.row_accumulate(0, cfg.frames_per_slot)
        c.row_slot_free(0)
        block_ns = 0
        for seq in range(1, 6):
            mark_all_valid(c, seq % 3)
            c.row_accumulate(seq, cfg.frames_per_slot)
            t0 = time.perf_counter_ns()
            c.row_slot_free(seq)
            block_ns += time.perf_counter_ns() - t0
        submit_ns = 0
        seq = 10
        for i in range(5):
            mark_all_valid(c, seq % 3)
            c.row_accumulate(seq, cfg.frames_per_slot)
            cnt.n = 0
            t0 = time.perf_counter_ns()
            assert lib.corr_row_slot_free_submit(c.h, seq, cnt.fn, None) == corr.OK
            submit_ns += time.perf_counter_ns() - t0
            assert cnt.wait(1, 5.0, poll=c.rows_pending), f"no callback for sequence {seq}"
            assert cnt.n == 1
            assert lib.corr_row_slot_free_finish(c.h, seq) == corr.OK
            seq += 1
        print(f"blocking release {block_ns} ns/5, submit {submit_ns} ns/5")
        assert submit_ns * 4 < block_ns
        for g in range(cfg.n_station):
            assert not c.station_view(g)[c.staging(g).validity_offset:].any()
        assert lib.corr_row_slot_free_finish(c.h, seq - 1) == corr.ERR_STATE
        assert lib.corr_row_slot_free_submit(c.h, seq - 1, corr.NOTIFY_FN(0), None) == \
            corr.ERR_ARG
        # as many outstanding as the ring holds, then many in turn
        cnt.n = 0
        first = seq
        for i in range(cfg.block_slots):
            mark_all_valid(c, seq % 3)
            c.row_accumulate(seq, cfg.frames_per_slot)
            assert lib.corr_row_slot_free_submit(c.h, seq, cnt.fn, None) == corr.OK
            seq += 1
        assert cnt.wait(cfg.block_slots, 10.0)
        for i in range(cfg.block_slots):
            assert lib.corr_row_slot_free_finish(c.h, first + i) == corr.OK
        many = 30
        for i in range(many):
            mark_all_valid(c, seq % 3)
            c.row_accumulate(seq, cfg.frames_per_slot)
            assert lib.corr_row_slot_free_submit(c.h, seq, cnt.fn, None) == corr.OK
            assert cnt.wait(cfg.block_slots + i + 1, 5.0)
            assert lib.corr_row_slot_free_finish(c.h, seq) == corr.OK
            seq += 1
        assert cnt.n == cfg.block_slots + many


def fill_pattern(c):
    # async_notify.c:296-306
    for g in range(c.n_station):
        s = c.staging(g)
        v = c.station_view(g)
        i = np.arange(s.validity_offset, dtype=np.int64)
        v[: s.validity_offset] = ((i * 31 + g * 7) % 9 - 4).astype(np.int8)
        v[s.validity_offset:] = -1


def one_row(c, cfg, first_seq):
    for i in range(cfg.block_slots):
        seq = first_seq + i
        mark_all_valid(c, seq % cfg.block_slots)
        c.row_accumulate(seq, cfg.frames_per_slot)
        c.row_slot_free(seq)
    c.row_end(cfg.frames_per_slot * cfg.block_slots)


def test_rows_wait_and_fetch_submit(corr):
    lib = corr.load()
    cfg = async_cfg(corr)
    with corr.Correlator(cfg) as c:
        fill_pattern(c)
        cnt = Counter(corr)
        one_row(c, cfg, 0)
        assert lib.corr_rows_wait_submit(c.h, cnt.fn, None) == corr.OK
        assert cnt.wait(1, 10.0) and cnt.n == 1
        assert lib.corr_rows_ready(c.h) == corr.OK
        ref = np.empty(c.fetch_bytes() // 8, dtype=np.complex64)
        assert c.rows_fetch(ref)[0] == 1
        assert np.abs(ref).max() > 0
        # the same work fetched without waiting
        one_row(c, cfg, cfg.block_slots)
        out = np.full(c.fetch_bytes() // 8, np.nan, dtype=np.complex64)
        rows = ctypes.c_int32(-1)
        cnt.n = 0
        assert lib.corr_rows_fetch_submit(c.h, out.ctypes.data, out.nbytes, ctypes.byref(rows),
                                          cnt.fn, None) == corr.OK
        assert rows.value == 1  # known at submit
        assert cnt.wait(1, 10.0) and cnt.n == 1
        assert lib.corr_rows_fetch_ready(c.h) == corr.OK
        one_row(c, cfg, 2 * cfg.block_slots)
        ref2 = np.empty_like(ref)
        assert c.rows_fetch(ref2)[0] == 1
        # one device: laps and paths are bit-identical
        # compared as bools first: a failing comparison of large bytes objects
        # would make pytest build a difflib diff of them
        same_ref = ref.tobytes() == ref2.tobytes()
        same_out = out.tobytes() == ref2.tobytes()
        assert same_ref and same_out
        # argument checks (async_notify.c:457-502)
        assert lib.corr_rows_fetch_submit(c.h, out.ctypes.data, 8, ctypes.byref(rows), cnt.fn,
                                          None) == corr.ERR_ARG
        assert lib.corr_rows_fetch_submit(c.h, None, out.nbytes, ctypes.byref(rows), cnt.fn,
                                          None) == corr.ERR_ARG
        assert lib.corr_rows_wait_submit(c.h, corr.NOTIFY_FN(0), None) == corr.ERR_ARG
        # nothing pending: 0 rows, no callback armed
        cnt.n = 0
        assert lib.corr_rows_fetch_submit(c.h, out.ctypes.data, out.nbytes, ctypes.byref(rows),
                                          cnt.fn, None) == corr.OK
        assert rows.value == 0
        time.sleep(0.05)
        assert cnt.n == 0
