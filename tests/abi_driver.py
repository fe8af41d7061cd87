"""Drive the drop-in ABI the way the reference's scheduler and golden test do
(corr-cuda/tests/golden.c:117-213): fill a staging slot, release it before its
next use, accumulate, close rows with the nominal n_cum, fetch the stack once.
"""

import numpy as np

ALL_VALID, HOST_ZEROED, MASK_MARKED = 0, 1, 2
MISSING_EVERY, MISSING_PHASE = 7, 3  # golden.c:97-104


def frame_is_missing(row, frame_in_row):
    return ((row * 3 + frame_in_row) % MISSING_EVERY) == MISSING_PHASE


def drive(corr, cfg, x, rows, n_cum, mode=ALL_VALID, lib=None, missing=None, stats=None, sample_bits=8):
    """x: int8 (rows, N, n_cum, L). Returns complex64 (rows, B, nf) from one fetch.
    `missing(row, frame_in_row)` overrides the missing-frame pattern of the
    HOST_ZEROED / MASK_MARKED modes; `stats`, a list, receives the context's
    device statistics before it is destroyed. sample_bits 4 drives a context with
    4-bit packed staging (x must lie in [-8, 7])."""
    is_missing = missing if missing is not None else frame_is_missing
    fps = cfg.frames_per_slot
    N = cfg.n_station
    if sample_bits == 4:
        x = corr.pack4(x)
    with corr.Correlator(cfg, lib=lib, sample_bits=sample_bits) as c:
        submits = 0
        for r in range(rows):
            for f0 in range(0, n_cum, fps):
                ft = min(fps, n_cum - f0)
                slot = submits % cfg.block_slots
                if submits >= cfg.block_slots:
                    c.row_slot_free(submits)  # golden.c:148-150 (release before fill)
                for g in range(N):
                    frames = c.slot_frames(g, slot)
                    mask = c.slot_mask(g, slot)
                    for f in range(ft):
                        gone = is_missing(r, f0 + f)
                        if mode != ALL_VALID and gone:
                            frames[f, :] = 0 if mode == HOST_ZEROED else 0x5A
                        else:
                            frames[f, :] = x[r, g, f0 + f]
                        if not (mode == MASK_MARKED and gone):
                            mask[f >> 3] |= np.uint8(1 << (f & 7))
                c.row_accumulate(submits, ft)
                submits += 1
            c.row_end(n_cum)
        first_held = max(0, submits - cfg.block_slots)
        for s in range(submits - 1, first_held - 1, -1):
            c.row_slot_free(s)
        c.rows_wait()
        assert c.rows_pending() == rows
        got, out = c.rows_fetch()
        assert got == rows
        if stats is not None:
            stats.append(c.device_stats(0))
        return out.reshape(cfg.rows_per_fetch, c.n_baseline, cfg.n_freq)[:rows].copy()


def valid_matrix(rows, N, n_cum, missing=None):
    """Per (row, station, frame) validity of the MASK_MARKED/HOST_ZEROED runs."""
    is_missing = missing if missing is not None else frame_is_missing
    v = np.ones((rows, N, n_cum), dtype=np.uint8)
    for r in range(rows):
        for f in range(n_cum):
            if is_missing(r, f):
                v[r, :, f] = 0
    return v
