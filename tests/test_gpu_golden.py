"""End-to-end numerical parity through the drop-in ABI, re-expressing the
reference's golden test (corr-cuda/tests/golden.c, configs at
corr-cuda/CMakeLists.txt:161-163; input recipe corr-cuda/tests/gen_golden.py:57-62).

Checks, as the reference does:
  1. relative RMS against the FP64 oracle below 1e-4 (golden.c:41-48);
  2. autocorrelation |Im| <= 1e-3 max|Re| (golden.c:336-349, 398-403);
  3. host-zeroed and mask-marked missing frames give byte-identical rows, which
     differ from the all-valid run (golden.c:351-396).
Also checked against numpy's FP64 rfft (an independent library), and, for
chunks in which every frame is missing, that they are skipped (no copy) with
the same bytes as a run that transforms them as zeros.
"""

import numpy as np
import pytest

from abi_driver import ALL_VALID, HOST_ZEROED, MASK_MARKED, drive, valid_matrix
from conftest import rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-4


def golden_input(rows, N, n_cum, L, seed=20260914):
    rng = np.random.default_rng(seed)
    x = rng.integers(-4, 5, size=(rows, N, n_cum, L)).astype(np.int8)
    x[0, 0, 0, 0] = -128
    x[0, 0, 0, 1] = 127
    return x


@pytest.mark.parametrize("N,nf,n_cum,fps,rows", [
    (4, 2048, 300, 128, 2),   # golden_partial_slot
    (15, 4096, 200, 64, 2),   # golden_15st
    (4, 6144, 150, 64, 2),    # n_freq with an odd factor: 3 x 2048
    (3, 14336, 70, 32, 1),    # 7 x 2048
    (2, 18432, 40, 16, 1),    # 9 x 2048, staged through HBM
    (2, 32768, 20, 8, 1),     # 2 x 16384, staged
    (2, 67584, 12, 4, 1),     # 33 x 2048, staged, radix-33 first stage
])
def test_golden(corr, oracle, N, nf, n_cum, fps, rows):
    print(f"golden N={N} nf={nf} n_cum={n_cum} fps={fps} rows={rows}", flush=True)
    L = 2 * nf
    x = golden_input(rows, N, n_cum, L)
    cfg = corr.make_config(N, nf, fps, min(fps, 128), block_slots=3, rows_per_fetch=rows)
    got = drive(corr, cfg, x, rows, n_cum, ALL_VALID)

    ref = np.stack([oracle.fx_f64(x[r], frames_div=n_cum) for r in range(rows)])
    err = rel_rms(got, ref)
    print(f"  rel RMS vs FP64 oracle = {err:.3e}")
    assert err < TOL

    if N == 4:  # independent FP64 restatement with numpy's FFT (gen_golden.py:68-78)
        Y = np.fft.rfft(x.astype(np.float64), axis=3)[..., :nf]
        pairs = oracle.pairs(N)
        ref_np = np.stack([np.stack([(Y[r, j] * np.conj(Y[r, k])).sum(0) for j, k in pairs])
                           for r in range(rows)]) / n_cum
        assert rel_rms(got, ref_np) < TOL

    pairs = corr.baseline_pairs(N)
    autos = [b for b, (j, k) in enumerate(pairs) if j == k]
    max_re = np.abs(got[:, autos].real).max()
    max_im = np.abs(got[:, autos].imag).max()
    assert max_im <= 1e-3 * max_re

    zeroed = drive(corr, cfg, x, rows, n_cum, HOST_ZEROED)
    masked = drive(corr, cfg, x, rows, n_cum, MASK_MARKED)
    same = zeroed.tobytes() == masked.tobytes()  # a bool: no difflib diff of large bytes on failure
    assert same, "device zeroing differs from host zeroing"
    assert rel_rms(zeroed, got) > 1e-6, "declaring frames missing changed nothing"

    valid = valid_matrix(rows, N, n_cum)
    ref_m = np.stack([oracle.fx_f64(x[r], frames_div=n_cum, valid=valid[r]) for r in range(rows)])
    assert rel_rms(masked, ref_m) < TOL


def _skip_pattern(r, f):
    """Row 0: the whole second slot (frames 32..63) missing on every station.
    Row 1: frames 8..15 (one whole chunk) and 70..75 (part of a chunk)."""
    if r == 0:
        return 32 <= f < 64
    return 8 <= f < 16 or 70 <= f < 76


def test_whole_missing_chunks_skipped(corr, oracle):
    """A chunk in which no station has a present frame is neither copied nor
    transformed (the reference's next step for whole-missing slots,
    docs/DESIGN-zero-fill-decoupling.md:219-231). The rows must equal the
    host-zeroed run byte for byte and the FP64 oracle on the validity pattern,
    and the copied bytes must drop by exactly the five skipped chunks."""
    N, nf, fps, chunk, n_cum, rows = 4, 2048, 32, 8, 96, 2
    L = 2 * nf
    print(f"skip whole-missing chunks N={N} nf={nf} fps={fps} chunk={chunk}", flush=True)
    x = golden_input(rows, N, n_cum, L, seed=77)
    cfg = corr.make_config(N, nf, fps, chunk, block_slots=3, rows_per_fetch=rows)
    zs, ms = [], []
    zeroed = drive(corr, cfg, x, rows, n_cum, HOST_ZEROED, missing=_skip_pattern, stats=zs)
    masked = drive(corr, cfg, x, rows, n_cum, MASK_MARKED, missing=_skip_pattern, stats=ms)
    same = zeroed.tobytes() == masked.tobytes()
    assert same, "skipping differs from transforming zeros"
    valid = valid_matrix(rows, N, n_cum, _skip_pattern)
    ref = np.stack([oracle.fx_f64(x[r], frames_div=n_cum, valid=valid[r]) for r in range(rows)])
    assert rel_rms(masked, ref) < TOL
    mask_per_slot = (fps + 7) // 8
    per_chunk = N * (chunk * L + mask_per_slot)
    print(f"  copied bytes: host-zeroed {zs[0].bytes_h2d}, mask-marked {ms[0].bytes_h2d}", flush=True)
    assert zs[0].bytes_h2d - ms[0].bytes_h2d == 5 * per_chunk
    assert zs[0].transfers - ms[0].transfers == 5


@pytest.mark.parametrize("N,nf,n_cum,fps,chunk", [
    (4, 2048, 150, 64, 32),
    (3, 6144, 70, 32, 32),    # the mixed-radix transform
    (5, 16384, 40, 16, 16),   # the half-buffer transform
])
def test_packed_staging(corr, oracle, N, nf, n_cum, fps, chunk):
    """4-bit staging (corrx_create_packed, the reference's proposed next step,
    HANDOFF.md:1437-1439): samples in [-8, 7] give rows byte-identical to the
    8-bit context on the same samples (each chunk is unpacked to int8 on the
    GPU and transformed by the same F-engine), with exactly half the sample
    bytes copied; missing frames behave as in the 8-bit path."""
    rows = 2
    L = 2 * nf
    print(f"packed staging N={N} nf={nf} n_cum={n_cum} fps={fps} chunk={chunk}", flush=True)
    rng = np.random.default_rng(nf + N)
    x = rng.integers(-8, 8, size=(rows, N, n_cum, L)).astype(np.int8)
    x[0, 0, 0, :4] = [-8, 7, 7, -8]
    cfg = corr.make_config(N, nf, fps, chunk, block_slots=3, rows_per_fetch=rows)
    s8, s4 = [], []
    got8 = drive(corr, cfg, x, rows, n_cum, ALL_VALID, stats=s8)
    got4 = drive(corr, cfg, x, rows, n_cum, ALL_VALID, stats=s4, sample_bits=4)
    same = got4.tobytes() == got8.tobytes()
    assert same, f"4-bit staging changed the rows (max |difference| {np.max(np.abs(got4 - got8)):.3e})"
    ref = np.stack([oracle.fx_f64(x[r], frames_div=n_cum) for r in range(rows)])
    err = rel_rms(got4, ref)
    print(f"  rel RMS vs FP64 oracle = {err:.3e}", flush=True)
    assert err < TOL
    mask_bytes = s8[0].transfers * N * ((fps + 7) // 8)
    assert s4[0].transfers == s8[0].transfers
    assert 2 * (s4[0].bytes_h2d - mask_bytes) == s8[0].bytes_h2d - mask_bytes
    masked4 = drive(corr, cfg, x, rows, n_cum, MASK_MARKED, sample_bits=4)
    masked8 = drive(corr, cfg, x, rows, n_cum, MASK_MARKED)
    same = masked4.tobytes() == masked8.tobytes()
    assert same, "4-bit staging changed the mask-marked rows"
    valid = valid_matrix(rows, N, n_cum)
    ref_m = np.stack([oracle.fx_f64(x[r], frames_div=n_cum, valid=valid[r]) for r in range(rows)])
    assert rel_rms(masked4, ref_m) < TOL
