"""bench.py's contract on the CPU: it refuses more than one GPU before touching
one, and the committed rocprofv3 summaries it reads (kernel-trace averages for
roofline.profile, PMC traffic for roofline.traffic) are of the workload it
runs by default and name every kernel its roofline can pick."""

import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.parametrize("form", ["fp16", "fp32", "int8"])
def test_profiles_match_the_default_workload(form):
    import bench
    avg, src = bench.traced_kernels(24, 16384, 2048)
    assert src == os.path.join("profiles", "fx_24in_nf16384_trace.json")
    traffic = bench.measured_traffic(24, 16384, 2048, form)
    for k in (f"fengine_{form}", f"xengine_{form}"):
        assert avg[k] > 0
        assert traffic[k] > 0
    # another geometry has no committed profile
    assert bench.traced_kernels(40, 16384, 2048) == ({}, None)
    assert bench.measured_traffic(24, 16384, 512, form) == {}


def test_valu_issue_profile():
    """The committed F-engine VALU figures are of the default workload and
    plausible: at most about one VALU instruction per CU-cycle (a CU's issue
    rate), about a thousand per wave and frame."""
    import bench
    v, src = bench.valu_issue(24, 16384, 2048)
    assert src == os.path.join("profiles", "fx_24in_nf16384_util.json")
    for form in ("fp16", "fp32", "int8"):
        f = v[f"fengine_{form}"]
        assert 0.5 < f["valu_per_cu_cycle"] <= 1.15
        assert 800 < f["valu_per_wave_and_frame"] < 1300
        assert 1.0 < f["clock_ghz"] < 2.5
    assert bench.valu_issue(40, 16384, 2048) == ({}, None)


def test_traffic_near_the_algorithmic_bytes():
    """The calibrated PMC bytes of each kernel are its algorithmic bytes to
    within 2% (no operand re-read from HBM)."""
    N, nf, T = 24, 16384, 2048
    B = N * (N + 1) // 2
    import bench
    for form in ("fp16", "fp32", "int8"):
        t = bench.measured_traffic(N, nf, T, form)
        f_alg = N * T * 2 * nf + N * T * nf * bench.SPEC_BYTES[form]
        x_alg = bench.SPEC_BYTES[form] * N * nf * T + 2 * bench.INTEG_BYTES[form] * B * nf
        assert abs(t[f"fengine_{form}"] / f_alg - 1) < 0.02
        assert abs(t[f"xengine_{form}"] / x_alg - 1) < 0.02
