"""The drop-in ABI driven from C on the GPU: tests/c_abi/abi_check.c, linked
against lib/libcorr_cuda.so, run in its full mode (numbering, config
rejections, staging geometry, slot and row state rules, the asynchronous
notifications, device statistics, zero in -> zero out, and impulse inputs
against their closed-form visibilities). The reference keeps this contract in
C programs of its own (corr-cuda/tests/abi_smoke.c, async_notify.c)."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CDIR = os.path.join(ROOT, "tests", "c_abi")


def test_c_harness_on_device(corr):
    exe = os.path.join(CDIR, "_build", "abi_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", CDIR], check=True)
    print("C harness: abi_check run", flush=True)
    r = subprocess.run([exe, "run"], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
