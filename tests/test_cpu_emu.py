"""The F- and X-engine kernel bodies (csrc/fft_body.hpp, csrc/xengine_body.hpp)
compiled for the CPU and run one workgroup at a time (a host thread per lane, a
barrier object for the workgroup barrier) under AddressSanitizer and UBSan,
with the engine's own launch plans: every index stays in bounds and the results
match FP64 (tests/cpu_emu/)."""

import os
import subprocess

import pytest

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_emu")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", EMU], check=True)
    return os.path.join(EMU, "_build")


@pytest.mark.parametrize("prog", ["fengine_emu", "xengine_emu"])
def test_kernel_body_emulation(built, prog):
    r = subprocess.run([os.path.join(built, prog)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout
