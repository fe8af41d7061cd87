"""The F- and X-engine kernel bodies (csrc/fft_body.hpp, csrc/xengine_body.hpp)
compiled for the CPU and run one workgroup at a time (a host thread per lane, a
barrier object for the workgroup barrier) under AddressSanitizer and UBSan,
with the engine's own launch plans: every index stays in bounds and the results
match FP64 (tests/cpu_emu/)."""

import os
import subprocess
import sys

import pytest

EMU = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_emu")


@pytest.fixture(scope="module")
def built():
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    subprocess.run(["make", "-s", "-j", jobs, "-C", EMU], check=True)
    return os.path.join(EMU, "_build")


@pytest.mark.parametrize("prog,parts", [("fengine_emu", 4), ("xengine_emu", 1), ("xmfma_emu", 6)])
def test_kernel_body_emulation(built, prog, parts):
    """`parts` processes at once, each running every parts-th check."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    args = [[]] if parts == 1 else [[str(i), str(parts)] for i in range(parts)]
    procs = [subprocess.Popen([os.path.join(built, prog)] + a, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                              text=True, env=env) for a in args]
    for p in procs:
        out, err = p.communicate(timeout=600)
        print(out)
        assert p.returncode == 0, out + err
        assert "OK" in out


def test_generated_code_has_no_clobbered_upper_halves():
    """The gfx950 code of every kernel, compiled with the library's own flags,
    passes the register-hazard check of tests/isa_lint.py."""
    import isa_lint
    csrc = os.path.join(os.path.dirname(EMU), "..", "21cma_abf_correlator_amd", "csrc")
    subprocess.run(["make", "-s", "-C", csrc, "asm"], check=True)
    text = open(os.path.join(csrc, "..", "lib", "obj", "engine_kernels.s")).read()
    kernels = isa_lint.parse_kernels(text)
    assert any("fengine_fp32_kernel" in k for k in kernels)
    assert any("xengine_fp32_kernel" in k for k in kernels)
    findings = isa_lint.lint_text(text)
    assert not findings, findings


def test_fengine_stores_take_scalar_descriptors():
    """The F-engine's spectrum stores and gain loads are buffer accesses whose
    base is the same in every lane. Where the compiler cannot prove that (a
    base merged after the branches on the bin-0 lane), it wraps each access in
    a loop over the wave's distinct descriptors: about 12 instructions per
    access on the VALU-bound kernel, and the loop serialises the emission's
    loads. No F-engine kernel may contain one."""
    import isa_lint
    csrc = os.path.join(os.path.dirname(EMU), "..", "21cma_abf_correlator_amd", "csrc")
    subprocess.run(["make", "-s", "-C", csrc, "asm"], check=True)
    text = open(os.path.join(csrc, "..", "lib", "obj", "engine_kernels.s")).read()
    kernels = {k: v for k, v in isa_lint.parse_kernels(text).items()
               if "fengine_form_kernel" in k or "fengine_fp32_kernel" in k}
    assert len(kernels) >= 12
    loops = {k: isa_lint.waterfall_loops(v) for k, v in kernels.items()}
    assert not any(loops.values()), {k: v for k, v in loops.items() if v}


def test_waterfall_detector():
    """The detector flags the loop form the compiler emits for a buffer store
    with a per-lane descriptor, and not an ordinary loop."""
    import isa_lint
    loop = """k:
\ts_mov_b64 s[8:9], exec
.LBB0_1:
\tv_readfirstlane_b32 s4, v0
\tv_readfirstlane_b32 s5, v1
\tv_readfirstlane_b32 s6, v2
\tv_readfirstlane_b32 s7, v3
\tv_cmp_eq_u64_e32 vcc, s[4:5], v[0:1]
\tv_cmp_eq_u64_e64 s[0:1], s[6:7], v[2:3]
\ts_and_b64 s[0:1], vcc, s[0:1]
\ts_and_saveexec_b64 s[0:1], s[0:1]
\tbuffer_store_dword v16, v26, s[4:7], 0 offen
\ts_xor_b64 exec, exec, s[0:1]
\ts_cbranch_execnz .LBB0_1
\ts_mov_b64 exec, s[8:9]
\ts_endpgm
.Lfunc_end0:
"""
    (insts,) = isa_lint.parse_kernels(loop).values()
    assert isa_lint.waterfall_loops(insts) == [".LBB0_1"]
    plain = loop.replace("\tv_readfirstlane_b32 s4, v0\n", "").replace("\tv_readfirstlane_b32 s5, v1\n", "") \
        .replace("\tv_readfirstlane_b32 s6, v2\n", "").replace("\tv_readfirstlane_b32 s7, v3\n", "")
    (insts,) = isa_lint.parse_kernels(plain).values()
    assert isa_lint.waterfall_loops(insts) == []


def test_isa_lint_catches_the_known_miscompile():
    """A load result reused as the upper half of a zero-extended index, on one
    path only (the pattern ROCm 7.2's allocator produced for the F-engine)."""
    import isa_lint
    bad = """k:
\tv_mov_b32_e32 v1, 0
\ts_cbranch_scc0 .LBB0_2
\tglobal_load_ubyte v1, v1, s[6:7]
\tv_readfirstlane_b32 s6, v1
\ts_branch .LBB0_3
.LBB0_2:
\tv_mov_b32_e32 v1, 0
.LBB0_3:
\tv_lshlrev_b64 v[40:41], 3, v[0:1]
\tglobal_store_dwordx2 v[40:41], v[2:3], off
\ts_endpgm
.Lfunc_end0:
"""
    assert len(isa_lint.lint_text(bad)) == 1
    ok = bad.replace("\tv_readfirstlane_b32 s6, v1\n", "\tv_readfirstlane_b32 s6, v1\n\tv_mov_b32_e32 v1, 0\n")
    assert isa_lint.lint_text(ok) == []


def test_abi_path_in_cpu_build(built):
    """The whole drop-in path -- ABI layer, engine, launch geometry, copies -- in
    the CPU build under AddressSanitizer, driven like the reference's golden
    test: parity with the FP64 oracle and host-zeroed == mask-marked bytes."""
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True, check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_abi_emu.py")], capture_output=True,
                       text=True, timeout=900, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.strip().endswith("OK")
