"""The drop-in boundary as the reference's host builds and links it.

corr-rs/build.rs:20-57 runs CMake on the sibling directory `corr-cuda`
(CMAKE_BUILD_TYPE=Release, CMAKE_LIBRARY_OUTPUT_DIRECTORY and
CMAKE_RUNTIME_OUTPUT_DIRECTORY set to its output directory), builds
`--target corr_cuda`, and links `dylib=corr_cuda`. These tests do exactly that
with this library's CMake project laid out as a self-contained `corr-cuda/`
directory (csrc/ plus include/ inside it), then:
  - the shared library leaves no symbol of its own undefined (a source left
    out of the target shows up as an undefined corr:: symbol);
  - every HIP compile carries the no-packed-FP32 option;
  - the C harness tests/c_abi/abi_check.c compiles against include/corr_cuda.h
    with the struct sizes asserted at compile time, links with all 26 entry
    points referenced, and starts with every symbol bound (LD_BIND_NOW=1).
The reference's own C tests are linked the same way
(corr-cuda/CMakeLists.txt:84-163). No device is touched.
"""

import glob
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "21cma_abf_correlator_amd", "csrc")
HARNESS = os.path.join(ROOT, "tests", "c_abi", "abi_check.c")


@pytest.fixture(scope="module")
def cmake_build(tmp_path_factory):
    top = tmp_path_factory.mktemp("boundary")
    src = top / "corr-cuda"
    src.mkdir()
    for f in os.listdir(CSRC):
        if f.endswith((".hip", ".cpp", ".hpp", ".txt")):
            shutil.copy(os.path.join(CSRC, f), src / f)
    shutil.copytree(os.path.join(ROOT, "include"), src / "include")
    lib_out = top / "corr-cuda-lib"
    lib_out.mkdir()
    bld = top / "build"
    env = dict(os.environ)
    subprocess.run(["cmake", "-S", str(src), "-B", str(bld), "-DCMAKE_BUILD_TYPE=Release",
                    f"-DCMAKE_LIBRARY_OUTPUT_DIRECTORY={lib_out}",
                    f"-DCMAKE_RUNTIME_OUTPUT_DIRECTORY={lib_out}",
                    "-DCMAKE_EXPORT_COMPILE_COMMANDS=ON"],
                   check=True, capture_output=True, env=env)
    subprocess.run(["cmake", "--build", str(bld), "--target", "corr_cuda", "--config", "Release",
                    "-j8"], check=True, capture_output=True, env=env)
    return dict(top=top, lib_out=lib_out, bld=bld)


def test_cmake_target_builds_the_library(cmake_build):
    libs = sorted(glob.glob(str(cmake_build["lib_out"] / "libcorr_cuda.so*")))
    assert libs, "libcorr_cuda.so was not produced where build.rs looks for it"


def test_no_undefined_symbol_of_the_library(cmake_build):
    so = str(cmake_build["lib_out"] / "libcorr_cuda.so")
    und = subprocess.run(["nm", "-DC", "--undefined-only", so], capture_output=True, text=True,
                         check=True).stdout
    own = [ln for ln in und.splitlines() if "corr::" in ln or " corr_" in ln or " corrx_" in ln]
    assert not own, f"undefined symbols of the library itself: {own}"
    defined = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                             check=True).stdout
    for name in ("corr_create", "corr_rows_fetch", "corrx_batch_run", "corrx_integrate"):
        assert f" T {name}\n" in defined, name


def test_every_hip_compile_drops_packed_fp32(cmake_build):
    cmds = json.load(open(cmake_build["bld"] / "compile_commands.json"))
    files = sorted(os.path.basename(c["file"]) for c in cmds)
    assert files == sorted(["batch.cpp", "corr_abi.cpp", "corrx_abi.cpp", "engine.cpp",
                            "engine_kernels.hip", "noalloc.hip"])
    for c in cmds:
        line = c.get("command") or " ".join(c["arguments"])
        assert "-packed-fp32-ops" in line and "gfx950" in line, c["file"]


def _c_harness(out_dir, lib_dir):
    exe = os.path.join(str(out_dir), "abi_check")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-Werror",
                    "-I" + os.path.join(ROOT, "include"), HARNESS, "-o", exe,
                    "-L" + str(lib_dir), "-lcorr_cuda", "-lm", "-Wl,-rpath," + str(lib_dir)],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60,
                       env=dict(os.environ, LD_BIND_NOW="1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "abi=7" in r.stdout and "linked 26 entry points" in r.stdout, r.stdout


def test_c_harness_links_the_cmake_library(cmake_build):
    _c_harness(cmake_build["top"], cmake_build["lib_out"])


def test_c_harness_links_the_makefile_library(corr, tmp_path):
    _c_harness(tmp_path, os.path.dirname(corr.LIB_PATH))
