"""The library's host code (CSR structure, reorderings, mesh generator and files, Delaunay, spectral
ordering) under AddressSanitizer, LeakSanitizer and UndefinedBehaviorSanitizer: tests/sanitize_host.cpp
drives every host-only entry point with structural checks, built here with g++ from the host sources
of parallel_fem_codes_amd/csrc (no device code is compiled or run)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "parallel_fem_codes_amd", "csrc")
SOURCES = ["structure.cpp", "meshio.cpp", "delaunay.cpp", "spectral.cpp", "errors.cpp"]


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_host_code_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
           "-I" + os.path.join(ROOT, "include")]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + [os.path.join(ROOT, "tests", "sanitize_host.cpp")]
    cmd += ["-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, cwd=str(tmp_path), env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "OK: 0 failed checks" in r.stdout
