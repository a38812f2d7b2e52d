"""Host-side AddressSanitizer + UndefinedBehaviorSanitizer build of the library's CPU-reachable
code (template bank, SIGPROC reader) and of the oracle's C restatement (oracle/fdas_cpu.c), run on
small ragged cases (tests/cpp/host_sanitize_test.cpp).  CPU only; the GPU kernels are not built
with a sanitizer."""

import shutil
import subprocess

import pytest

from conftest import ROOT


def test_host_code_under_asan_and_ubsan(tmp_path):
    gxx, gcc = shutil.which("g++"), shutil.which("gcc")
    if not gxx or not gcc:
        pytest.skip("gcc / g++ not available")
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
    csrc = ROOT / "fdas_amd" / "csrc"
    cpu_o = tmp_path / "fdas_cpu.o"
    r = subprocess.run([gcc, *san, "-std=c11", "-fopenmp", "-DREAL=float", "-c", str(ROOT / "oracle" / "fdas_cpu.c"),
                        "-o", str(cpu_o)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = tmp_path / "host_sanitize_test"
    r = subprocess.run([gxx, *san, "-std=c++17", "-fopenmp", f"-I{ROOT / 'include'}", f"-I{csrc}",
                        str(ROOT / "tests" / "cpp" / "host_sanitize_test.cpp"), str(csrc / "fdas_templates.cpp"),
                        str(csrc / "fdas_io.cpp"), str(cpu_o), "-o", str(exe), "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
           "PATH": "/usr/bin:/bin"}
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-5000:]
    assert "host sanitize test ok" in r.stdout
    assert "runtime error" not in r.stderr
