"""The host-side C++ (CPU platform KSG and APSP, prefix selection, row
statistics, CuArray's host side, netchem's PDB / DCD readers) under
AddressSanitizer and
UndefinedBehaviorSanitizer: tests/native/host_sanitize.cpp is compiled with
g++ (no HIP code involved) and must run clean, its own result checks
included."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
def test_host_code_is_clean_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-pthread",
           os.path.join(ROOT, "tests", "native", "host_sanitize.cpp"),
           os.path.join(ROOT, "netsci_amd", "csrc", "host_cpu.cpp"),
           os.path.join(ROOT, "netsci_amd", "csrc", "host", "netchem.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    gold = os.path.join(ROOT, "tests", "golden", "netchem")
    run = subprocess.run([exe, gold], cwd=str(tmp_path), capture_output=True, text=True,
                         env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "host sanitizer run: ok" in run.stdout
