"""A C++ caller written with the reference's netcalc signatures compiles
against netsci_amd/csrc/host/netcalc.h and links against libnsx.so
(tests/native/reference_signatures.cpp); its host-only checks run without a
GPU: recoverSingleShortestPath in the reference's (int**, int**, CuArray<int>*,
int, int, int) form (netcalc/include/hedetniemi.h:173-180) equals the vector
form, the CPU pair loop returns the platform, and xd != 2 raises on the GPU
platform."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "netsci_amd", "libnsx.so")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(LIB),
                    reason="needs g++ and the built libnsx.so")
def test_reference_signatures_compile_and_run(tmp_path):
    exe = str(tmp_path / "reference_signatures")
    cmd = ["g++", "-std=c++17", "-O1", "-pthread",
           os.path.join(ROOT, "tests", "native", "reference_signatures.cpp"),
           os.path.join(ROOT, "netsci_amd", "csrc", "host", "netcalc.cpp"),
           "-L" + os.path.dirname(LIB), "-lnsx", "-Wl,-rpath," + os.path.dirname(LIB),
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "reference signatures: ok" in run.stdout
