"""Build the native parts of netsci_amd in-tree.

* ``libnsx.so`` — hipcc for gfx950: the HIP kernels, the C-ABI of
  include/netsci_hip.h and the host (CPU platform) implementations.
* ``_netsci*.so`` — g++: the C++ CuArray / netcalc host layer and its
  pybind11 bindings, linked against libnsx.so.
* ``libnsx_probe.so`` — hipcc: the FP32 VALU issue-rate probe that
  bench.py --valu-probe times (a measurement tool, not part of the C ABI).

Run ``python -m netsci_amd.build`` (or ``__graft_entry__.build()``).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ARCH = "gfx950"

HIP_SOURCES = ["nsx_runtime.hip", "ksg.hip", "ksg_d1.hip", "ksg_d2.hip", "ksg_d3.hip",
               "ksg_walk_d1.hip", "ksg_walk_d2.hip", "ksg_walk_d3.hip", "apsp.hip", "screener.hip", "stats.hip"]
HOST_SOURCES = ["host_cpu.cpp"]
PY_SOURCES = ["host/netcalc.cpp", "host/netchem.cpp", "host/py_module.cpp"]

# Packed FP32 instructions stay off: hipcc would otherwise form v_pk_* ops.
# Kernel arguments (up to 16 SGPRs) are preloaded into SGPRs at wave launch
# (gfx950), so short kernels do not start with a dependent argument load.
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-mllvm", "-amdgpu-kernarg-preload-count=16",
               "-Wall", "-Wno-unused-function"]
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-pthread", "-Wall"]


def _hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build libnsx")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def lib_path():
    return os.path.join(HERE, "libnsx.so")


def probe_path():
    return os.path.join(HERE, "libnsx_probe.so")


def ext_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_netsci" + suffix)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s.replace(".hip", ".o"))
        jobs.append(([hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o], o))
    for s in HOST_SOURCES:
        o = os.path.join(OBJ, s.replace(".cpp", ".o"))
        jobs.append((["g++"] + GXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o], o))
    workers = min(len(jobs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [o for _, o in jobs]
    tmp = lib_path() + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs +
         ["-pthread", "-L/opt/rocm/lib", "-lrocblas", "-lrocprofiler-sdk-roctx",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, lib_path())
    tmp = probe_path() + ".tmp"
    _run([hipcc] + HIPCC_FLAGS + ["-shared", os.path.join(CSRC, "valu_probe.hip"), "-o", tmp])
    os.replace(tmp, probe_path())

    import pybind11
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]
    srcs = [os.path.join(CSRC, s) for s in PY_SOURCES]
    tmp = ext_path() + ".tmp"
    _run(["g++"] + GXX_FLAGS + ["-shared", "-fvisibility=hidden"] + inc + srcs +
         ["-L" + HERE, "-lnsx", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, ext_path())
    return lib_path(), ext_path()


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", lib_path(), ext_path())
