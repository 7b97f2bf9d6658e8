"""ctypes bindings to oracle/liboracle.so (test infrastructure only).

Each wrapper names the reference file:line its C++ function restates; see
oracle/ksg_oracle.cpp for the restatement itself and what pins it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_psi.argtypes = [_f32p, ctypes.c_int]
        L.orc_ksg_counts.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _i32p, _i32p, _f32p, _f32p]
        L.orc_mi_from_counts.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int]
        L.orc_mi_from_counts.restype = ctypes.c_float
        L.orc_mi.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_mi.restype = ctypes.c_float
        L.orc_gc_from_mi.argtypes = [ctypes.c_float, ctypes.c_int]
        L.orc_gc_from_mi.restype = ctypes.c_float
        L.orc_pairs.argtypes = [_f32p, ctypes.c_longlong, _i32p, ctypes.c_longlong,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, _f32p]
        L.orc_corr_to_adj.argtypes = [_f32p, ctypes.c_int, _f32p]
        L.orc_row_mean.argtypes = [_f32p, ctypes.c_longlong, ctypes.c_longlong, _f32p]
        L.orc_row_std.argtypes = [_f32p, ctypes.c_longlong, ctypes.c_longlong, _f32p, _f32p]
        L.orc_hedetniemi.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     _f32p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_recover_path.argtypes = [_i32p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, _i32p]
        L.orc_recover_path.restype = ctypes.c_int
        L.orc_argsort_like_reference.argtypes = [_f32p, ctypes.c_int, _i32p]
        L.orc_heapsort_like_libstdcxx.argtypes = [_f32p, ctypes.c_int, _i32p]
        L.orc_antiqsort.argtypes = [ctypes.c_int, _f32p]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def psi(n):
    """psi.cpp:6-19."""
    out = np.zeros(n + 1, np.float32)
    lib().orc_psi(out, n)
    return out


def ksg_counts(xa, xb, k, n, d):
    """mutual_information_cpu.cpp:22-109 → (nXa, nXb, epsXa, epsXb) per frame."""
    xa, xb = _f32(xa).ravel(), _f32(xb).ravel()
    na = np.zeros(n, np.int32)
    nb = np.zeros(n, np.int32)
    ea = np.zeros(n, np.float32)
    eb = np.zeros(n, np.float32)
    lib().orc_ksg_counts(xa, xb, k, n, d, na, nb, ea, eb)
    return na, nb, ea, eb


def mi_from_counts(na, nb, k, n):
    """mutual_information_cpu.cpp:110-118."""
    return float(lib().orc_mi_from_counts(np.ascontiguousarray(na, np.int32),
                                          np.ascontiguousarray(nb, np.int32), k, n))


def mi(xa, xb, k, n, d):
    """mutual_information_cpu.cpp:6-121 (mutualInformationCpu)."""
    return float(lib().orc_mi(_f32(xa).ravel(), _f32(xb).ravel(), k, n, d))


def gc_from_mi(m, d):
    """generalized_correlation_cpu.cpp:8-27."""
    return float(lib().orc_gc_from_mi(m, d))


def pairs(X, ab, k, n, d, gc=True):
    """generalized_correlation.cpp:104-173 / mutual_information.cpp:103-172."""
    X = _f32(X)
    ab = np.ascontiguousarray(ab, np.int32).reshape(-1, 2)
    out = np.zeros(ab.shape[0], np.float32)
    lib().orc_pairs(X, X.shape[1], ab, ab.shape[0], k, n, d, int(gc), out)
    return out


def corr_to_adj(C):
    """hedetniemi_gpu.cu:8-26."""
    C = _f32(C)
    A = np.zeros_like(C)
    lib().orc_corr_to_adj(C, C.shape[0], A)
    return A


def row_mean(a):
    """math_utils_gpu.cu:23-68: per-row mean in the reference's summation order."""
    a = _f32(a)
    u = np.zeros(a.shape[0], np.float32)
    lib().orc_row_mean(a, a.shape[0], a.shape[1], u)
    return u


def row_std(a):
    """math_utils_gpu.cu:70-156: per-row mean and population deviation."""
    a = _f32(a)
    u = np.zeros(a.shape[0], np.float32)
    sigma = np.zeros(a.shape[0], np.float32)
    lib().orc_row_std(a, a.shape[0], a.shape[1], u, sigma)
    return u, sigma


def hedetniemi(A, L, tol, want_hi=False, want_paths=True):
    """hedetniemi_gpu.cu:47-107 (Jacobi form) → (H, Hi or None, paths or None)."""
    A = _f32(A)
    N = A.shape[0]
    H = np.zeros_like(A)
    Hi = np.zeros((L, N, N), np.float32) if want_hi else None
    paths = np.zeros(N * L * N, np.int32) if want_paths else None
    lib().orc_hedetniemi(A, N, L, tol, H,
                         Hi.ctypes.data if Hi is not None else None,
                         paths.ctypes.data if paths is not None else None)
    return H, Hi, paths


def recover_path(paths, N, L, i, j):
    """hedetniemi.cpp:71-108."""
    out = np.zeros(L + 2, np.int32)
    m = lib().orc_recover_path(np.ascontiguousarray(paths, np.int32), N, L, i, j, out)
    return out[:m].copy()


def argsort_like_reference(key):
    key = _f32(key)
    out = np.zeros(key.size, np.int32)
    lib().orc_argsort_like_reference(key, key.size, out)
    return out


def heapsort_like_libstdcxx(key):
    key = _f32(key)
    out = np.zeros(key.size, np.int32)
    lib().orc_heapsort_like_libstdcxx(key, key.size, out)
    return out


def antiqsort(n):
    """An input that drives std::sort into its heap-sort fallback."""
    out = np.zeros(n, np.float32)
    lib().orc_antiqsort(n, out)
    return out
