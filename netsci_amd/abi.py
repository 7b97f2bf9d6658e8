"""ctypes access to libnsx.so's C ABI (include/netsci_hip.h).

This is the binding a Python caller of the C ABI would write (the netcalc /
cuarray modules go through the C++ host layer instead). It holds device
buffers as plain pointers and raises RuntimeError with nsx_last_error() on
any non-zero status — there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsx.so")

# (name, argtypes) for every exported function, in header order.
_VP, _I, _LL, _F, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float, ctypes.c_size_t
_IP, _LLP, _FP = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "nsx_last_error": [],
    "nsx_device_count": [_IP],
    "nsx_init": [_I],
    "nsx_device_info": [ctypes.c_char_p, _I, _IP, _IP],
    "nsx_alloc": [ctypes.POINTER(_VP), _SZ],
    "nsx_free": [_VP],
    "nsx_host_alloc": [ctypes.POINTER(_VP), _SZ],
    "nsx_host_free": [_VP],
    "nsx_memcpy_h2d": [_VP, _VP, _SZ],
    "nsx_memcpy_d2h": [_VP, _VP, _SZ],
    "nsx_memcpy_d2d": [_VP, _VP, _SZ, _VP],
    "nsx_stream_sync": [_VP],
    "nsx_device_sync": [],
    "nsx_event_create": [ctypes.POINTER(_VP)],
    "nsx_event_destroy": [_VP],
    "nsx_event_record": [_VP, _VP],
    "nsx_event_elapsed_ms": [_VP, _VP, _FP],
    "nsx_ksg_mi": [_VP, _LL, _LL, _I, _I, _VP, _LL, _I, _VP, _VP, _VP],
    "nsx_ksg_mi_windows": [_VP, _LL, _LL, _I, _IP, _IP, _I, _VP, _LL, _I, _VP, _VP],
    "nsx_ksg_mi_noalloc": [_VP, _LL, _LL, _I, _I, _VP, _LL, _I, _VP, _VP],
    "nsx_ksg_mi_chunked": [_VP, _LL, _LL, _I, _I, _VP, _LL, _I, _VP, _VP, _LLP, _I, _VP, _VP,
                           _VP],
    "nsx_ksg_last_tie_frames": [_LLP],
    "nsx_ksg_set_config": [_I, _I, _I, _I, _I],
    "nsx_ksg_set_method": [_I, _I, _I],
    "nsx_ksg_last_method": [_IP, _LLP],
    "nsx_ksg_last_phase_ms": [ctypes.POINTER(ctypes.c_double)],
    "nsx_ksg_last_call_ms": [ctypes.POINTER(ctypes.c_double)],
    "nsx_ksg_last_list_info": [_IP, _LLP, _LLP],
    "nsx_ksg_set_table_budget": [_LL],
    "nsx_ksg_set_list_cap": [_LL],
    "nsx_release_workspace": [],
    "nsx_ksg_plan_create": [ctypes.POINTER(_VP), _VP, _LL, _LL, _I, _IP, _IP, _I, _VP, _LL, _I,
                            _VP, _I],
    "nsx_ksg_plan_launch": [_VP, _VP],
    "nsx_ksg_plan_tie_frames": [_VP, _LLP],
    "nsx_ksg_plan_config": [_VP, _IP, _IP, _IP, _LLP, ctypes.POINTER(ctypes.c_double), _IP],
    "nsx_ksg_plan_destroy": [_VP],
    "nsx_ksg_plan_method": [_VP, _IP],
    "nsx_gc_from_mi_host": [_VP, _LL, _I, _VP],
    "nsx_row_mean": [_VP, _LL, _LL, _VP, _VP],
    "nsx_row_std": [_VP, _LL, _LL, _VP, _VP, _VP],
    "nsx_cpu_row_stats": [_VP, _LL, _LL, _VP, _VP],
    "nsx_psi_host": [_VP, _I],
    "nsx_screen_features": [_VP, _LL, _LL, _I, _I, _VP, _LL, _VP, _VP],
    "nsx_mlp_param_count": [_IP],
    "nsx_mlp_forward": [_VP, _VP, _LL, _VP, _VP],
    "nsx_mlp_grad": [_VP, _VP, _VP, _VP, _LL, _VP, _VP, _VP],
    "nsx_mlp_adam": [_VP, _VP, _VP, _VP, _I, _F, _VP],
    "nsx_screen_select": [_VP, _LL, _F, _F, ctypes.c_uint, _VP, _VP, _VP, _VP, _LLP, _LLP, _VP],
    "nsx_cpu_ksg_mi": [_VP, _LL, _LL, _I, _I, _VP, _LL, _I, _VP, _I],
    "nsx_select_prefix_host": [_VP, _I, _I, _VP, _IP],
    "nsx_hedetniemi": [_VP, _I, _I, _F, _VP, _VP, _VP, _VP],
    "nsx_hedetniemi_lengths": [_VP, _I, _I, _VP, _VP],
    "nsx_corr_to_adj": [_VP, _I, _VP, _VP],
    "nsx_cpu_hedetniemi": [_VP, _I, _I, _F, _VP, _VP, _I],
    "nsx_cpu_corr_to_adj": [_VP, _I, _VP],
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m netsci_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name == "nsx_last_error" else ctypes.c_int
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().nsx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def device_count():
    c = ctypes.c_int(0)
    lib().nsx_device_count(ctypes.byref(c))
    return c.value


def init(device=0):
    check(lib().nsx_init(device), "nsx_init")


def device_info():
    name = ctypes.create_string_buffer(256)
    cu, clk = ctypes.c_int(0), ctypes.c_int(0)
    check(lib().nsx_device_info(name, 256, ctypes.byref(cu), ctypes.byref(clk)), "device info")
    return name.value.decode(), cu.value, clk.value


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class DeviceArray:
    """A device buffer with a fixed dtype and shape (freed on close/del)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().nsx_alloc(ctypes.byref(p), max(self.nbytes, 4)), "nsx_alloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            check(lib().nsx_memcpy_h2d(d.ptr, _ptr(a), a.nbytes), "upload")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().nsx_memcpy_d2h(_ptr(out), self.ptr, self.nbytes), "download")
        return out

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().nsx_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ksg_mi(dX, ld, dstride, d, n, dab, P, k, dmi, dcounts=None, stream=None, frame_offset=0,
           out_offset=0):
    """nsx_ksg_mi; frame_offset shifts X to a window, out_offset (floats)
    shifts the MI output."""
    x = ctypes.c_void_p(dX.ptr.value + 4 * frame_offset)
    o = ctypes.c_void_p(dmi.ptr.value + 4 * out_offset)
    check(lib().nsx_ksg_mi(x, ld, dstride, d, n, dab.ptr, P, k, o,
                           dcounts.ptr if dcounts is not None else None, stream), "nsx_ksg_mi")


def ksg_mi_windows(dX, ld, dstride, d, windows, dab, P, k, dmi, stream=None):
    w = np.ascontiguousarray(windows, np.int32).reshape(-1, 2)
    off = np.ascontiguousarray(w[:, 0])
    ln = np.ascontiguousarray(w[:, 1])
    check(lib().nsx_ksg_mi_windows(dX.ptr, ld, dstride, d,
                                   off.ctypes.data_as(_IP), ln.ctypes.data_as(_IP), len(w),
                                   dab.ptr, P, k, dmi.ptr, stream), "nsx_ksg_mi_windows")


def set_config(qpl=0, chunk=128, wpb=4, blocks_per_cu=0, symmetric=0):
    """Force the eager KSG launch shape (qpl=0 restores tuned/default shapes);
    symmetric=1 selects the pass 2 that evaluates each unordered frame pair
    once."""
    check(lib().nsx_ksg_set_config(qpl, chunk, wpb, blocks_per_cu, int(symmetric)),
          "set config")


class KsgPlan:
    """A hipGraph-captured batched KSG launch (nsx_ksg_plan_*)."""

    def __init__(self, dX, ld, dstride, d, windows, dab, P, k, dmi, autotune=True):
        w = np.ascontiguousarray(windows, np.int32).reshape(-1, 2)
        self._off = np.ascontiguousarray(w[:, 0])
        self._len = np.ascontiguousarray(w[:, 1])
        self.ptr = ctypes.c_void_p()
        check(lib().nsx_ksg_plan_create(ctypes.byref(self.ptr), dX.ptr, ld, dstride, d,
                                        self._off.ctypes.data_as(_IP),
                                        self._len.ctypes.data_as(_IP), len(w), dab.ptr, P, k,
                                        dmi.ptr, int(bool(autotune))), "plan create")

    def launch(self, stream=None):
        check(lib().nsx_ksg_plan_launch(self.ptr, stream), "plan launch")

    def tie_frames(self):
        c = ctypes.c_longlong(0)
        check(lib().nsx_ksg_plan_tie_frames(self.ptr, ctypes.byref(c)), "plan ties")
        return c.value

    def config(self):
        q, c, w, sy = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        b, t = ctypes.c_longlong(), ctypes.c_double()
        check(lib().nsx_ksg_plan_config(self.ptr, ctypes.byref(q), ctypes.byref(c),
                                        ctypes.byref(w), ctypes.byref(b), ctypes.byref(t),
                                        ctypes.byref(sy)),
              "plan config")
        return {"queries_per_lane": q.value, "frame_tile": c.value, "waves_per_workgroup": w.value,
                "symmetric_pass2": sy.value, "workgroups": b.value, "tune_ms": t.value}

    def method(self):
        """2: the graph runs the sorted-list path; 1: the all-frames kernel."""
        m = ctypes.c_int(0)
        check(lib().nsx_ksg_plan_method(self.ptr, ctypes.byref(m)), "plan method")
        return m.value

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().nsx_ksg_plan_destroy(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


METHOD_AUTO, METHOD_ALL_FRAMES, METHOD_LISTS = 0, 1, 2


def set_method(method=METHOD_AUTO, walk_waves=0, walk_batch=0):
    """Algorithm of the single-window KSG call: METHOD_AUTO (sorted lists
    when they apply), METHOD_ALL_FRAMES, or METHOD_LISTS (lists or an error).
    Results are identical for every choice."""
    check(lib().nsx_ksg_set_method(int(method), int(walk_waves), int(walk_batch)), "set method")


def last_method():
    """(method the last ksg_mi call ran: 1 all-frames / 2 lists, list entries
    walked, both lists)."""
    m, e = ctypes.c_int(0), ctypes.c_longlong(0)
    lib().nsx_ksg_last_method(ctypes.byref(m), ctypes.byref(e))
    return m.value, e.value


def last_phase_ms():
    """Device ms of the last list-path call: {build, walk, ties, psi_sums}."""
    out = (ctypes.c_double * 4)()
    lib().nsx_ksg_last_phase_ms(out)
    return {"build": out[0], "walk": out[1], "ties": out[2], "psi_sums": out[3]}


def last_call_ms():
    """Host wall time (ms) of the last ksg_mi / ksg_mi_windows call."""
    v = ctypes.c_double(0)
    lib().nsx_ksg_last_call_ms(ctypes.byref(v))
    return v.value


def last_list_info():
    """(node blocks, table slots, node tables built) of the last list-path call;
    blocks == 1 when every used node's tables fit at once."""
    b, s, nb = ctypes.c_int(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
    lib().nsx_ksg_last_list_info(ctypes.byref(b), ctypes.byref(s), ctypes.byref(nb))
    return b.value, s.value, nb.value


def set_table_budget(nbytes=0):
    """Cap on the device memory of the list path's tables (0 = default)."""
    check(lib().nsx_ksg_set_table_budget(int(nbytes)), "set table budget")


def set_list_cap(entries=0):
    """Entries kept per sorted list of the list path's tables (0 = default:
    whole rows up to 512 frames, max(512, n / 8) beyond). Any cap gives the
    same results: walks that reach the end of a cut list go to the tie
    kernels."""
    check(lib().nsx_ksg_set_list_cap(int(entries)), "set list cap")


def table_slot_bytes(n, cap=0):
    """Device bytes of one node's list-path tables for windows of n frames
    (ksg.hip's slot_bytes_for) under list cap `cap` (0 = default): rows of
    n rounded up to 8, each with the first entries of its sorted list and a
    rank per frame (2 bytes each), and one float prior per row."""
    rows = (n + 7) // 8 * 8
    if cap > 0:
        ors = min(rows, max(8, (cap + 7) // 8 * 8))
    else:
        ors = min(rows, max(512, (n // 8 + 7) // 8 * 8))
    return rows * (ors + rows) * 2 + rows * 4


def release_workspace():
    """Free the library's device workspaces and tables."""
    check(lib().nsx_release_workspace(), "release workspace")


def last_tie_frames():
    c = ctypes.c_longlong(0)
    lib().nsx_ksg_last_tie_frames(ctypes.byref(c))
    return c.value


def gc_from_mi(mi, d):
    mi = np.ascontiguousarray(mi, np.float32)
    out = np.empty_like(mi)
    check(lib().nsx_gc_from_mi_host(_ptr(mi), mi.size, d, _ptr(out)), "gc transform")
    return out


def select_prefix(key, m, with_heap_count=False):
    key = np.ascontiguousarray(key, np.float32)
    out = np.empty(m, np.int32)
    heaps = ctypes.c_int(0)
    check(lib().nsx_select_prefix_host(_ptr(key), key.size, m, _ptr(out), ctypes.byref(heaps)),
          "select prefix")
    return (out, heaps.value) if with_heap_count else out


def cpu_ksg_mi(X, ab, k, n, d, threads=0, dstride=None):
    X = np.ascontiguousarray(X, np.float32)
    ab = np.ascontiguousarray(ab, np.int32).reshape(-1, 2)
    out = np.empty(len(ab), np.float32)
    check(lib().nsx_cpu_ksg_mi(_ptr(X), X.shape[1], dstride or n, d, n, _ptr(ab), len(ab), k,
                               _ptr(out), threads), "cpu ksg")
    return out


def screen_features(dX, ld, dstride, N, n, dab, P, dfeat, stream=None):
    check(lib().nsx_screen_features(dX.ptr, ld, dstride, N, n, dab.ptr, P, dfeat.ptr, stream),
          "nsx_screen_features")


def mlp_param_count():
    c = ctypes.c_int(0)
    lib().nsx_mlp_param_count(ctypes.byref(c))
    return c.value


def mlp_forward(dparams, dfeat, P, dout, stream=None):
    check(lib().nsx_mlp_forward(dparams.ptr, dfeat.ptr, P, dout.ptr, stream), "nsx_mlp_forward")


def mlp_grad(dparams, dfeat, dy, dw, B, dgrad, dloss, stream=None):
    check(lib().nsx_mlp_grad(dparams.ptr, dfeat.ptr, dy.ptr, dw.ptr, B, dgrad.ptr, dloss.ptr,
                             stream), "nsx_mlp_grad")


def mlp_adam(dparams, dgrad, dm, dv, step, lr, stream=None):
    check(lib().nsx_mlp_adam(dparams.ptr, dgrad.ptr, dm.ptr, dv.ptr, step, lr, stream),
          "nsx_mlp_adam")


def screen_select(dpred, P, threshold, audit_rate, seed, dab, dab_out, didx_out, dflag,
                  stream=None):
    ns, na = ctypes.c_longlong(0), ctypes.c_longlong(0)
    check(lib().nsx_screen_select(dpred.ptr, P, threshold, audit_rate, seed, dab.ptr,
                                  dab_out.ptr, didx_out.ptr, dflag.ptr, ctypes.byref(ns),
                                  ctypes.byref(na), stream), "nsx_screen_select")
    return ns.value, na.value


def hedetniemi(dA, N, L, tol, dH, dHi=None, dpaths=None, stream=None):
    check(lib().nsx_hedetniemi(dA.ptr, N, L, tol, dH.ptr, dHi.ptr if dHi is not None else None,
                               dpaths.ptr if dpaths is not None else None, stream),
          "nsx_hedetniemi")


def hedetniemi_lengths(dA, N, L, dH, stream=None):
    check(lib().nsx_hedetniemi_lengths(dA.ptr, N, L, dH.ptr, stream), "nsx_hedetniemi_lengths")


def corr_to_adj(dC, N, dA, stream=None):
    check(lib().nsx_corr_to_adj(dC.ptr, N, dA.ptr, stream), "nsx_corr_to_adj")


def sync():
    check(lib().nsx_device_sync(), "device sync")


class Timer:
    """hipEvent pair on a stream: elapsed device time in ms."""

    def __init__(self, stream=None):
        self.stream = stream
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().nsx_event_create(ctypes.byref(self.a)), "event")
        check(lib().nsx_event_create(ctypes.byref(self.b)), "event")

    def start(self):
        check(lib().nsx_event_record(self.a, self.stream), "record")

    def stop(self):
        check(lib().nsx_event_record(self.b, self.stream), "record")

    def ms(self):
        v = ctypes.c_float(0)
        check(lib().nsx_event_elapsed_ms(self.a, self.b, ctypes.byref(v)), "elapsed")
        return v.value
