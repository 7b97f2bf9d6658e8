"""Checkpointed GC on the GPU platform (GPU required): the file protocol of
generalized_correlation.cpp:8-101, the values equal to the oracle's pair
loop (the reference CPU path restated) and to the CPU platform's run of the
same call."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from netsci_amd import abi, cuarray, netcalc, synth

pytestmark = pytest.mark.gpu


def _fa(x):
    c = cuarray.FloatCuArray()
    c.fromNumpy2D(np.ascontiguousarray(x, np.float32))
    return c


def _ia(x):
    c = cuarray.IntCuArray()
    c.fromNumpy2D(np.ascontiguousarray(x, np.int32))
    return c


def test_gpu_checkpoints_equal_cpu(tmp_path):
    print("checkpoint gpu", flush=True)
    N, n, d, k = 7, 180, 3, 4
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=12)
    ab = synth.all_pairs(N)  # 21 pairs
    out = {}
    for plat, sub in ((netcalc.GPU_PLATFORM, "gpu"), (netcalc.CPU_PLATFORM, "cpu")):
        os.makedirs(tmp_path / sub)
        base = str(tmp_path / sub / "ck.npy")
        R = cuarray.FloatCuArray()
        rc = netcalc.generalizedCorrelationWithCheckpointing(
            X=_fa(X), R=R, ab=_ia(ab), k=k, n=n, xd=2, d=d, platform=plat,
            checkpointFrequency=5, checkpointFileName=base)
        assert rc == plat
        out[sub] = (R.toNumpy1D(), sorted(os.listdir(tmp_path / sub)),
                    np.load(str(tmp_path / sub / "ck_20.npy")))
    want = oracle.pairs(X, ab, k, n, d, gc=True)
    np.testing.assert_array_equal(out["gpu"][0], want)
    np.testing.assert_array_equal(out["gpu"][0], out["cpu"][0])
    # The last checkpoint (after pair 20) holds every value.
    np.testing.assert_array_equal(out["gpu"][2].ravel(), want)
    assert out["gpu"][1] == out["cpu"][1] == ["ck_20.npy", "ck_ab.npy"]
    np.testing.assert_array_equal(out["gpu"][2], out["cpu"][2])


_CHUNK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong)


def _chunked(X, ab, k, n, d, ends, stop_at=None):
    """nsx_ksg_mi_chunked through the C ABI; returns the final values and a
    copy of the host output taken at every callback."""
    X = np.ascontiguousarray(X, np.float32)
    P = len(ab)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(np.ascontiguousarray(ab, np.int32))
    dmi = abi.DeviceArray((P,), np.float32)
    host = np.full(P, np.nan, np.float32)
    e = np.ascontiguousarray(ends, np.int64)
    seen = []

    def cb(user, end):
        seen.append((int(end), host.copy()))
        return 1 if stop_at is not None and end >= stop_at else 0

    fn = _CHUNK_FN(cb)
    rc = abi.lib().nsx_ksg_mi_chunked(
        dX.ptr, X.shape[1], n, d, n, dab.ptr, P, k, dmi.ptr, host.ctypes.data,
        e.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(e), fn, None, None)
    return rc, host, seen


@pytest.mark.parametrize("method,budget_slots", [(abi.METHOD_AUTO, 0), (abi.METHOD_AUTO, 4),
                                                 (abi.METHOD_ALL_FRAMES, 0)])
def test_chunked_pair_loop_values_at_every_chunk(method, budget_slots):
    """Every callback sees exactly the finished chunks' values (equal to
    nsx_ksg_mi's and the oracle's) and nothing after them; on the list path
    the tables are built once (blocked tables and the all-frames method run
    each chunk as its own call)."""
    print(f"chunked method={method} budget={budget_slots}", flush=True)
    N, n, d, k = 12, 300, 3, 6
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=31)
    ab = synth.all_pairs(N)  # 66 pairs
    ends = [1, 11, 21, 31, 41, 51, 61, 66]
    want = oracle.pairs(X, ab, k, n, d, gc=False)
    abi.set_method(method)
    abi.set_table_budget(budget_slots * abi.table_slot_bytes(n))
    try:
        rc, host, seen = _chunked(X, ab, k, n, d, ends)
    finally:
        abi.set_method(abi.METHOD_AUTO)
        abi.set_table_budget(0)
    assert rc == 0, abi.lib().nsx_last_error()
    np.testing.assert_array_equal(host, want)
    assert [s for s, _ in seen] == ends
    for end, snap in seen:
        np.testing.assert_array_equal(snap[:end], want[:end])
        assert np.isnan(snap[end:]).all()


def test_chunked_callback_stops_the_call():
    N, n, d, k = 6, 200, 3, 4
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=32)
    ab = synth.all_pairs(N)
    rc, _, seen = _chunked(X, ab, k, n, d, [5, 10, 15], stop_at=10)
    assert rc != 0 and b"callback" in abi.lib().nsx_last_error()
    assert [s for s, _ in seen] == [5, 10]
