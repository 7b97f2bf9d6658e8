"""Batched (windowed) KSG: every launch shape gives the same bits, and a
hipGraph-captured plan replays exactly (GPU required)."""
import numpy as np
import pytest

import oracle
from netsci_amd import abi, synth

pytestmark = pytest.mark.gpu

SHAPES = [(q, c, w, s) for q in (1, 2) for c in (64, 128, 256) for w in (1, 4, 8) for s in (0, 1)]


def _problem():
    N, nT, d, k = 7, 331, 3, 5
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=21)
    ab = synth.all_pairs(N)
    return X, ab, nT, d, k


def test_every_launch_shape_is_bitexact():
    X, ab, n, d, k = _problem()
    want = oracle.pairs(X, ab, k, n, d, gc=False)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((len(ab),), np.float32)
    abi.set_method(abi.METHOD_ALL_FRAMES)
    try:
        for q, c, w, s in SHAPES:
            print(f"shape qpl={q} tile={c} wpb={w} symmetric={s}", flush=True)
            abi.set_config(q, c, w, 0, s)
            abi.ksg_mi(dX, X.shape[1], n, d, n, dab, len(ab), k, dmi)
            np.testing.assert_array_equal(dmi.numpy(), want, err_msg=f"shape {(q, c, w, s)}")
    finally:
        abi.set_config(0)
        abi.set_method(abi.METHOD_AUTO)


@pytest.mark.parametrize("method", [abi.METHOD_AUTO, abi.METHOD_ALL_FRAMES])
def test_plan_replays_match_eager(method):
    """Three graph replays separated by eager work and synchronisation; each
    replay's output equals the eager windowed call's. With the automatic
    choice the graph runs the sorted-list path, else the all-frames kernel."""
    X, ab, nT, d, k = _problem()
    win = synth.bootstrap_windows(R=6, n_total=nT, lo=40, hi=nT, seed=2)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    R, P = len(win), len(ab)
    eager = abi.DeviceArray((R, P), np.float32)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, P, k, eager)
    want = eager.numpy()
    out = abi.DeviceArray((R, P), np.float32)
    abi.set_method(method)
    try:
        plan = abi.KsgPlan(dX, X.shape[1], nT, d, win, dab, P, k, out, autotune=True)
    finally:
        abi.set_method(abi.METHOD_AUTO)
    assert plan.method() == (2 if method == abi.METHOD_AUTO else 1)
    print("plan config", plan.method(), plan.config(), flush=True)
    other = abi.DeviceArray((P,), np.float32)
    for rep in range(3):
        abi.check(abi.lib().nsx_memcpy_d2d(out.ptr, other.ptr, 4 * P, None), "clobber")
        abi.sync()
        plan.launch()
        abi.sync()
        np.testing.assert_array_equal(out.numpy(), want, err_msg=f"replay {rep}")
        # eager work between replays
        abi.ksg_mi(dX, X.shape[1], nT, d, nT, dab, P, k, other)
        abi.sync()
    for r, (off, ln) in enumerate(win[:2]):
        Xw = np.concatenate([X[:, di * nT + off: di * nT + off + ln] for di in range(d)], axis=1)
        np.testing.assert_array_equal(want[r], oracle.pairs(Xw, ab, k, int(ln), d, gc=False))
    plan.close()


def _windows_eager(dX, ld, nT, d, win, dab, P, k, method):
    out = abi.DeviceArray((len(win), P), np.float32)
    abi.set_method(method)
    try:
        abi.ksg_mi_windows(dX, ld, nT, d, win, dab, P, k, out)
    finally:
        abi.set_method(abi.METHOD_AUTO)
    return out.numpy()


def test_list_plan_follows_new_coordinates_and_falls_back():
    """A list-path graph recomputes every window's tables on each replay: new
    coordinates in X (same shape) give the eager results on them; a
    non-finite coordinate makes the graph's guarded all-frames launch run
    (the eager automatic choice does the same); finite coordinates again
    take the list path."""
    X, ab, nT, d, k = _problem()
    win = np.array([[0, 331], [10, 200], [5, 40], [17, 100]], np.int32)
    R, P = len(win), len(ab)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    out = abi.DeviceArray((R, P), np.float32)
    plan = abi.KsgPlan(dX, X.shape[1], nT, d, win, dab, P, k, out, autotune=False)
    assert plan.method() == 2
    X2, _ = synth.calpha_trajectory(N=X.shape[0], n=nT, seed=99)
    X3 = X2.copy()
    X3[1, 17] = np.inf  # frame 17, x of node 1: in every window
    assert all(off <= 17 < off + ln for off, ln in win)
    for rep, Xr in enumerate((X2, X3, X)):
        print(f"list plan replay {rep}", flush=True)
        abi.check(abi.lib().nsx_memcpy_h2d(dX.ptr, np.ascontiguousarray(Xr).ctypes.data,
                                           Xr.nbytes), "refill X")
        plan.launch()
        abi.sync()
        got = out.numpy()
        want = _windows_eager(dX, X.shape[1], nT, d, win, dab, P, k, abi.METHOD_AUTO)
        assert abi.last_method()[0] == (1 if rep == 1 else 2)
        np.testing.assert_array_equal(got, want, err_msg=f"replay {rep}")
    plan.close()


def test_blocked_list_plan():
    """A list-path graph whose tables hold 3 node slots (a blocked schedule
    captured once) equals the eager windows on every replay."""
    N, nT, d, k = 9, 400, 3, 6
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=8)
    ab = synth.all_pairs(N)
    win = np.array([[0, 400], [33, 250]], np.int32)
    R, P = len(win), len(ab)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    want = _windows_eager(dX, X.shape[1], nT, d, win, dab, P, k, abi.METHOD_ALL_FRAMES)
    out = abi.DeviceArray((R, P), np.float32)
    abi.set_table_budget(3 * abi.table_slot_bytes(nT))
    try:
        plan = abi.KsgPlan(dX, X.shape[1], nT, d, win, dab, P, k, out, autotune=False)
    finally:
        abi.set_table_budget(0)
    assert plan.method() == 2
    for rep in range(2):
        plan.launch()
        abi.sync()
        np.testing.assert_array_equal(out.numpy(), want, err_msg=f"replay {rep}")
    plan.close()
