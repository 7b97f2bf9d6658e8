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


def test_plan_replays_match_eager():
    """Three graph replays separated by eager work and synchronisation; each
    replay's output equals the eager windowed call's."""
    X, ab, nT, d, k = _problem()
    win = synth.bootstrap_windows(R=6, n_total=nT, lo=40, hi=nT, seed=2)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    R, P = len(win), len(ab)
    eager = abi.DeviceArray((R, P), np.float32)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, P, k, eager)
    want = eager.numpy()
    out = abi.DeviceArray((R, P), np.float32)
    plan = abi.KsgPlan(dX, X.shape[1], nT, d, win, dab, P, k, out, autotune=True)
    print("plan config", plan.config(), flush=True)
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
