"""The allocation-free KSG entry point nsx_ksg_mi_noalloc (GPU required).

It has no counterpart in the reference; it exists for callers that cannot
allocate device memory. Called through the C ABI, its output must equal
nsx_ksg_mi's bit for bit (and the oracle's), on two calls in a row, the
second with a smaller input, because every call shares the GPU's static
scratch and must not depend on what the previous call left there.
"""
import ctypes

import numpy as np
import pytest

import oracle
from netsci_amd import abi, synth

pytestmark = pytest.mark.gpu


def _noalloc(dX, ld, dstride, d, n, dab, P, k, dmi):
    abi.check(abi.lib().nsx_ksg_mi_noalloc(dX.ptr, ld, dstride, d, n, dab.ptr, P, k, dmi.ptr,
                                           None), "nsx_ksg_mi_noalloc")
    abi.sync()


def _case(N, n, d, k, seed, pairs=None):
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=seed)
    X = np.ascontiguousarray(X[:, : d * n])
    ab = synth.all_pairs(N) if pairs is None else pairs
    return X, ab


@pytest.mark.parametrize("d,k", [(3, 6), (1, 4), (2, 16)])
def test_noalloc_equals_nsx_ksg_mi_twice(d, k):
    print(f"noalloc d={d} k={k}", flush=True)
    for N, n, seed in ((24, 1500, 5 + d), (9, 400, 9 + d)):  # the second call is smaller
        X, ab = _case(N, n, d, k, seed)
        P = len(ab)
        dX = abi.DeviceArray.from_numpy(X)
        dab = abi.DeviceArray.from_numpy(ab)
        ref = abi.DeviceArray((P,), np.float32)
        got = abi.DeviceArray((P,), np.float32)
        abi.ksg_mi(dX, X.shape[1], n, d, n, dab, P, k, ref)
        _noalloc(dX, X.shape[1], n, d, n, dab, P, k, got)
        np.testing.assert_array_equal(got.numpy(), ref.numpy())
        sample = ab[:: max(1, P // 6)]
        idx = np.arange(0, P, max(1, P // 6))
        np.testing.assert_array_equal(got.numpy()[idx], oracle.pairs(X, sample, k, n, d, gc=False))


def test_noalloc_ties_and_long_rows():
    """Ties at the k-th neighbour (coarse grid) use the static tie rows; a
    long window caps the launch to the waves whose rows fit."""
    print("noalloc ties and n=20000", flush=True)
    rng = np.random.default_rng(3)
    n, d, k = 600, 3, 6
    X = (np.round(rng.standard_normal((4, d * n)) * 3) / 3).astype(np.float32)
    ab = np.array([[0, 1], [2, 3], [1, 3]], np.int32)
    dX, dab = abi.DeviceArray.from_numpy(X), abi.DeviceArray.from_numpy(ab)
    got = abi.DeviceArray((3,), np.float32)
    _noalloc(dX, d * n, n, d, n, dab, 3, k, got)
    np.testing.assert_array_equal(got.numpy(), oracle.pairs(X, ab, k, n, d, gc=False))
    n = 20000
    X, _ = synth.calpha_trajectory(N=2, n=n, seed=4)
    ab = np.array([[0, 1]], np.int32)
    dX, dab = abi.DeviceArray.from_numpy(X), abi.DeviceArray.from_numpy(ab)
    ref, got = abi.DeviceArray((1,), np.float32), abi.DeviceArray((1,), np.float32)
    abi.set_method(abi.METHOD_ALL_FRAMES)
    try:
        abi.ksg_mi(dX, X.shape[1], n, 3, n, dab, 1, 6, ref)
    finally:
        abi.set_method(abi.METHOD_AUTO)
    _noalloc(dX, X.shape[1], n, 3, n, dab, 1, 6, got)
    np.testing.assert_array_equal(got.numpy(), ref.numpy())


def test_noalloc_rejects_oversize():
    X = np.zeros((2, 10), np.float32)
    dX, dab = abi.DeviceArray.from_numpy(X), abi.DeviceArray.from_numpy(np.array([[0, 1]], np.int32))
    out = abi.DeviceArray((1,), np.float32)
    L = abi.lib()
    assert L.nsx_ksg_mi_noalloc(dX.ptr, 1 << 19, 1 << 19, 1, (1 << 18) + 1, dab.ptr, 1, 4,
                                out.ptr, None) == 1
    assert b"262144" in L.nsx_last_error()
