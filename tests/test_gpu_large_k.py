"""KSG with k above the register heaps' 16 (GPU required).

The reference's GPU path takes any k (mutual_information_gpu.cu:480 runs
k + 1 argmin sweeps). Here k > 16 runs one wave per query frame (bisection
on the distance bits, libstdc++-order selection on exact ties): with the
rank tables on the list path, with a counting pass on the all-frames
method. Per-frame counts and MI must equal the oracle's
(oracle/ksg_oracle.cpp, the reference CPU path restated) bit for bit,
including the smallest window n = k + 1, ties and duplicate frames.
"""
import numpy as np
import pytest

import oracle
from netsci_amd import abi, cuarray, netcalc, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["lists", "all-frames"])
def method(request):
    abi.set_method(abi.METHOD_LISTS if request.param == "lists" else abi.METHOD_ALL_FRAMES)
    yield request.param
    abi.set_method(abi.METHOD_AUTO)


def _check(X, ab, k, n, d):
    X = np.ascontiguousarray(X, np.float32)
    ab = np.ascontiguousarray(ab, np.int32)
    P = len(ab)
    dX, dab = abi.DeviceArray.from_numpy(X), abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((P,), np.float32)
    dcnt = abi.DeviceArray((P, n, 2), np.int32)
    abi.ksg_mi(dX, X.shape[1], n, d, n, dab, P, k, dmi, dcnt)
    mi, cnt = dmi.numpy(), dcnt.numpy()
    for p, (a, b) in enumerate(ab):
        na, nb, _, _ = oracle.ksg_counts(X[a, : d * n], X[b, : d * n], k, n, d)
        np.testing.assert_array_equal(cnt[p, :, 0], na)
        np.testing.assert_array_equal(cnt[p, :, 1], nb)
        assert mi[p] == np.float32(oracle.mi(X[a, : d * n], X[b, : d * n], k, n, d))
    return abi.last_method()[0], abi.last_tie_frames()


@pytest.mark.parametrize("k", [17, 32, 64])
@pytest.mark.parametrize("extra", [1, 5, 300])
def test_large_k_counts(method, k, extra):
    n, d = k + extra, 3
    print(f"large k={k} n={n} ({method})", flush=True)
    X, _ = synth.calpha_trajectory(N=4, n=n, seed=k + extra)
    ab = np.array([[0, 1], [2, 3], [1, 1], [3, 0]], np.int32)
    got, _ = _check(X, ab, k, n, d)
    assert got == 2


@pytest.mark.parametrize("d", [1, 2, 3])
def test_large_k_ties_and_duplicates(method, d):
    """A coarse grid (many equal distances at the k-th neighbour) and
    duplicated frames."""
    n, k = 500, 40
    print(f"large k ties d={d} ({method})", flush=True)
    rng = np.random.default_rng(90 + d)
    X = (np.round(rng.standard_normal((3, d, n)) * 2) / 2).astype(np.float32)
    X[:, :, 400:450] = X[:, :, 0:50]
    X = X.reshape(3, d * n)
    _, ties = _check(X, np.array([[0, 1], [1, 2], [2, 0]], np.int32), k, n, d)
    assert ties > 0


def test_large_k_windows_and_pair_loop(method):
    """Windows and the netcalc pair loop (GPU platform) with k = 20."""
    N, nT, d, k = 5, 700, 3, 20
    print(f"large k windows ({method})", flush=True)
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=66)
    ab = synth.all_pairs(N)
    win = np.array([[0, 700], [100, 21], [50, 333]], np.int32)
    dX, dab = abi.DeviceArray.from_numpy(X), abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((len(win), len(ab)), np.float32)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, len(ab), k, dmi)
    got = dmi.numpy()
    for r, (off, ln) in enumerate(win):
        Xw = np.concatenate([X[:, di * nT + off: di * nT + off + ln] for di in range(d)], axis=1)
        np.testing.assert_array_equal(got[r], oracle.pairs(Xw, ab, k, int(ln), d, gc=False))
    Xc, R, abc = cuarray.FloatCuArray(), cuarray.FloatCuArray(), cuarray.IntCuArray()
    Xc.fromNumpy2D(X[:, : d * nT].copy())
    abc.fromNumpy2D(ab)
    assert netcalc.generalizedCorrelation(X=Xc, R=R, ab=abc, k=k, n=nT, xd=2, d=d,
                                          platform=netcalc.GPU_PLATFORM) == 0
    np.testing.assert_array_equal(R.toNumpy1D(), oracle.pairs(X, ab, k, nT, d, gc=True))


def test_large_k_limits():
    """k > 16 needs n <= 8192; plans and the allocation-free entry keep
    k <= 16."""
    X = np.zeros((2, 9000), np.float32)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(np.array([[0, 1]], np.int32))
    dmi = abi.DeviceArray((1,), np.float32)
    with pytest.raises(RuntimeError, match="8192"):
        abi.ksg_mi(dX, 9000, 9000, 1, 9000, dab, 1, 17, dmi)
    with pytest.raises(RuntimeError, match="k <= 16"):
        abi.KsgPlan(dX, 9000, 9000, 1, np.array([[0, 100]], np.int32), dab, 1, 17, dmi,
                    autotune=False)
