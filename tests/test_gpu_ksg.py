"""Parity of the gfx950 KSG path with the oracle (GPU required).

The bar is bit-exactness: per-frame n_x / n_y counts and the MI float must
equal the oracle's (oracle/ksg_oracle.cpp, the reference CPU path restated),
ties included. GC is compared bit-exactly too (host double transform).
"""
import numpy as np
import pytest

import oracle
from netsci_amd import abi, cuarray, netcalc, synth
from tests.conftest import curve_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=[0, 1], ids=["pass2", "sympass2"])
def pass2_variant(request):
    """Every test runs with the plain pass 2 (each ordered frame pair) and
    with the symmetric one (each unordered pair once, both frames credited);
    both must give the oracle's bits."""
    abi.set_config(2, 128, 4, 0, request.param)
    yield request.param
    abi.set_config(0)


def _flush(msg):
    print(msg, flush=True)


def _gpu_counts(X, ab, k, n, d, dstride=None):
    """MI and per-frame counts through the C ABI."""
    X = np.ascontiguousarray(X, np.float32)
    ab = np.ascontiguousarray(ab, np.int32).reshape(-1, 2)
    P = len(ab)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((P,), np.float32)
    dcnt = abi.DeviceArray((P, n, 2), np.int32)
    abi.ksg_mi(dX, X.shape[1], dstride or n, d, n, dab, P, k, dmi, dcnt)
    return dmi.numpy(), dcnt.numpy(), abi.last_tie_frames()


def test_golden_fixtures_gpu(golden_1000, golden_2000):
    """tests/netcalc/python/test_MutualInformation.py:13-64 on the GPU."""
    for X, n, want in ((golden_1000, 1000, 0.8449974060058594),
                       (golden_2000, 2000, 0.8485465049743652)):
        _flush(f"golden n={n}")
        Xa, Xb = cuarray.FloatCuArray(), cuarray.FloatCuArray()
        Xa.fromNumpy1D(X[:n])
        Xb.fromNumpy1D(X[n:])
        gpu = netcalc.mutualInformationGpu(Xa=Xa, Xb=Xb, k=4, n=n, xd=2, d=1)
        cpu = netcalc.mutualInformationCpu(Xa=Xa, Xb=Xb, k=4, n=n, xd=2, d=1)
        assert gpu == cpu == want


@pytest.mark.parametrize("d,n,variant", [(1, 1000, "gc"), (2, 1000, "mi"), (2, 2000, "gc"),
                                         (3, 1000, "gc"), (3, 2000, "gc")])
def test_reference_curves_gpu_equals_cpu(d, n, variant):
    """The reference's CPU == GPU cases (test_GeneralizedCorrelation.py:12-226,
    test_MutualInformation.py:67-154) plus the oracle's value."""
    _flush(f"curves d={d} n={n}")
    xa, xb = curve_inputs(n, d, variant)
    Xa, Xb = cuarray.FloatCuArray(), cuarray.FloatCuArray()
    Xa.fromNumpy2D(xa.reshape(d, n))
    Xb.fromNumpy2D(xb.reshape(d, n))
    gpu = netcalc.generalizedCorrelationGpu(Xa=Xa, Xb=Xb, k=4, n=n, xd=2, d=d)
    cpu = netcalc.generalizedCorrelationCpu(Xa=Xa, Xb=Xb, k=4, n=n, xd=2, d=d)
    want = oracle.gc_from_mi(oracle.mi(xa, xb, 4, n, d), d)
    assert gpu == cpu == want


@pytest.mark.parametrize("d,k", [(1, 4), (2, 3), (3, 6), (3, 1), (3, 16)])
def test_counts_bitexact_random(d, k):
    """Per-frame counts equal the oracle's on C-alpha-like data."""
    n = 300
    _flush(f"counts d={d} k={k}")
    X3, _ = synth.calpha_trajectory(N=6, n=n, seed=11 + d + k)
    X = np.ascontiguousarray(X3[:, : d * n])  # first d coordinates, row stride d*n
    ab = np.array([[0, 1], [2, 3], [4, 5], [1, 4], [3, 3]], np.int32)
    mi, cnt, _ = _gpu_counts(X, ab, k, n, d)
    for p, (a, b) in enumerate(ab):
        na, nb, _, _ = oracle.ksg_counts(X[a, : d * n], X[b, : d * n], k, n, d)
        np.testing.assert_array_equal(cnt[p, :, 0], na)
        np.testing.assert_array_equal(cnt[p, :, 1], nb)
        assert mi[p] == np.float32(oracle.mi(X[a, : d * n], X[b, : d * n], k, n, d))


def test_ties_resolved_in_std_sort_order():
    """Coordinates on a coarse grid make k-th-neighbour ties common; counts
    must still equal the oracle's (std::sort order) frame by frame."""
    n, d, k = 400, 3, 6
    _flush("ties")
    rng = np.random.default_rng(5)
    X = (np.round(rng.standard_normal((4, d * n)) * 4) / 4).astype(np.float32)
    ab = np.array([[0, 1], [2, 3], [0, 3]], np.int32)
    mi, cnt, ties = _gpu_counts(X, ab, k, n, d)
    assert ties > 0
    for p, (a, b) in enumerate(ab):
        na, nb, _, _ = oracle.ksg_counts(X[a], X[b], k, n, d)
        np.testing.assert_array_equal(cnt[p, :, 0], na)
        np.testing.assert_array_equal(cnt[p, :, 1], nb)
        assert mi[p] == np.float32(oracle.mi(X[a], X[b], k, n, d))


def test_duplicate_frames():
    """Exact duplicate frames (zero joint distance to another frame)."""
    n, d, k = 200, 3, 4
    _flush("duplicates")
    rng = np.random.default_rng(9)
    base = rng.standard_normal((2, d, n)).astype(np.float32)
    base[:, :, 100:150] = base[:, :, 0:50]  # 50 duplicated frames in both rows
    X = base.reshape(2, d * n)
    mi, cnt, _ = _gpu_counts(X, np.array([[0, 1]], np.int32), k, n, d)
    na, nb, _, _ = oracle.ksg_counts(X[0], X[1], k, n, d)
    np.testing.assert_array_equal(cnt[0, :, 0], na)
    np.testing.assert_array_equal(cnt[0, :, 1], nb)
    assert mi[0] == np.float32(oracle.mi(X[0], X[1], k, n, d))


def test_pair_loop_gc_matches_oracle():
    """netcalc.generalizedCorrelation (GPU platform) over many pairs equals
    the oracle's pair loop; the return value is the platform (0)."""
    N, n, d, k = 24, 240, 3, 6
    _flush("pair loop")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=2)
    ab = synth.all_pairs(N)
    Xc, R, abc = cuarray.FloatCuArray(), cuarray.FloatCuArray(), cuarray.IntCuArray()
    Xc.fromNumpy2D(X)
    abc.fromNumpy2D(ab)
    assert netcalc.generalizedCorrelation(X=Xc, R=R, ab=abc, k=k, n=n, xd=2, d=d,
                                          platform=netcalc.GPU_PLATFORM) == 0
    got = R.toNumpy1D()
    want = oracle.pairs(X, ab, k, n, d, gc=True)
    np.testing.assert_array_equal(got, want)


def test_windows_equal_sliced_calls():
    """Each window of the batched launch equals a plain call on its frames."""
    N, nT, d, k = 6, 600, 3, 6
    _flush("windows")
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=4)
    ab = synth.all_pairs(N)
    win = synth.bootstrap_windows(R=5, n_total=nT, lo=100, hi=600, seed=8)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((len(win), len(ab)), np.float32)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, len(ab), k, dmi)
    got = dmi.numpy()
    for r, (off, ln) in enumerate(win):
        Xw = np.concatenate([X[:, di * nT + off: di * nT + off + ln] for di in range(d)], axis=1)
        want = oracle.pairs(Xw, ab, k, int(ln), d, gc=False)
        np.testing.assert_array_equal(got[r], want)


def test_full_size_symmetry_and_sample():
    """At the metric's size (n=2000, d=3, k=6): MI(a,b) == MI(b,a) bitwise for
    64 pairs, and two pairs equal the oracle."""
    N, n, d, k = 40, 2000, 3, 6
    _flush("full size")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=1)
    rng = np.random.default_rng(0)
    ab = rng.choice(N, size=(64, 2), replace=True).astype(np.int32)
    ab[ab[:, 0] == ab[:, 1], 1] = (ab[ab[:, 0] == ab[:, 1], 1] + 1) % N
    both = np.concatenate([ab, ab[:, ::-1]]).astype(np.int32)
    mi, _, _ = _gpu_counts(X, both, k, n, d)
    np.testing.assert_array_equal(mi[:64], mi[64:])
    for p in range(2):
        a, b = ab[p]
        assert mi[p] == np.float32(oracle.mi(X[a], X[b], k, n, d))


_LONG = {}


def _long_window_reference():
    """Oracle counts for two pairs at n = 9001 (computed once per session)."""
    if not _LONG:
        N, n, d, k = 3, 9001, 3, 6
        X, _ = synth.calpha_trajectory(N=N, n=n, seed=77)
        ab = np.array([[0, 1], [1, 2]], np.int32)
        ref = [oracle.ksg_counts(X[a], X[b], k, n, d)[:2] for a, b in ab]
        mi = [np.float32(oracle.mi_from_counts(na, nb, k, n)) for na, nb in ref]
        _LONG.update(X=X, ab=ab, n=n, d=d, k=k, ref=ref, mi=mi)
    return _LONG


def test_long_window_counts_bitexact():
    """n = 9001 frames (35 LDS tiles of 256 plus a ragged tail, 71 query
    blocks of 128): per-frame counts and MI equal the oracle's."""
    _flush("long window n=9001")
    L = _long_window_reference()
    mi, cnt, _ = _gpu_counts(L["X"], L["ab"], L["k"], L["n"], L["d"])
    for p in range(2):
        np.testing.assert_array_equal(cnt[p, :, 0], L["ref"][p][0])
        np.testing.assert_array_equal(cnt[p, :, 1], L["ref"][p][1])
        assert mi[p] == L["mi"][p]


def test_self_pair_closed_form():
    """Xa == Xb without duplicate frames: n_x = n_y = k for every frame, so
    MI = psi[n] + psi[k] - 1/k - (sum of 2 psi[k] in frame order)/n."""
    n, d, k = 500, 3, 5
    _flush("self pair")
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, d * n)).astype(np.float32)
    mi, cnt, _ = _gpu_counts(X, np.array([[0, 0]], np.int32), k, n, d)
    assert (cnt[0] == k).all()
    psi = oracle.psi(n)
    acc = np.float32(0)
    for _ in range(n):
        acc = np.float32(acc + np.float32(psi[k] + psi[k]))
    acc = np.float32(acc / np.float32(n))
    want = np.float32(np.float32(np.float32(psi[n] + psi[k]) - np.float32(1.0 / k)) - acc)
    assert mi[0] == want


def test_invalid_arguments_raise():
    X = np.zeros((2, 100), np.float32)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(np.array([[0, 1]], np.int32))
    dmi = abi.DeviceArray((1,), np.float32)
    with pytest.raises(RuntimeError, match="k"):
        abi.ksg_mi(dX, 100, 100, 1, 100, dab, 1, 17, dmi)
    with pytest.raises(RuntimeError, match="d = 1, 2 or 3"):
        abi.ksg_mi(dX, 100, 25, 4, 25, dab, 1, 4, dmi)
