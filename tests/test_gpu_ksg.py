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


@pytest.fixture(autouse=True, params=["lists", "pass2", "sympass2"])
def method_variant(request):
    """Every test runs on the sorted-list path (where it applies) and on the
    all-frames kernel with the plain pass 2 (each ordered frame pair) and the
    symmetric one (each unordered pair once, both frames credited); all must
    give the oracle's bits."""
    if request.param == "lists":
        abi.set_method(abi.METHOD_AUTO)
    else:
        abi.set_method(abi.METHOD_ALL_FRAMES)
        abi.set_config(2, 128, 4, 0, int(request.param == "sympass2"))
    yield request.param
    abi.set_method(abi.METHOD_AUTO)
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
    with pytest.raises(RuntimeError, match="more than k frames"):
        abi.ksg_mi(dX, 100, 100, 1, 100, dab, 1, 100, dmi)
    with pytest.raises(RuntimeError, match="d = 1, 2 or 3"):
        abi.ksg_mi(dX, 100, 25, 4, 25, dab, 1, 4, dmi)


# ---- sorted-list path specifics ------------------------------------------------


def _check_pairs(X, ab, k, n, d, expect_method=None):
    mi, cnt, ties = _gpu_counts(X, ab, k, n, d)
    if expect_method is not None:
        assert abi.last_method()[0] == expect_method
    for p, (a, b) in enumerate(ab):
        na, nb, _, _ = oracle.ksg_counts(X[a, : d * n], X[b, : d * n], k, n, d)
        np.testing.assert_array_equal(cnt[p, :, 0], na)
        np.testing.assert_array_equal(cnt[p, :, 1], nb)
        assert mi[p] == np.float32(oracle.mi(X[a, : d * n], X[b, : d * n], k, n, d))
    return ties


@pytest.mark.parametrize("n,k", [(7, 6), (8, 6), (9, 4), (16, 15), (17, 16), (511, 6), (512, 6),
                                 (513, 6), (1024, 3), (1025, 6), (2047, 6), (2048, 6)])
def test_list_path_sizes(method_variant, n, k):
    """Row strides and sort widths at their boundaries (8, 16 or 32 keys per
    lane; 8-entry list blocks), the smallest n = k + 1 and the largest n."""
    d = 3
    _flush(f"list sizes n={n} k={k} ({method_variant})")
    X3, _ = synth.calpha_trajectory(N=4, n=n, seed=100 + n)
    ab = np.array([[0, 1], [2, 3], [1, 1]], np.int32)
    _check_pairs(X3, ab, k, n, d, 2 if method_variant == "lists" else 1)


def test_list_path_limits_and_fallback(method_variant):
    """Above 8192 frames, or with a non-finite coordinate, the automatic
    choice runs the all-frames kernel; METHOD_LISTS raises instead."""
    if method_variant != "lists":
        pytest.skip("method selection is exercised in the lists variant")
    d, k = 1, 4
    _flush("list fallback n=8193")
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2, 8193)).astype(np.float32)
    ab = np.array([[0, 1]], np.int32)
    _gpu_counts(X, ab, k, 8193, d)
    assert abi.last_method()[0] == 1
    abi.set_method(abi.METHOD_LISTS)
    with pytest.raises(RuntimeError, match="8192"):
        _gpu_counts(X, ab, k, 8193, d)
    _flush("list fallback non-finite")
    d, k = 3, 6
    X3, _ = synth.calpha_trajectory(N=2, n=300, seed=6)
    X3[1, 17] = np.inf
    with pytest.raises(RuntimeError, match="finite"):
        _gpu_counts(X3, np.array([[0, 1]], np.int32), k, 300, d)
    abi.set_method(abi.METHOD_AUTO)
    mi, _, _ = _gpu_counts(X3, np.array([[0, 0]], np.int32), k, 300, d)
    assert abi.last_method()[0] == 2  # node 1 (the non-finite row) is not used
    mi, _, _ = _gpu_counts(X3, np.array([[0, 1]], np.int32), k, 300, d)
    assert abi.last_method()[0] == 1


@pytest.mark.parametrize("d,n,k", [(3, 2049, 6), (3, 2055, 6), (3, 3000, 4), (3, 4096, 6),
                                   (3, 4097, 6), (3, 5000, 6), (3, 6000, 3), (2, 7000, 6),
                                   (1, 8192, 4)])
def test_list_path_wide_tables(method_variant, d, n, k):
    """Rows sorted by 2 or 4 waves together (n > 2048): the cross-wave merges,
    ranks whose equal-key runs cross a wave boundary, and one pair per
    workgroup in LDS."""
    if method_variant != "lists":
        pytest.skip("wide tables only exist on the list path")
    _flush(f"wide tables d={d} n={n} k={k}")
    rng = np.random.default_rng(n + d)
    if d == 3:
        X, _ = synth.calpha_trajectory(N=3, n=n, seed=n)
    else:  # a coarse grid: many equal keys, runs across wave blocks
        X = (np.round(rng.standard_normal((3, d * n)) * 6) / 6).astype(np.float32)
    ab = np.array([[0, 1], [1, 2]], np.int32)
    _check_pairs(X, ab, k, n, d, 2)


def test_list_path_long_windows(method_variant):
    """Bootstrap windows of uneven lengths up to 4100 frames go through the
    list path one window at a time (each with its own tables); every window
    equals the oracle on its frames."""
    if method_variant != "lists":
        pytest.skip("windows on the list path")
    N, nT, d, k = 3, 4200, 3, 6
    _flush("long windows on the list path")
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=31)
    ab = synth.all_pairs(N)
    win = np.array([[0, 4100], [100, 2500], [7, 3001], [4000, 200]], np.int32)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((len(win), len(ab)), np.float32)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, len(ab), k, dmi)
    assert abi.last_method()[0] == 2
    got = dmi.numpy()
    for r, (off, ln) in enumerate(win):
        Xw = np.concatenate([X[:, di * nT + off: di * nT + off + ln] for di in range(d)], axis=1)
        want = oracle.pairs(Xw, ab, k, int(ln), d, gc=False)
        np.testing.assert_array_equal(got[r], want)


@pytest.mark.parametrize("waves,batch", [(1, 1), (2, 64), (8, 16), (4, 3)])
def test_list_path_walk_shapes(method_variant, waves, batch):
    """Waves per walk workgroup and the refill batch change the schedule,
    never the bits; heavy ties exercise the cooperative tie path."""
    if method_variant != "lists":
        pytest.skip("walk shapes only apply to the list path")
    n, d, k = 600, 3, 6
    _flush(f"walk shape waves={waves} batch={batch}")
    rng = np.random.default_rng(41)
    X = (np.round(rng.standard_normal((5, d * n)) * 3) / 3).astype(np.float32)
    ab = np.array([[0, 1], [2, 3], [4, 0], [3, 3]], np.int32)
    abi.set_method(abi.METHOD_LISTS, waves, batch)
    ties = _check_pairs(X, ab, k, n, d, 2)
    assert ties > 0


def test_list_path_tie_overflow(method_variant):
    """More tied queries than the tie list holds (at least 1024 entries, or
    1/32 of the queries): the rest are marked in the counts and resolved by
    the overflow pass, with the same bits."""
    if method_variant != "lists":
        pytest.skip("the tie list only exists on the list path")
    n, d, k = 400, 3, 6
    _flush("tie list overflow")
    rng = np.random.default_rng(43)
    X = np.round(rng.standard_normal((4, d * n))).astype(np.float32)
    ab = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [0, 2], [1, 3]], np.int32)
    ties = _check_pairs(X, ab, k, n, d, 2)
    assert ties > 1024


def test_list_path_extreme_spans(method_variant):
    """The walk's heap codes sit in a window below each pair's coordinate
    span: a node with constant coordinates (span 0, every marginal distance
    0), spans of 1e20 and 1e-20, and pairs mixing them must still give the
    oracle's counts (code collisions go to the exact tie resolution)."""
    if method_variant != "lists":
        pytest.skip("heap codes only exist on the list path")
    n, d, k = 300, 3, 4
    _flush("extreme spans")
    rng = np.random.default_rng(47)
    X = rng.standard_normal((5, d * n)).astype(np.float32)
    X[0] = 1.5
    X[2] *= np.float32(1e20)
    X[3] *= np.float32(1e-20)
    X[4] = np.round(X[4] * 2) / 2
    ab = np.array([[0, 1], [1, 2], [2, 3], [0, 3], [1, 3], [3, 4], [0, 4], [2, 2]], np.int32)
    _check_pairs(X, ab, k, n, d, 2)


def test_list_path_sparse_node_indices(method_variant):
    """Pairs over a few scattered rows of a tall X: the tables cover only the
    used nodes (slot map), and entries walked are reported."""
    if method_variant != "lists":
        pytest.skip("node slots only exist on the list path")
    n, d, k = 256, 3, 6
    _flush("sparse node indices")
    X, _ = synth.calpha_trajectory(N=1200, n=n, seed=12)
    ab = np.array([[1199, 3], [3, 700], [700, 1199], [5, 5]], np.int32)
    _check_pairs(X, ab, k, n, d, 2)
    method, entries = abi.last_method()
    assert method == 2 and 0 < entries <= len(ab) * n * n


@pytest.mark.parametrize("d", [1, 2])
def test_list_path_low_dimensions(method_variant, d):
    """d = 1 and d = 2 with ties (coarse grid) on every variant."""
    n, k = 700, 4
    _flush(f"low dimensions d={d} ({method_variant})")
    rng = np.random.default_rng(60 + d)
    X = (np.round(rng.standard_normal((3, d * n)) * 8) / 8).astype(np.float32)
    ab = np.array([[0, 1], [1, 2], [2, 0]], np.int32)
    _check_pairs(X, ab, k, n, d, 2 if method_variant == "lists" else 1)
