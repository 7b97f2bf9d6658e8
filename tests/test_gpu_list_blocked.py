"""The list path when its tables do not all fit, with sorted lists cut
short, and walks whose workgroups retire and restage pairs through their
LDS ring (GPU required).

Every case compares all pairs bitwise with the all-frames kernel (an
independent algorithm over every frame pair) and with the oracle
(oracle/ksg_oracle.cpp, the reference CPU path restated) on all or a seeded
sample of the pairs.
"""
import numpy as np
import pytest

import oracle
from netsci_amd import abi, synth

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _reset():
    abi.set_method(abi.METHOD_AUTO)
    abi.set_table_budget(0)
    abi.set_list_cap(0)
    yield
    abi.set_method(abi.METHOD_AUTO)
    abi.set_table_budget(0)
    abi.set_list_cap(0)


def _slot_bytes(n):
    return abi.table_slot_bytes(n)


def _mi(X, ab, k, n, d, method, counts=False, dstride=None):
    X = np.ascontiguousarray(X, np.float32)
    ab = np.ascontiguousarray(ab, np.int32).reshape(-1, 2)
    P = len(ab)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((P,), np.float32)
    dcnt = abi.DeviceArray((P, n, 2), np.int32) if counts else None
    abi.set_method(method)
    abi.ksg_mi(dX, X.shape[1], dstride or n, d, n, dab, P, k, dmi, dcnt)
    got_method = abi.last_method()[0]
    info = abi.last_list_info()
    abi.set_method(abi.METHOD_AUTO)
    return dmi.numpy(), (dcnt.numpy() if counts else None), got_method, info


def _oracle_sample(X, ab, mi, k, n, d, count, seed):
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(ab), size=min(count, len(ab)), replace=False)
    want = oracle.pairs(X, ab[idx], k, n, d, gc=False)
    np.testing.assert_array_equal(mi[idx], want)


@pytest.mark.parametrize("slots", [2, 3, 5, 8])
def test_blocked_tables_all_pairs(slots):
    """A table budget of a few node slots forces the blocked schedule (blocks
    of slots // 2 nodes, the row block resident while the others stream);
    every pair equals the all-frames kernel and the oracle."""
    N, n, d, k = 13, 300, 3, 6
    _flush(f"blocked tables slots={slots}")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=21 + slots)
    ab = np.concatenate([synth.all_pairs(N), np.array([[4, 4], [12, 0]], np.int32)])
    abi.set_table_budget(slots * _slot_bytes(n))
    got, _, method, (blocks, held, builds) = _mi(X, ab, k, n, d, abi.METHOD_LISTS)
    assert method == 2
    bn = slots // 2
    nb = (N + bn - 1) // bn
    assert held == slots and blocks == nb
    size = [min(bn, N - b * bn) for b in range(nb)]
    assert builds == size[0] + sum(size[J] for I in range(nb) for J in range(I + 1, nb))
    abi.set_table_budget(0)
    ref, _, method, _ = _mi(X, ab, k, n, d, abi.METHOD_ALL_FRAMES)
    assert method == 1
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, oracle.pairs(X, ab, k, n, d, gc=False))


def test_blocked_tables_counts_and_scattered_nodes():
    """Blocked schedule with per-frame counts requested (written through the
    pair permutation) over scattered node indices of a tall X."""
    n, d, k = 256, 3, 4
    _flush("blocked tables with counts")
    X, _ = synth.calpha_trajectory(N=600, n=n, seed=5)
    rng = np.random.default_rng(3)
    nodes = np.sort(rng.choice(600, size=9, replace=False))
    ab = nodes[rng.integers(0, 9, size=(40, 2))].astype(np.int32)
    abi.set_table_budget(4 * _slot_bytes(n))
    got, cnt, method, (blocks, held, _) = _mi(X, ab, k, n, d, abi.METHOD_LISTS, counts=True)
    assert method == 2 and blocks > 1 and held == 4
    for p, (a, b) in enumerate(ab):
        na, nb_, _, _ = oracle.ksg_counts(X[a, : d * n], X[b, : d * n], k, n, d)
        np.testing.assert_array_equal(cnt[p, :, 0], na)
        np.testing.assert_array_equal(cnt[p, :, 1], nb_)
    np.testing.assert_array_equal(got, oracle.pairs(X, ab, k, n, d, gc=False))


def test_blocked_tables_windows():
    """Windows on the blocked schedule: each window rebuilds its blocks."""
    N, nT, d, k = 9, 700, 3, 6
    _flush("blocked tables with windows")
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=14)
    ab = synth.all_pairs(N)
    win = np.array([[0, 700], [50, 333], [300, 400]], np.int32)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((len(win), len(ab)), np.float32)
    abi.set_table_budget(4 * _slot_bytes(700))
    abi.set_method(abi.METHOD_LISTS)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, len(ab), k, dmi)
    assert abi.last_method()[0] == 2 and abi.last_list_info()[0] > 1
    got = dmi.numpy()
    for r, (off, ln) in enumerate(win):
        Xw = np.concatenate([X[:, di * nT + off: di * nT + off + ln] for di in range(d)], axis=1)
        np.testing.assert_array_equal(got[r], oracle.pairs(Xw, ab, k, int(ln), d, gc=False))


def test_no_room_for_two_nodes_falls_back():
    """A budget below two node slots: the automatic choice runs the
    all-frames kernel, METHOD_LISTS raises."""
    N, n, d, k = 4, 200, 3, 6
    _flush("table budget below two nodes")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=8)
    ab = synth.all_pairs(N)
    abi.set_table_budget(_slot_bytes(n))
    got, _, method, info = _mi(X, ab, k, n, d, abi.METHOD_AUTO)
    assert method == 1 and info == (0, 0, 0)
    np.testing.assert_array_equal(got, oracle.pairs(X, ab, k, n, d, gc=False))
    with pytest.raises(RuntimeError, match="two nodes"):
        _mi(X, ab, k, n, d, abi.METHOD_LISTS)


def test_ring_restaging_many_pairs_per_workgroup():
    """1,128 pairs over about 256 walk workgroups with 3 LDS ring buffers:
    each workgroup retires and restages pairs several times (ring slot
    reuse, per-buffer done counts, the barrier-free publish of new pairs)."""
    N, n, d, k = 48, 300, 3, 6
    _flush("ring restaging N=48 n=300")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=48)
    ab = synth.all_pairs(N)
    got, _, method, (blocks, _, _) = _mi(X, ab, k, n, d, abi.METHOD_LISTS)
    assert method == 2 and blocks == 1
    ref, _, _, _ = _mi(X, ab, k, n, d, abi.METHOD_ALL_FRAMES)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, oracle.pairs(X, ab, k, n, d, gc=False))


def test_ring_one_buffer_long_rows():
    """n = 4096 at d = 3: one pair's records fill the LDS ring (one buffer),
    and 276 pairs give some workgroups a second pair to stage."""
    N, n, d, k = 24, 4096, 3, 6
    _flush("ring of one buffer N=24 n=4096")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=24)
    ab = synth.all_pairs(N)
    got, _, method, _ = _mi(X, ab, k, n, d, abi.METHOD_LISTS)
    assert method == 2
    ref, _, _, _ = _mi(X, ab, k, n, d, abi.METHOD_ALL_FRAMES)
    np.testing.assert_array_equal(got, ref)
    _oracle_sample(X, ab, got, k, n, d, 3, seed=1)


def test_ring_restaging_two_windows():
    """Two windows of 1,128 pairs each on the list path (one after another,
    each restaging pairs through the ring)."""
    N, nT, d, k = 48, 500, 3, 5
    _flush("ring restaging, two windows")
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=49)
    ab = synth.all_pairs(N)
    win = np.array([[0, 500], [120, 301]], np.int32)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((len(win), len(ab)), np.float32)
    abi.set_method(abi.METHOD_LISTS)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, len(ab), k, dmi)
    assert abi.last_method()[0] == 2
    got = dmi.numpy()
    abi.set_method(abi.METHOD_ALL_FRAMES)
    abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, len(ab), k, dmi)
    np.testing.assert_array_equal(got, dmi.numpy())
    for r, (off, ln) in enumerate(win):
        Xw = np.concatenate([X[:, di * nT + off: di * nT + off + ln] for di in range(d)], axis=1)
        _oracle_sample(Xw, ab, got[r], k, int(ln), d, 64, seed=r)


def test_tables_persist_between_calls():
    """The tables are a grow-only buffer: a second, smaller call reuses it
    (no reallocation), and releasing the workspaces frees it; results are
    unchanged throughout."""
    N, n, d, k = 10, 400, 3, 6
    _flush("table buffer reuse")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=10)
    ab = synth.all_pairs(N)
    first, _, _, (_, held1, _) = _mi(X, ab, k, n, d, abi.METHOD_LISTS)
    second, _, _, (_, held2, _) = _mi(X, ab[:7], k, n, d, abi.METHOD_LISTS)
    np.testing.assert_array_equal(first[:7], second)
    abi.release_workspace()
    third, _, _, _ = _mi(X, ab, k, n, d, abi.METHOD_LISTS)
    np.testing.assert_array_equal(first, third)
    assert held1 == N and held2 <= N


@pytest.mark.parametrize("n,d,k,cap", [(300, 3, 6, 8), (300, 3, 6, 24), (1000, 2, 4, 16),
                                       (2600, 3, 6, 64), (700, 1, 6, 40)])
def test_cut_lists_resolved_exactly(n, d, k, cap):
    """Sorted lists cut to a few entries: most walks reach the end of their
    cut list before the stopping rule holds and go to the tie kernels, which
    resolve them on the whole row; every pair still equals the default
    tables' result, the all-frames kernel and the oracle, with and without
    per-frame counts."""
    N = 9
    _flush(f"cut lists n={n} d={d} k={k} cap={cap}")
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=70 + cap)
    if d < 3:
        X = np.ascontiguousarray(X[:, :d * n])
    ab = np.concatenate([synth.all_pairs(N), np.array([[2, 2], [8, 1]], np.int32)])
    want, wcnt, m0, _ = _mi(X, ab, k, n, d, abi.METHOD_LISTS, counts=True)
    ties0 = abi.last_tie_frames()
    assert m0 == 2
    abi.set_list_cap(cap)
    got, gcnt, m1, _ = _mi(X, ab, k, n, d, abi.METHOD_LISTS, counts=True)
    ties1 = abi.last_tie_frames()
    assert m1 == 2
    assert ties1 > ties0  # walks cut short went to the tie kernels
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(gcnt, wcnt)
    af, _, m2, _ = _mi(X, ab, k, n, d, abi.METHOD_ALL_FRAMES)
    assert m2 == 1
    np.testing.assert_array_equal(got, af)
    _oracle_sample(X, ab, got, k, n, d, 8, seed=cap)
