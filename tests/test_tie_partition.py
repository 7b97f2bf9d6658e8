"""The wave-parallel partition of the GPU tie resolution, restated on the CPU.

netsci_amd/csrc/ksg_walk.h (partition_wave, sort_prefix_wave) runs
libstdc++'s unguarded Hoare partition with one wave: it lists the positions
where the left scan stops (not less than the pivot) and where the right scan
stops (not greater than the pivot, the pivot included), pairs the t-th left
stop with the t-th right stop counted from the right, does the swaps before
the first pair that has crossed at once, and returns min(left stop, previous
right stop) of that pair. This test restates those steps in numpy and
requires (a) the same array and cut as the sequential partition on
tie-heavy and random inputs, and (b) through the same sub-range steps as
sort_prefix_wave (parallel partitions above 128 records), the first m
positions std::sort itself leaves (oracle.argsort_like_reference), the
heap-sort fallback included. A planted mistake (right stops counted from
the left) must fail.
"""
import numpy as np
import pytest

import oracle


def _seq_partition(key, idx, f1, l):
    """libstdc++ __unguarded_partition(first=f1, last=l, pivot=f1-1)."""
    p = key[f1 - 1]
    f = f1
    while True:
        while key[f] < p:
            f += 1
        l -= 1
        while p < key[l]:
            l -= 1
        if not f < l:
            return f
        key[f], key[l] = key[l], key[f]
        idx[f], idx[l] = idx[l], idx[f]
        f += 1


def _wave_partition(key, idx, f1, l, bug=None):
    p = key[f1 - 1]
    pos = np.arange(f1 - 1, l)
    v = key[f1 - 1:l]
    lasc = pos[(pos >= f1) & ~(v < p)]
    rasc = pos[~(p < v)]
    nl, nr = len(lasc), len(rasc)
    T = 1
    while True:
        rs_t = rasc[T - 1] if bug == "right_from_left" else (rasc[nr - T] if T <= nr else None)
        if T > nl or T > nr or lasc[T - 1] >= rs_t:
            break
        T += 1
    for t in range(1, T):
        pl = lasc[t - 1]
        pr = rasc[t - 1] if bug == "right_from_left" else rasc[nr - t]
        key[pl], key[pr] = key[pr], key[pl]
        idx[pl], idx[pr] = idx[pr], idx[pl]
    ls = lasc[T - 1] if T <= nl else 1 << 30
    prev = (rasc[T - 2] if bug == "right_from_left" else rasc[nr - (T - 1)]) if T >= 2 else l
    return min(ls, prev)


def _median_to_first(key, idx, res, x, y, z):
    def sw(a):
        key[res], key[a] = key[a], key[res]
        idx[res], idx[a] = idx[a], idx[res]
    if key[x] < key[y]:
        if key[y] < key[z]:
            sw(y)
        elif key[x] < key[z]:
            sw(z)
        else:
            sw(x)
    elif key[x] < key[z]:
        sw(x)
    elif key[y] < key[z]:
        sw(z)
    else:
        sw(y)


_HEAP_RANGES = [0]


def _prefix_wave(key, m, bug=None):
    """sort_prefix_wave's steps; returns the frame indices of positions [0, m)."""
    key = key.astype(np.float32).copy()
    idx = np.arange(len(key))
    n = len(key)
    m = min(m, n)
    bound = n
    stack = [(0, n, 2 * (n.bit_length() - 1))]
    while stack:
        f, l, depth = stack.pop()
        while l - f > 16:
            if depth == 0:
                # the heap-sort fallback sorts the whole sub-range; std::sort's
                # partial_sort order is what oracle.heapsort_like_libstdcxx gives
                sub = oracle.heapsort_like_libstdcxx(key[f:l])
                _HEAP_RANGES[0] += 1
                key[f:l], idx[f:l] = key[f:l][sub], idx[f:l][sub]
                break
            depth -= 1
            mid = f + (l - f) // 2
            _median_to_first(key, idx, f, f + 1, mid, l - 1)
            if l - f > 128:
                cut = _wave_partition(key, idx, f + 1, l, bug)
            else:
                cut = _seq_partition(key, idx, f + 1, l)
            if cut >= m:
                bound = min(bound, cut)
            else:
                stack.append((cut, l, depth))
            l = cut
    # final insertion sort of [0, bound), strict <
    for i in range(1, bound):
        kv, iv = key[i], idx[i]
        j = i
        while j > 0 and kv < key[j - 1]:
            key[j], idx[j] = key[j - 1], idx[j - 1]
            j -= 1
        key[j], idx[j] = kv, iv
    return idx[:m]


def _cases(rng, n):
    yield rng.integers(0, 3, n).astype(np.float32)
    yield rng.integers(0, max(1, n // 10), n).astype(np.float32)
    yield rng.standard_normal(n).astype(np.float32)
    yield np.zeros(n, np.float32)
    yield np.arange(n, dtype=np.float32)
    yield np.arange(n, 0, -1).astype(np.float32)


@pytest.mark.parametrize("n", [3, 17, 129, 130, 300, 1000, 2000])
def test_wave_partition_equals_sequential(n):
    rng = np.random.default_rng(n)
    for key in _cases(rng, n):
        # pivot at 0 after a median-of-three, as the sort calls it
        k1, i1 = key.copy(), np.arange(n)
        if n >= 3:
            _median_to_first(k1, i1, 0, 1, n // 2, n - 1)
        k2, i2 = k1.copy(), i1.copy()
        if n < 3:
            continue
        c1 = _seq_partition(k1, i1, 1, n)
        c2 = _wave_partition(k2, i2, 1, n)
        assert c1 == c2
        np.testing.assert_array_equal(i1, i2)


@pytest.mark.parametrize("n", [100, 500, 2000, 5000])
def test_wave_prefix_equals_std_sort(n):
    rng = np.random.default_rng(100 + n)
    for key in list(_cases(rng, n)) + [oracle.antiqsort(n)]:
        want = oracle.argsort_like_reference(key)
        for m in (1, 7, 17):
            np.testing.assert_array_equal(_prefix_wave(key, m), want[:m])


@pytest.mark.parametrize("n", [257, 1000])
def test_wave_prefix_heap_fallback(n):
    """McIlroy's adversary over the whole array (m = n) reaches the depth
    limit's heap-sort fallback."""
    key = oracle.antiqsort(n)
    want = oracle.argsort_like_reference(key)
    before = _HEAP_RANGES[0]
    np.testing.assert_array_equal(_prefix_wave(key, n), want)
    assert _HEAP_RANGES[0] > before


def test_planted_mistake_is_caught():
    rng = np.random.default_rng(3)
    bad = 0
    for _ in range(10):
        key = rng.integers(0, 5, 1000).astype(np.float32)
        want = oracle.argsort_like_reference(key)
        try:
            got = _prefix_wave(key, 7, bug="right_from_left")
            bad += int(not np.array_equal(got, want[:7]))
        except (IndexError, ValueError):
            bad += 1
    assert bad > 0
