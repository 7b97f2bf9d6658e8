"""The product's libstdc++-order selection (netsci_amd/csrc/sort_select.h,
used for GPU tie resolution and the CPU platform) against std::sort itself
(oracle.argsort_like_reference): the first m positions must be identical,
ties included, for every m and input shape."""
import numpy as np
import pytest

import oracle
from netsci_amd import abi


@pytest.mark.parametrize("n", [1, 2, 5, 16, 17, 18, 33, 100, 1000, 5000])
def test_random_with_heavy_ties(n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        levels = max(1, n // (1 + trial % 7))
        key = rng.integers(0, levels, size=n).astype(np.float32)
        want = oracle.argsort_like_reference(key)
        for m in sorted({1, 2, 5, 7, 17, n} & set(range(1, n + 1))):
            np.testing.assert_array_equal(abi.select_prefix(key, m), want[:m])


def test_knn_rows_with_self_zero():
    """Rows shaped like a frame's joint distances: self at 0, duplicates,
    quantised values, selection of k+1 = 7."""
    rng = np.random.default_rng(7)
    for t in range(200):
        n = int(rng.integers(20, 3000))
        key = np.round(rng.exponential(1.0, n) * 8) / 8
        key[rng.integers(0, n)] = 0.0
        key = key.astype(np.float32)
        want = oracle.argsort_like_reference(key)
        np.testing.assert_array_equal(abi.select_prefix(key, 7), want[:7])


def test_sorted_and_reversed_and_constant():
    for n in (50, 513, 4096):
        for key in (np.arange(n, dtype=np.float32), np.arange(n, 0, -1).astype(np.float32),
                    np.zeros(n, np.float32), np.tile(np.array([3, 1, 2], np.float32), n)[:n]):
            want = oracle.argsort_like_reference(key)
            np.testing.assert_array_equal(abi.select_prefix(key, n), want)
            np.testing.assert_array_equal(abi.select_prefix(key, 7), want[:7])


@pytest.mark.parametrize("n", [64, 257, 1000, 4000])
def test_heap_sort_fallback(n):
    """McIlroy's adversary drives introsort past its depth limit; the full
    permutation (and every prefix) must still equal std::sort's."""
    key = oracle.antiqsort(n)
    want = oracle.argsort_like_reference(key)
    got, heaps = abi.select_prefix(key, n, with_heap_count=True)
    assert heaps > 0, "the adversarial input did not reach the heap-sort fallback"
    np.testing.assert_array_equal(got, want)
    for m in (1, 7, 17, n // 2):
        np.testing.assert_array_equal(abi.select_prefix(key, m), want[:m])
