// Tie-exact k-nearest selection: the order libstdc++'s std::sort leaves in the
// first m positions of an array, computed without sorting the whole array.
//
// The reference picks a frame's k nearest neighbours by std::sort-ing an iota
// index array by joint distance with `<` and reading positions 1..k
// (netcalc/src/mutual_information_cpu.cpp:48-77). std::sort is unstable, so
// when the k-th and (k+1)-th distances tie, which frame is chosen depends on
// libstdc++'s introsort. This header runs the same introsort steps (GCC's
// bits/stl_algo.h and bits/stl_heap.h: depth limit 2*floor(log2 n), median of
// three moved to the first slot, unguarded Hoare partition, threshold 16,
// heap-sort fallback, final insertion sort) on (key, index) records, but only
// descends into sub-ranges that can still reach positions [0, m). Two facts
// make the pruning exact:
//   * after a partition at cut c every record left of c compares not-greater
//     than every record right of c, so the final insertion sort never moves a
//     record across c;
//   * sub-ranges are independent, so skipping one does not change another.
// The result for positions [0, m) is therefore identical to a full std::sort.
//
// Compiled for the host (product CPU platform) and for gfx950 (GPU tie
// resolution). The algorithm's steps have to match libstdc++'s to give the
// same permutation; the code is written from the algorithm's description.
#pragma once

#if defined(__HIPCC__)
#define NSX_HD __host__ __device__
#else
#define NSX_HD
#endif

namespace nsx {

struct KeyIdx {
    float key;
    int idx;
};

NSX_HD inline bool kless(const KeyIdx &a, const KeyIdx &b) { return a.key < b.key; }

NSX_HD inline void kswap(KeyIdx *a, long x, long y) {
    KeyIdx t = a[x];
    a[x] = a[y];
    a[y] = t;
}

NSX_HD inline int floor_log2(long n) {
    int r = 0;
    while (n > 1) {
        n >>= 1;
        ++r;
    }
    return r;
}

// Move the median of a[x], a[y], a[z] into a[res] (libstdc++ decision tree).
NSX_HD inline void median_to_first(KeyIdx *a, long res, long x, long y, long z) {
    if (kless(a[x], a[y])) {
        if (kless(a[y], a[z])) kswap(a, res, y);
        else if (kless(a[x], a[z])) kswap(a, res, z);
        else kswap(a, res, x);
    } else if (kless(a[x], a[z])) {
        kswap(a, res, x);
    } else if (kless(a[y], a[z])) {
        kswap(a, res, z);
    } else {
        kswap(a, res, y);
    }
}

// Unguarded Hoare partition of [f, l) around the record at pivot.
NSX_HD inline long partition_unguarded(KeyIdx *a, long f, long l, long pivot) {
    while (true) {
        while (kless(a[f], a[pivot])) ++f;
        --l;
        while (kless(a[pivot], a[l])) --l;
        if (!(f < l)) return f;
        kswap(a, f, l);
        ++f;
    }
}

// Sift a value down from hole `hole` then back up (heap of length len at a).
NSX_HD inline void heap_adjust(KeyIdx *a, long hole, long len, KeyIdx v) {
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (kless(a[child], a[child - 1])) child--;
        a[hole] = a[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[hole] = a[child - 1];
        hole = child - 1;
    }
    long parent = (hole - 1) / 2;
    while (hole > top && kless(a[parent], v)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = v;
}

// The introsort fallback: std::partial_sort(f, l, l) = make_heap + sort_heap.
NSX_HD inline void heap_sort_range(KeyIdx *a, long len) {
    if (len < 2) return;
    for (long parent = (len - 2) / 2;; --parent) {
        KeyIdx v = a[parent];
        heap_adjust(a, parent, len, v);
        if (parent == 0) break;
    }
    for (long last = len - 1; last > 0; --last) {
        KeyIdx v = a[last];
        a[last] = a[0];
        heap_adjust(a, 0, last, v);
    }
}

// Insertion sort of [0, l) with strict `<` (what libstdc++'s final
// insertion sort does to each block between partition cuts).
NSX_HD inline void insertion_sort(KeyIdx *a, long l) {
    for (long i = 1; i < l; ++i) {
        KeyIdx v = a[i];
        long j = i;
        while (j > 0 && kless(v, a[j - 1])) {
            a[j] = a[j - 1];
            --j;
        }
        a[j] = v;
    }
}

// Arrange a[0..n) so that a[0..m) equals what std::sort(a, a + n, kless)
// puts there. Records beyond the returned bound are left partially ordered.
// Returns the bound B (m <= B <= n) up to which the prefix is final;
// heap_ranges (nullable) counts sub-ranges that took the heap-sort fallback.
// Needs no memory besides `a`; the explicit stack holds pending sub-ranges.
NSX_HD inline long sort_prefix(KeyIdx *a, long n, long m, int *heap_ranges = nullptr) {
    if (n <= 0) return 0;
    if (m > n) m = n;
    long bound = n;  // smallest partition cut >= m seen so far
    const int kThreshold = 16;
    const int kStack = 64;
    long st_f[kStack], st_l[kStack];
    int st_d[kStack];
    int sp = 0;
    st_f[sp] = 0;
    st_l[sp] = n;
    st_d[sp] = 2 * floor_log2(n);
    ++sp;
    while (sp > 0) {
        --sp;
        long f = st_f[sp], l = st_l[sp];
        int depth = st_d[sp];
        while (l - f > kThreshold) {
            if (depth == 0) {
                heap_sort_range(a + f, l - f);
                if (heap_ranges) ++*heap_ranges;
                break;
            }
            --depth;
            long mid = f + (l - f) / 2;
            median_to_first(a, f, f + 1, mid, l - 1);
            long cut = partition_unguarded(a, f + 1, l, f);
            if (cut >= m) {
                if (cut < bound) bound = cut;
            } else if (sp < kStack) {
                // The right part still reaches [0, m): process it too.
                st_f[sp] = cut;
                st_l[sp] = l;
                st_d[sp] = depth;
                ++sp;
            }
            l = cut;
        }
    }
    insertion_sort(a, bound);
    return bound;
}

}  // namespace nsx
