"""The sorted-list KSG algorithm restated on the CPU, against the oracle.

netsci_amd/csrc/ksg_walk.h replaces the all-frames k-NN search with a walk
over per-node lists of frames sorted by (marginal distance, frame) and the
marginal counts with upper-bound rank lookups. Its exactness rests on three
facts: a frame already walked in the other list is recognised by a
lexicographic (distance, frame) compare; once max(last walked keys) exceeds
the (k+1)-th smallest joint distance no unwalked frame can enter the k+1
nearest; and n_x + 1 = rank_a[i][j*] for the selected frame j* of largest
marginal distance. This test runs the same steps in numpy (one list block of
8 entries at a time, chosen by the kernel's remaining-entries estimate) and
requires the oracle's per-frame counts (mutual_information_cpu.cpp:22-109),
ties and duplicate frames included. It pins the algorithm, not the kernel;
tests/test_gpu_ksg.py pins the kernel.
"""
import numpy as np
import pytest

import oracle


def _marg(c, i):
    d = np.abs(c[0][i] - c[0])
    for r in c[1:]:
        d = np.maximum(d, np.abs(r[i] - r))
    return d.astype(np.float32)


def _tables(c, n):
    ords, ranks = [], []
    for i in range(n):
        key = _marg(c, i)
        o = np.lexsort((np.arange(n), key))
        ks = key[o]
        r = np.empty(n, np.int64)
        r[o] = np.searchsorted(ks, ks, side="right")
        ords.append(o)
        ranks.append(r)
    return np.array(ords), np.array(ranks)


def _walk_counts(xa, xb, n, d, k, bug=None):
    ca, cb = xa.reshape(d, n), xb.reshape(d, n)
    oa, ra = _tables(ca, n)
    ob, rb = _tables(cb, n)
    nx = np.zeros(n, np.int64)
    ny = np.zeros(n, np.int64)
    for i in range(n):
        DA, DB = _marg(ca, i), _marg(cb, i)
        h = [np.float32(np.inf)] * (k + 1)
        ids = [0] * (k + 1)
        last = {"a": (-1.0, -1), "b": (-1.0, -1)}  # (key, frame); -1: none yet
        pos = {"a": 0, "b": 0}
        done = False
        while not done:
            # list choice per 8-entry block, as the kernel
            if pos["a"] >= n:
                L = "b"
            elif pos["b"] >= n:
                L = "a"
            elif pos["a"] < 8 or pos["b"] < 8 or not np.isfinite(h[k]):
                L = "a" if pos["a"] <= pos["b"] else "b"
            else:
                with np.errstate(over="ignore"):
                    ea = pos["a"] * ((h[k] / max(last["a"][0], 1e-30)) ** d - 1)
                    eb = pos["b"] * ((h[k] / max(last["b"][0], 1e-30)) ** d - 1)
                L = "a" if ea <= eb else "b"
            O = "b" if L == "a" else "a"
            order = oa if L == "a" else ob
            for _ in range(8):
                if pos[L] >= n:
                    break
                j = order[i, pos[L]]
                pos[L] += 1
                kL, kO = (DA[j], DB[j]) if L == "a" else (DB[j], DA[j])
                seen = (kO, j) <= last[O] if last[O][1] >= 0 else False
                if bug == "no_seen":
                    seen = False
                elif bug == "key_only_seen":
                    seen = kO <= last[O][0] if last[O][1] >= 0 else False
                last[L] = (kL, j)
                if not seen and j != i:
                    z = max(DA[j], DB[j])
                    if z < h[k]:
                        m = k
                        while m > 0 and z < h[m - 1]:
                            h[m], ids[m] = h[m - 1], ids[m - 1]
                            m -= 1
                        h[m], ids[m] = z, j
                keys = [v[0] for v in last.values() if v[1] >= 0]
                if max(keys) > h[k] or (pos["a"] >= n and pos["b"] >= n):
                    done = True
                    break
        if h[k - 1] == h[k]:  # exact tie at the k-th distance: std::sort order
            sel = oracle.argsort_like_reference(np.maximum(DA, DB))[1:k + 1]
        else:
            sel = np.array(ids[:k])
        jx = sel[int(np.argmax(DA[sel]))]
        jy = sel[int(np.argmax(DB[sel]))]
        nx[i] = ra[i, jx] - 1
        ny[i] = rb[i, jy] - 1
    return nx, ny


def _case(kind, d, n, rng):
    if kind == "gauss":
        xa = rng.standard_normal(d * n)
        xb = 0.5 * xa + rng.standard_normal(d * n)
    elif kind == "grid":  # heavy ties
        xa = rng.integers(0, 4, d * n)
        xb = rng.integers(0, 3, d * n)
    elif kind == "dups":  # duplicate frames
        base = rng.standard_normal((d, n // 3 + 1))
        pick = rng.integers(0, base.shape[1], n)
        xa = base[:, pick].ravel()
        xb = 2 * base[:, pick].ravel()
    else:  # one coarse, one fine
        xa = rng.integers(0, 6, d * n) * 0.25
        xb = rng.standard_normal(d * n)
    return xa.astype(np.float32), xb.astype(np.float32)


@pytest.mark.parametrize("kind", ["gauss", "grid", "dups", "mixed"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_walk_counts_equal_oracle(kind, d):
    rng = np.random.default_rng(hash((kind, d)) % 2**32)
    for _ in range(3):
        n = int(rng.integers(9, 90))
        k = int(rng.integers(1, min(8, n - 1) + 1))
        xa, xb = _case(kind, d, n, rng)
        nx, ny = _walk_counts(xa, xb, n, d, k)
        ox, oy, _, _ = oracle.ksg_counts(xa, xb, k, n, d)
        np.testing.assert_array_equal(nx, ox)
        np.testing.assert_array_equal(ny, oy)


@pytest.mark.parametrize("bug,kind", [("no_seen", "dups"), ("key_only_seen", "grid")])
def test_planted_mistakes_are_caught(bug, kind):
    """A restatement that inserts a frame walked in both lists twice (no
    seen test), or treats every frame whose key equals the other list's last
    key as walked (no frame tie-break), must disagree with the oracle on
    duplicate-frame or tie-heavy inputs, so the comparison above can fail."""
    rng = np.random.default_rng(11)
    mismatched = 0
    for _ in range(4):
        d, n, k = 2, 60, 4
        xa, xb = _case(kind, d, n, rng)
        nx, ny = _walk_counts(xa, xb, n, d, k, bug=bug)
        ox, oy, _, _ = oracle.ksg_counts(xa, xb, k, n, d)
        mismatched += int(not (np.array_equal(nx, ox) and np.array_equal(ny, oy)))
    assert mismatched > 0
