"""The sorted-list KSG algorithm restated on the CPU, against the oracle.

netsci_amd/csrc/ksg_walk.h replaces the all-frames k-NN search with a walk
over one per-node list of frames sorted by (marginal distance, frame) and
the marginal counts with upper-bound rank lookups. Its exactness rests on
three facts: the joint distance is at least either marginal distance, so
once the walked list's last key exceeds the (k+1)-th smallest joint
distance no unwalked frame can enter the k+1 nearest (whichever list is
walked); the query frame itself is skipped; and n_x + 1 = rank_a[i][j*] for
the selected frame j* of largest marginal distance. The kernel keeps its k+1
nearest as 32-bit entries (distance code << id_bits | frame), the code a
monotone, lossy function of the distance's float bits in a window below the
pair's coordinate span; walks whose k-th and (k+1)-th codes are equal go to
the tie resolution, which separates code collisions from exact ties. This
test runs the same steps in numpy (one list block of 8 entries at a time,
the list chosen by the kernel's prior, or forced to either node) and
requires the oracle's per-frame counts (mutual_information_cpu.cpp:22-109),
ties and duplicate frames included, with the kernel's code step and with a
far coarser one that makes most walks collide. It pins the algorithm, not
the kernel; tests/test_gpu_ksg.py pins the kernel.
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


def _bits(x):
    return int(np.float32(x).view(np.uint32))


def _id_bits(n):
    b = 1
    while (1 << b) < n:
        b += 1
    return b


def _prior_pos(n):
    # ksg_walk.h prior_pos(n, build_width(n)): 64, or for rows of at most 64
    # frames the last multiple of 8 below n.
    return 64 if n > 64 else ((n - 1) // 8) * 8


def _walk_counts(xa, xb, n, d, k, bug=None, shift=None, choice="prior"):
    ca, cb = xa.reshape(d, n), xb.reshape(d, n)
    oa, ra = _tables(ca, n)
    ob, rb = _tables(cb, n)
    idb = _id_bits(n)
    if shift is None:
        shift = max(idb - 5, 0)
    span = max(float(np.max(np.float32(r.max()) - np.float32(r.min()))) for r in (*ca, *cb))
    window = 1 << (32 - idb + shift)
    base = max(_bits(span) - (window - 1), 0)
    mask = (1 << idb) - 1
    sentinel = 0xFFFFFFFF
    pp = _prior_pos(n)

    def code(key):
        return max(_bits(key) - base, 0) >> shift

    nx = np.zeros(n, np.int64)
    ny = np.zeros(n, np.int64)
    for i in range(n):
        DA, DB = _marg(ca, i), _marg(cb, i)
        # The list: node a's when its key at the prior position is at least
        # node b's (the kernel's choice), or forced.
        if choice == "prior":
            use_a = DA[oa[i, pp]] >= DB[ob[i, pp]]
        else:
            use_a = choice == "a"
        order, DL, DO = (oa, DA, DB) if use_a else (ob, DB, DA)
        h = [sentinel] * (k + 1)
        pos = 0
        while True:
            blk = order[i, pos:pos + 8]
            for j in blk:
                if j != i or bug == "self_inserted":
                    c = (code(max(DA[j], DB[j])) << idb) | int(j)
                    h = sorted(h + [c])[:k + 1]
            # the stopping rule once per block, on the block's last key, as
            # the kernel; a list walked to its end has evaluated every frame
            key7 = (DO if bug == "stop_on_other_key" else DL)[blk[-1]]
            pos += len(blk)
            if pos >= n or (code(key7) << idb) > h[k]:
                break
        if (h[k - 1] ^ h[k]) >> idb == 0 and bug != "collisions_as_distinct":
            # Tie resolution: exact k-th / (k+1)-th distances over the other
            # frames; equal ones take the std::sort order.
            z = np.maximum(DA, DB)
            zo = np.sort(np.delete(z, i))
            if zo[k - 1] < zo[k]:
                sel = np.array([j for j in range(n) if j != i and z[j] <= zo[k - 1]])
            else:
                sel = oracle.argsort_like_reference(z)[1:k + 1]
        else:
            sel = np.array([v & mask for v in h[:k]])
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


@pytest.mark.parametrize("coarse", [False, True], ids=["code", "coarse_code"])
@pytest.mark.parametrize("kind", ["gauss", "grid", "dups", "mixed"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_walk_counts_equal_oracle(kind, d, coarse):
    rng = np.random.default_rng(hash((kind, d)) % 2**32)
    for t in range(3):
        n = int(rng.integers(9, 90))
        k = int(rng.integers(1, min(8, n - 1) + 1))
        xa, xb = _case(kind, d, n, rng)
        ox, oy, _, _ = oracle.ksg_counts(xa, xb, k, n, d)
        # the kernel's choice, and either list forced: any list is exact
        for choice in ("prior", "a", "b") if t == 0 else ("prior",):
            # coarse: 2^19 float-bit units per code step, 1/16 binade
            nx, ny = _walk_counts(xa, xb, n, d, k, shift=19 if coarse else None, choice=choice)
            np.testing.assert_array_equal(nx, ox, err_msg=choice)
            np.testing.assert_array_equal(ny, oy, err_msg=choice)


def test_code_window_holds_every_key():
    """No key's code exceeds the code bits (it would spill into the frame
    bits): keys are at most the pair's span, including the extreme case of
    coordinates at both ends of the float range's finite half."""
    rng = np.random.default_rng(5)
    for n in (9, 100, 2000, 2048, 2049, 8192):
        idb = _id_bits(n)
        shift = max(idb - 5, 0)
        for scale in (1e-30, 1.0, 1e30):
            x = (rng.standard_normal(64) * scale).astype(np.float32)
            span = np.float32(x.max()) - np.float32(x.min())
            base = max(_bits(span) - ((1 << (32 - idb + shift)) - 1), 0)
            keys = np.abs(x[:, None] - x[None, :]).astype(np.float32).ravel()
            codes = np.maximum(keys.view(np.uint32).astype(np.int64) - base, 0) >> shift
            assert codes.max() < (1 << (32 - idb))


@pytest.mark.parametrize("bug,kind", [("self_inserted", "gauss"), ("stop_on_other_key", "mixed"),
                                      ("collisions_as_distinct", "gauss")])
def test_planted_mistakes_are_caught(bug, kind):
    """A restatement that inserts the query frame itself, stops on the
    other node's distance of the block's last entry instead of the walked
    list's key, or takes the k entries of a walk whose k-th and (k+1)-th
    codes collide as the k nearest, must disagree with the oracle on
    ordinary, mixed-scale or coarse-code inputs, so the comparison above can
    fail."""
    rng = np.random.default_rng(11)
    mismatched = 0
    for _ in range(4):
        d, n, k = 2, 60, 4
        xa, xb = _case(kind, d, n, rng)
        nx, ny = _walk_counts(xa, xb, n, d, k, bug=bug,
                              shift=19 if bug == "collisions_as_distinct" else None)
        ox, oy, _, _ = oracle.ksg_counts(xa, xb, k, n, d)
        mismatched += int(not (np.array_equal(nx, ox) and np.array_equal(ny, oy)))
    assert mismatched > 0
