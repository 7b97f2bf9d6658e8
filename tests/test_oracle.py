"""Pinning the oracle (oracle/ksg_oracle.cpp) before it judges anything.

Each check is chosen so that a plausible slip in the restatement — a wrong
psi recurrence, `<` for `<=`, self counted, a dropped term, a transposed
operand in the Hedetniemi recurrence — fails at least one of them.
"""
import numpy as np
import pytest

import oracle
from netsci_amd import synth
from tests.conftest import curve_inputs


def test_golden_known_answers(golden_1000, golden_2000):
    """The reference's own golden values (test_MutualInformation.py:37,64)."""
    assert oracle.mi(golden_1000[:1000], golden_1000[1000:], 4, 1000, 1) == 0.8449974060058594
    assert oracle.mi(golden_2000[:2000], golden_2000[2000:], 4, 2000, 1) == 0.8485465049743652


def test_psi_is_digamma():
    """psi[i] approximates the digamma function psi(i) (psi.cpp:6-19)."""
    from scipy.special import digamma
    p = oracle.psi(5000)
    assert p[0] == 0.0
    i = np.arange(1, 5001)
    np.testing.assert_allclose(p[1:], digamma(i), atol=2e-4)


def _brute_counts(xa, xb, k, n, d):
    """O(n^2) Python restatement with an explicit stable tie order: the k
    nearest are chosen by (distance, index); valid whenever no boundary tie."""
    xa = xa.reshape(d, n).astype(np.float32)
    xb = xb.reshape(d, n).astype(np.float32)
    na, nb = np.zeros(n, int), np.zeros(n, int)
    for i in range(n):
        dxa = np.max(np.abs(xa[:, i:i + 1] - xa), axis=0)
        dxb = np.max(np.abs(xb[:, i:i + 1] - xb), axis=0)
        dz = np.maximum(dxa, dxb)
        order = np.lexsort((np.arange(n), dz))
        sel = order[1:k + 1]
        ea, eb = dxa[sel].max(), dxb[sel].max()
        na[i] = np.sum(dxa <= ea) - 1
        nb[i] = np.sum(dxb <= eb) - 1
    return na, nb


@pytest.mark.parametrize("d,k,n", [(1, 3, 40), (2, 4, 50), (3, 6, 64)])
def test_counts_brute_force(d, k, n):
    rng = np.random.default_rng(d * 100 + k)
    xa = rng.standard_normal(d * n).astype(np.float32)
    xb = (0.6 * xa + rng.standard_normal(d * n)).astype(np.float32)
    na, nb, _, _ = oracle.ksg_counts(xa, xb, k, n, d)
    ba, bb = _brute_counts(xa, xb, k, n, d)
    np.testing.assert_array_equal(na, ba)
    np.testing.assert_array_equal(nb, bb)
    assert (na >= k).all() and (nb >= k).all()  # the k neighbours lie within eps


def test_mi_from_counts_matches_direct():
    rng = np.random.default_rng(1)
    n, d, k = 300, 3, 6
    xa = rng.standard_normal(d * n).astype(np.float32)
    xb = rng.standard_normal(d * n).astype(np.float32)
    na, nb, _, _ = oracle.ksg_counts(xa, xb, k, n, d)
    assert oracle.mi_from_counts(na, nb, k, n) == oracle.mi(xa, xb, k, n, d)


def test_symmetry():
    """MI(a, b) == MI(b, a) bitwise: dz is symmetric and float + commutes."""
    rng = np.random.default_rng(2)
    n, d = 400, 3
    xa = rng.standard_normal(d * n).astype(np.float32)
    xb = (xa + rng.standard_normal(d * n)).astype(np.float32)
    assert oracle.mi(xa, xb, 6, n, d) == oracle.mi(xb, xa, 6, n, d)


def test_self_pair_closed_form():
    """Xa == Xb (no duplicates): every n_x = n_y = k."""
    rng = np.random.default_rng(3)
    n, d, k = 250, 2, 5
    x = rng.standard_normal(d * n).astype(np.float32)
    na, nb, _, _ = oracle.ksg_counts(x, x, k, n, d)
    assert (na == k).all() and (nb == k).all()


def test_gaussian_closed_form():
    """C0-style Gaussian nodes: KSG MI tracks -0.5 ln(1 - rho^2)."""
    X, rho = synth.gaussian_nodes(M=8, n=1000, seed=4)
    ab = synth.all_pairs(8)
    mi = oracle.pairs(X, ab, 4, 1000, 1, gc=False)
    exact = -0.5 * np.log(1 - rho[ab[:, 0], ab[:, 1]] ** 2)
    assert np.abs(mi - exact).max() < 0.08
    assert np.corrcoef(mi, exact)[0, 1] > 0.98


def test_gc_transform():
    """GC = sqrt(1 - exp(-2 I / d)) in double, 0 for I <= 0."""
    assert oracle.gc_from_mi(-0.1, 3) == 0.0
    assert oracle.gc_from_mi(0.0, 3) == 0.0
    for mi, d in [(0.5, 1), (0.2, 3), (1.7, 2)]:
        want = np.float32(np.sqrt(1.0 - np.float64(np.float32(np.exp(-(2.0 / d) * np.float64(np.float32(mi)))))))
        assert oracle.gc_from_mi(mi, d) == want


def test_independent_is_near_zero():
    rng = np.random.default_rng(6)
    n, d = 1000, 3
    mi = oracle.mi(rng.standard_normal(d * n).astype(np.float32),
                   rng.standard_normal(d * n).astype(np.float32), 6, n, d)
    assert abs(mi) < 0.05


def test_curves_deterministic():
    xa, xb = curve_inputs(500, 2, "mi")
    a = oracle.mi(xa, xb, 4, 500, 2)
    b = oracle.mi(xa, xb, 4, 500, 2)
    assert a == b and a > 1.0


# ---- Hedetniemi -----------------------------------------------------------

def _graph(N, density, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 3.0, size=(N, N)).astype(np.float32)
    A[rng.random((N, N)) > density] = np.inf
    np.fill_diagonal(A, 0.0)
    return A


def test_hedetniemi_converges_to_textbook_apsp():
    """At convergence H equals Floyd-Warshall / Dijkstra (scipy, float64)
    within float32 rounding; for asymmetric A the recurrence yields the
    shortest-path matrix up to the transposition it applies each step."""
    from scipy.sparse.csgraph import shortest_path
    A = _graph(40, 0.12, 1)
    S = (A + A.T) / 2  # symmetric, so transposition is invisible
    S = S.astype(np.float32)
    H, _, _ = oracle.hedetniemi(S, 40, 1e-4, want_paths=False)
    ref = shortest_path(np.where(np.isinf(S), 0, S).astype(np.float64), directed=True)
    np.testing.assert_allclose(H, ref, rtol=1e-5)


def test_hedetniemi_recurrence_orientation():
    """One step on an asymmetric 3-node graph, worked by hand:
    H1[i][j] = min_k A[j][k] + A[k][i]."""
    inf = np.inf
    A = np.array([[0, 1, inf], [inf, 0, 2], [5, inf, 0]], np.float32)
    H, Hi, _ = oracle.hedetniemi(A, 1, 0.0, want_hi=True, want_paths=False)
    np.testing.assert_array_equal(Hi[0], A)
    want = np.empty((3, 3), np.float32)
    for i in range(3):
        for j in range(3):
            want[i, j] = min(A[j, k] + A[k, i] for k in range(3))
    np.testing.assert_array_equal(H, want)
    # H1[i][j] is the shortest <= 2-edge walk j -> i in A: 0->1 costs 1,
    # 1->0 only via 2 (A[1][2] + A[2][0] = 7).
    assert H[1, 0] == 1 and H[0, 1] == 7


def test_hedetniemi_paths_are_valid_shortest_paths():
    A = _graph(30, 0.15, 2)
    S = np.minimum(A, A.T)
    L = 30
    H, _, P = oracle.hedetniemi(S, L, 1e-4)
    for i, j in [(0, 7), (3, 19), (12, 29), (5, 5)]:
        path = oracle.recover_path(P, 30, L, i, j)
        assert path[0] == i and path[-1] == j
        if i == j:
            assert list(path) == [i]
            continue
        if not np.isfinite(H[i, j]):
            continue
        w = sum(float(S[path[t], path[t + 1]]) for t in range(len(path) - 1))
        assert abs(w - float(H[i, j])) < 1e-3


def test_corr_to_adj():
    C = np.array([[1.0, 0.5, 0.0], [0.5, 1.0, 0.25], [0.0, 0.25, 1.0]], np.float32)
    A = oracle.corr_to_adj(C)
    np.testing.assert_array_equal(A, np.array([[0, 2, np.inf], [2, 0, 4], [np.inf, 4, 0]],
                                              np.float32))
