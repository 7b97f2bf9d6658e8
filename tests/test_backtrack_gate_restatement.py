"""The gated Hedetniemi backtrack restated on the CPU, against the oracle.

netsci_amd/csrc/apsp.hip's sparse backtrack takes, for each (i, j) and p
from L-1 down to 0, the first q != j of the candidate list cand(k) =
{q : A[q][k] < +inf} (ascending) with fl(fl(Hi[p][i][q] + A[q][k]) - len)
< tol (hedetniemi_gpu.cu:86-104). When every diagonal entry of A is +0.0,
one Jacobi step gives H_{p+1}[i][j] <= H_p[j][i] (k = j is a candidate with
weight 0), so H_{p+2} <= H_p entry by entry and
    LB[i][q] = min(H_{L-1}[i][q], H_{L-2}[i][q])
bounds every snapshot H_0 .. H_{L-1} from below (NaN entries of H_0 pass no
test and need no bound). Rounding is monotone, so
    M_i[k] = min over q < k in cand(k) of fl(LB[i][q] + A[q][k])
satisfies: fl(M_i[k] - len) >= tol  =>  no q < k passes at any step. The
kernel then tests q = k alone (weight +0.0, so len is unchanged when it
passes) and, if that fails, resumes the scan after k's own entry; otherwise
it scans cand(k) from the start. This test runs those steps in Python and
requires the oracle's paths on directed and undirected graphs with NaN,
-0.0, -inf and unreachable pairs, and a zero-diagonal-free graph (gate off);
planted mistakes (a bound from H_{L-1} alone, a stay test that forgets
q != j, a resumed scan that skips an entry) must change some path. It pins
the algorithm; tests/test_gpu_apsp.py pins the kernel.
"""
import numpy as np
import pytest

import oracle

F = np.float32


def _graph(N, density, seed, directed=False, special=False):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.05, 2.0, size=(N, N)).astype(F)
    keep = rng.random((N, N)) < density
    band = np.abs(np.arange(N)[:, None] - np.arange(N)[None, :]) <= 2
    keep |= band & (rng.random((N, N)) < 0.9)
    if not directed:
        keep = keep | keep.T
        A = np.minimum(A, A.T)
    A[~keep] = np.inf
    if special:
        A[rng.random((N, N)) < 0.01] = np.nan
        A[rng.random((N, N)) < 0.02] = -0.0
        A[rng.random((N, N)) < 0.004] = -np.inf
    np.fill_diagonal(A, 0.0)
    return A


def _gated_paths(A, H, Hi, L, tol, bug=None):
    N = A.shape[0]
    tol = F(tol)
    cand = [np.nonzero(A[:, k] < np.inf)[0] for k in range(N)]
    gate_ok = bool(np.all(np.diag(A).view(np.uint32) == 0))
    paths = np.zeros(N * L * N, np.int32)
    for i in range(N):
        if gate_ok:
            if L >= 2 and bug != "bound_from_last_only":
                lb = np.fmin(Hi[L - 1][i], Hi[L - 2][i])
            else:
                lb = Hi[L - 1][i]
            M = np.full(N, np.inf, F)
            with np.errstate(invalid="ignore", over="ignore"):
                for k in range(N):
                    for q in cand[k]:
                        if q >= k:
                            break
                        t = F(lb[q] + A[q, k])
                        if np.isnan(t):
                            M[k] = -np.inf
                            break
                        M[k] = min(M[k], t)
        for j in range(N):
            base = (i * N + j) * L
            if i == j:
                paths[base:base + L] = i
                continue
            ln = F(H[i, j])
            k = j
            for p in range(L - 1, -1, -1):
                row = Hi[p][i]
                start = 0
                found = -1
                with np.errstate(invalid="ignore", over="ignore"):
                    if gate_ok and F(M[k] - ln) >= tol:
                        if (k != j or bug == "stay_ignores_j") and F(F(row[k] + F(0.0)) - ln) < tol:
                            found = k
                        else:
                            start = int(np.searchsorted(cand[k], k)) + 1
                            if bug == "resume_skips_one":
                                start += 1
                    if found < 0:
                        for q in cand[k][start:]:
                            if q == j:
                                continue
                            if F(F(row[q] + A[q, k]) - ln) < tol:
                                found = q
                                break
                    if found >= 0:
                        ln = F(ln - A[found, k])
                        k = found
                paths[base + p] = k
    return paths


CASES = [
    # N, L, density, directed, special, seed
    (40, 9, 0.04, False, False, 1),
    (45, 12, 0.03, True, False, 2),
    (36, 7, 0.08, True, True, 3),
    (50, 2, 0.05, False, False, 4),
    (30, 1, 0.1, True, False, 5),
    (48, 14, 0.02, True, False, 6),
]


@pytest.mark.parametrize("N,L,density,directed,special,seed", CASES)
def test_gated_backtrack_equals_oracle(N, L, density, directed, special, seed):
    A = _graph(N, density, seed, directed, special)
    for tol in (1e-4, 0.0):
        H, Hi, P = oracle.hedetniemi(A, L, tol, want_hi=True)
        np.testing.assert_array_equal(_gated_paths(A, H, Hi, L, tol), P)


def test_gate_off_without_zero_diagonal():
    """A positive diagonal entry disables the gate (the bound needs
    A[k][k] = +0.0); the plain scan still equals the oracle."""
    A = _graph(40, 0.05, 8, directed=True)
    A[3, 3] = F(0.25)
    H, Hi, P = oracle.hedetniemi(A, 8, 1e-4, want_hi=True)
    np.testing.assert_array_equal(_gated_paths(A, H, Hi, 8, 1e-4), P)


def test_lower_bound_holds_on_every_snapshot():
    """H_{p+2} <= H_p entry by entry (non-NaN entries) on a zero-diagonal
    graph, so min(H_{L-1}, H_{L-2}) bounds every snapshot from below."""
    A = _graph(120, 0.02, 9, directed=True, special=True)
    L = 16
    _, Hi, _ = oracle.hedetniemi(A, L, 1e-4, want_hi=True, want_paths=False)
    lb = np.fmin(Hi[L - 1], Hi[L - 2])
    for p in range(L):
        ok = np.isnan(Hi[p]) | (lb <= Hi[p])
        assert ok.all(), p


@pytest.mark.parametrize("bug", ["bound_from_last_only", "stay_ignores_j", "resume_skips_one"])
def test_planted_mistakes_change_paths(bug):
    differs = False
    for N, L, density, directed, special, seed in CASES:
        A = _graph(N, density, seed, directed, special)
        H, Hi, P = oracle.hedetniemi(A, L, 1e-4, want_hi=True)
        if not np.array_equal(_gated_paths(A, H, Hi, L, 1e-4, bug=bug), P):
            differs = True
            break
    assert differs, bug
