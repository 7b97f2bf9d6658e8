"""Parity of the gfx950 Hedetniemi path with the oracle (GPU required).

Distances, snapshots and backtracked paths must be bit-exact against the
oracle's Jacobi restatement of hedetniemi_gpu.cu:47-107.
"""
import ctypes

import numpy as np
import pytest

import oracle
from netsci_amd import abi, cuarray, netcalc

pytestmark = pytest.mark.gpu


def _graph(N, density, seed, negzero=False):
    rng = np.random.default_rng(seed)
    A = rng.uniform(0.1, 3.0, size=(N, N)).astype(np.float32)
    A[rng.random((N, N)) > density] = np.inf
    np.fill_diagonal(A, 0.0)
    if negzero:
        A[rng.random((N, N)) < 0.02] = -0.0
    return A


def _gpu(A, L, tol):
    N = A.shape[0]
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    dHi = abi.DeviceArray((L, N, N), np.float32)
    dP = abi.DeviceArray((N * L * N,), np.int32)
    abi.hedetniemi(dA, N, L, tol, dH, dHi, dP)
    return dH.numpy(), dHi.numpy(), dP.numpy()


# Sparse graphs take the lane-per-(i, j) candidate-list backtrack; the last
# case (mean in-degree 270 > 128) takes the wave-per-(i, j) dense kernel.
@pytest.mark.parametrize("N,L,density,negzero", [(37, 6, 0.15, False), (100, 12, 0.05, False),
                                                 (130, 5, 0.3, True), (64, 1, 0.5, False),
                                                 (300, 4, 0.9, False)])
def test_hedetniemi_bitexact(N, L, density, negzero):
    print(f"apsp N={N} L={L}", flush=True)
    A = _graph(N, density, seed=N + L, negzero=negzero)
    H, Hi, P = _gpu(A, L, 1e-4)
    Ho, Hio, Po = oracle.hedetniemi(A, L, 1e-4, want_hi=True)
    np.testing.assert_array_equal(H.view(np.uint32), Ho.view(np.uint32))
    np.testing.assert_array_equal(Hi.view(np.uint32), Hio.view(np.uint32))
    np.testing.assert_array_equal(P, Po)


def test_sparse_special_values_bitexact():
    """NaN and -0.0 weights on a sparse graph: the neighbour lists drop
    exactly the entries the reference's tests can never take (+inf, NaN), and
    the sparse step keeps the strict-`<` first minimum between -0.0 and +0.0,
    so distances and paths still match the oracle bit for bit."""
    N, L = 90, 6
    print(f"apsp special values N={N}", flush=True)
    rng = np.random.default_rng(21)
    A = _graph(N, 0.08, seed=21)
    A[rng.random((N, N)) < 0.01] = np.nan
    A[rng.random((N, N)) < 0.01] = -0.0
    np.fill_diagonal(A, 0.0)
    H, Hi, P = _gpu(A, L, 1e-4)
    Ho, Hio, Po = oracle.hedetniemi(A, L, 1e-4, want_hi=True)
    np.testing.assert_array_equal(H.view(np.uint32), Ho.view(np.uint32))
    np.testing.assert_array_equal(Hi.view(np.uint32), Hio.view(np.uint32))
    np.testing.assert_array_equal(P, Po)


def test_contact_graph_paths_bitexact():
    """The C-alpha contact-graph shape the path workload uses (-log GC on
    pairs closer than 7.5 A, about ten candidates per node), L = hop diameter
    + 1, against the oracle."""
    from netsci_amd import synth
    N, n = 400, 600
    print(f"apsp contact graph N={N}", flush=True)
    X, mean = synth.calpha_trajectory(N=N, n=n, seed=11)
    pairs = synth.contact_pairs(mean)
    # Edge weights from the GPU KSG path (bit-exact against the oracle in
    # test_gpu_ksg.py); only the graph's shape matters here.
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(pairs)
    dmi = abi.DeviceArray((len(pairs),), np.float32)
    abi.ksg_mi(dX, X.shape[1], n, 3, n, dab, len(pairs), 6, dmi)
    gc = abi.gc_from_mi(dmi.numpy(), 3)
    A = synth.contact_adjacency(gc, pairs, N)
    L = synth.hop_diameter(A) + 1
    H, Hi, P = _gpu(A, L, 1e-4)
    Ho, Hio, Po = oracle.hedetniemi(A, L, 1e-4, want_hi=True)
    np.testing.assert_array_equal(H.view(np.uint32), Ho.view(np.uint32))
    np.testing.assert_array_equal(P, Po)


@pytest.mark.parametrize("L", [7, 8])
def test_paths_unaligned_buffer_bitexact(L):
    """A paths pointer that is not 16-byte aligned (4 bytes into a buffer)
    takes the one-store-per-step backtrack; the 64-byte-block stores of the
    aligned case must give the same entries, and the word before the pointer
    stays untouched."""
    N = 120
    print(f"apsp unaligned paths L={L}", flush=True)
    A = _graph(N, 0.06, seed=40 + L)
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    dHi = abi.DeviceArray((L, N, N), np.float32)
    buf = abi.DeviceArray.from_numpy(np.full(N * L * N + 1, -7, np.int32))

    class _View:
        ptr = ctypes.c_void_p(buf.ptr.value + 4)

    abi.hedetniemi(dA, N, L, 1e-4, dH, dHi, _View)
    got = buf.numpy()
    _, _, Po = oracle.hedetniemi(A, L, 1e-4, want_hi=True)
    assert got[0] == -7
    np.testing.assert_array_equal(got[1:], Po)
    _, _, P = _gpu(A, L, 1e-4)
    np.testing.assert_array_equal(P, Po)


@pytest.mark.parametrize("L", [3, 4])
def test_sparse_steps_long_lists(L):
    """A sparse graph near the dense threshold (N = 1900, about 115 entries
    per neighbour list, 16-wide j tiles): lengths (even and odd L, both
    ping-pong parities) and snapshots must equal the oracle."""
    N = 1900
    print(f"apsp sparse steps, long lists N={N} L={L}", flush=True)
    A = _graph(N, 0.06, seed=19 + L)
    assert ((A < np.inf).sum(axis=1) + 3) // 4 * 4 @ np.ones(N) / N < 128  # sparse path
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    abi.hedetniemi_lengths(dA, N, L, dH)
    Ho, Hio, _ = oracle.hedetniemi(A, L, 1e-4, want_hi=True, want_paths=False)
    np.testing.assert_array_equal(dH.numpy().view(np.uint32), Ho.view(np.uint32))
    dHi = abi.DeviceArray((L, N, N), np.float32)
    abi.hedetniemi(dA, N, L, 1e-4, dH, dHi, None)
    np.testing.assert_array_equal(dH.numpy().view(np.uint32), Ho.view(np.uint32))
    np.testing.assert_array_equal(dHi.numpy().view(np.uint32), Hio.view(np.uint32))


def test_sparse_steps_mixed_tiles():
    """A banded graph whose first 64 rows are dense-ish (neighbour lists of
    about 200 entries beside lists of 11): waves of one step walk lists of
    very different lengths; lengths and snapshots must equal the oracle."""
    N, L = 700, 5
    print(f"apsp mixed tiles N={N} L={L}", flush=True)
    rng = np.random.default_rng(77)
    A = rng.uniform(0.1, 2.0, (N, N)).astype(np.float32)
    band = np.abs(np.arange(N)[:, None] - np.arange(N)[None, :]) <= 5
    keep = band.copy()
    keep[:64] |= rng.random((64, N)) < 0.3
    A[~keep] = np.inf
    np.fill_diagonal(A, 0.0)
    assert ((A < np.inf).sum(axis=1) + 3) // 4 * 4 @ np.ones(N) / N < 128  # sparse path
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    dHi = abi.DeviceArray((L, N, N), np.float32)
    Ho, Hio, _ = oracle.hedetniemi(A, L, 1e-4, want_hi=True, want_paths=False)
    abi.hedetniemi_lengths(dA, N, L, dH)
    np.testing.assert_array_equal(dH.numpy().view(np.uint32), Ho.view(np.uint32))
    abi.hedetniemi(dA, N, L, 1e-4, dH, dHi, None)
    np.testing.assert_array_equal(dH.numpy().view(np.uint32), Ho.view(np.uint32))
    np.testing.assert_array_equal(dHi.numpy().view(np.uint32), Hio.view(np.uint32))


@pytest.mark.parametrize("N,L,diag", [(160, 11, "zero"), (160, 11, "positive"),
                                      (160, 11, "negzero"), (96, 1, "zero"), (96, 2, "zero"),
                                      (700, 9, "zero"), (300, 45, "zero"), (258, 37, "zero")])
def test_backtrack_stay_gate(N, L, diag):
    """Directed graphs (H_p far from symmetric, so the gate's bound needs
    both H_{L-1} and H_{L-2}) with a band of short edges, whose paths mostly
    stay at a node for many steps. With an all-+0.0 diagonal the backtrack
    skips the scans the gate rules out; a positive or -0.0 diagonal entry
    turns the gate off. L = 37 and 45 fill several 16-entry path blocks per
    (i, j) (N = 258: rows that are not 16-byte multiples stage through
    registers). Paths must equal the oracle in every case."""
    print(f"apsp stay gate N={N} L={L} diag={diag}", flush=True)
    rng = np.random.default_rng(N + L)
    A = rng.uniform(0.05, 2.0, size=(N, N)).astype(np.float32)
    band = np.abs(np.arange(N)[:, None] - np.arange(N)[None, :]) <= 3
    keep = (band & (rng.random((N, N)) < 0.9)) | (rng.random((N, N)) < 0.01)
    A[~keep] = np.inf
    np.fill_diagonal(A, 0.0)
    if diag == "positive":
        A[5, 5] = 0.25
    elif diag == "negzero":
        A[7, 7] = -0.0
    H, Hi, P = _gpu(A, L, 1e-4)
    Ho, Hio, Po = oracle.hedetniemi(A, L, 1e-4, want_hi=True)
    np.testing.assert_array_equal(H.view(np.uint32), Ho.view(np.uint32))
    np.testing.assert_array_equal(P, Po)


def test_lengths_only_matches():
    N, L = 150, 9
    print("apsp lengths", flush=True)
    A = _graph(N, 0.04, seed=3)
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    abi.hedetniemi_lengths(dA, N, L, dH)
    Ho, _, _ = oracle.hedetniemi(A, L, 0.0, want_paths=False)
    np.testing.assert_array_equal(dH.numpy().view(np.uint32), Ho.view(np.uint32))


@pytest.mark.parametrize("N,L", [(5200, 5), (8200, 3)])
def test_large_n_backtrack_variants_sampled(N, L):
    """N = 5200 takes the backtrack whose list starts stay in global memory
    (its row buffers and bound table fill the workgroup's LDS) and N = 8200
    the one that reads snapshot rows from L2 instead of LDS. The dense oracle
    is too slow here, so 400 seeded (i, j) are checked bit for bit against
    the recurrence itself: every sampled H_{p+1}[i][j] refolded from the
    GPU's own H_p (ascending k, strict `<`), and every sampled paths row
    re-derived by the first-hit rule (hedetniemi_gpu.cu:86-104) from the
    GPU's own snapshots."""
    print(f"apsp large-N backtrack N={N} L={L}", flush=True)
    rng = np.random.default_rng(N)
    band = 3
    A = np.full((N, N), np.inf, np.float32)
    for o in range(-band, band + 1):
        if o:
            r = np.arange(max(0, -o), N - max(0, o))  # rows whose column r + o exists
            A[r, r + o] = rng.uniform(0.05, 2.0, len(r)).astype(np.float32)
    far = rng.integers(0, N, size=(4 * N, 2))
    A[far[:, 0], far[:, 1]] = rng.uniform(0.5, 3.0, len(far)).astype(np.float32)
    np.fill_diagonal(A, 0.0)
    tol = np.float32(1e-4)
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    dHi = abi.DeviceArray((L, N, N), np.float32)
    dP = abi.DeviceArray((N * L * N,), np.int32)
    abi.hedetniemi(dA, N, L, float(tol), dH, dHi, dP)
    H, Hi = dH.numpy(), dHi.numpy()
    np.testing.assert_array_equal(Hi[0].view(np.uint32), A.view(np.uint32))
    cand_row = [np.nonzero(A[j] < np.inf)[0] for j in range(N)]     # step candidates of j
    cand_col = [np.nonzero(A[:, k] < np.inf)[0] for k in range(N)]  # backtrack candidates of k
    ij = rng.integers(0, N, size=(400, 2))
    row = np.empty(L, np.int32)
    with np.errstate(invalid="ignore", over="ignore"):
        for i, j in ij:
            i, j = int(i), int(j)
            # Steps: H_{p+1}[i][j] = first minimum over k of fl(A[j][k] + H_p[k][i]).
            for p in range(L):
                nxt = H if p + 1 == L else Hi[p + 1]
                c = np.float32(np.inf)
                for k in cand_row[j]:
                    t = np.float32(A[j, k] + Hi[p][k, i])
                    if t < c:
                        c = t
                assert nxt[i, j].view(np.uint32) == c.view(np.uint32), (i, j, p)
            # Backtrack.
            off = ((i * L * N) + j * L) * 4
            abi.check(abi.lib().nsx_memcpy_d2h(row.ctypes.data, ctypes.c_void_p(dP.ptr.value + off),
                                                L * 4), "paths row")
            if i == j:
                assert (row == i).all()
                continue
            ln = np.float32(H[i, j])
            k = j
            for p in range(L - 1, -1, -1):
                hrow = Hi[p][i]
                for q in cand_col[k]:
                    if q != j and np.float32(np.float32(hrow[q] + A[q, k]) - ln) < tol:
                        ln = np.float32(ln - A[q, k])
                        k = int(q)
                        break
                assert row[p] == k, (i, j, p, row.tolist())


@pytest.mark.parametrize("L,in_place", [(0, False), (1, False), (4, False), (3, True),
                                         (4, True)])
def test_lengths_first_step_reads_A(L, in_place):
    """The lengths entry point's first step reads A itself (no H_0 copy);
    when H is A's own buffer, A is copied into the buffer that step 0 does
    not write. Both parities of L, L = 0 (H = A) and L = 1, against the
    oracle, with -0.0 weights so the exact fold runs."""
    N = 140
    print(f"apsp lengths L={L} in_place={in_place}", flush=True)
    A = _graph(N, 0.05, seed=50 + L, negzero=True)
    dA = abi.DeviceArray.from_numpy(A)
    dH = dA if in_place else abi.DeviceArray((N, N), np.float32)
    abi.hedetniemi_lengths(dA, N, L, dH)
    Ho, _, _ = oracle.hedetniemi(A, L, 0.0, want_paths=False) if L else (A, None, None)
    np.testing.assert_array_equal(dH.numpy().view(np.uint32), Ho.view(np.uint32))


def test_netcalc_gpu_equals_cpu_platform_and_paths_decode():
    N, L = 60, 8
    print("apsp netcalc", flush=True)
    A = _graph(N, 0.08, seed=7)
    Ac = cuarray.FloatCuArray()
    Ac.fromNumpy2D(A)
    out = {}
    for plat in (netcalc.GPU_PLATFORM, netcalc.CPU_PLATFORM):
        H, P = cuarray.FloatCuArray(), cuarray.IntCuArray()
        netcalc.hedetniemiAllShortestPaths(A=Ac, H=H, paths=P, maxPathLength=L, tolerance=1e-4,
                                           platform=plat)
        out[plat] = (H.toNumpy2D(), P.toNumpy1D(), P)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    for i, j in [(0, 5), (3, 41), (10, 10)]:
        got = netcalc.recoverSingleShortestPath(paths=out[0][2], maxPathLength=L, i=i, j=j)
        np.testing.assert_array_equal(got, oracle.recover_path(out[0][1], N, L, i, j))


def test_corr_to_adj_gpu():
    N = 70
    print("adjacency", flush=True)
    rng = np.random.default_rng(1)
    C = rng.uniform(0, 1, (N, N)).astype(np.float32)
    C[rng.random((N, N)) < 0.2] = 0.0
    Cc, A = cuarray.FloatCuArray(), cuarray.FloatCuArray()
    Cc.fromNumpy2D(C)
    netcalc.correlationToAdjacency(A=A, C=Cc, n=N, platform=netcalc.GPU_PLATFORM)
    np.testing.assert_array_equal(A.toNumpy2D(), oracle.corr_to_adj(C))
