"""Seeded synthetic inputs for the five benchmark / parity workloads.

Every array is float32 in the reference's node-row layout: row a of X holds
[x frames | y frames | z frames] of node a (netchem/src/network.cpp:65-69),
so dimension di of frame f is X[a, di * n + f].

* ``gaussian_nodes`` — M scalar nodes (d=1) sampled from a rank-1 factor
  model x_i = s_i z + sqrt(1 - s_i^2) e_i with s_i ~ U[0, 0.95], so the
  pairwise correlations rho_ij = s_i s_j span [0, 0.9]; the analytic MI
  -0.5 ln(1 - rho^2) is the sanity reference (64 nodes, n=1000, k=4).
* ``calpha_trajectory`` — N C-alpha-like nodes (d=3): the mean structure is
  a seeded random-walk chain with 3.8 A steps; fluctuations are R=20
  collective modes plus isotropic N(0, 0.5^2) A noise. Mode r moves a
  contiguous chain segment (Gaussian envelope centred at c_r, width w_r
  ~ U[0.5, 1.5] * N / 40 nodes) rigidly along a random unit direction v_r with
  amplitude a_r(t) ~ N(0, sigma_r^2), sigma_r ~ U[0.4, 1.2] A. Frames are
  i.i.d. Nodes that share a mode's core are strongly coupled; most pairs are
  not, as in a protein with a few correlated domains.
* ``bootstrap_windows`` — R contiguous frame windows [s_r, s_r + n_r) with
  n_r ~ U[lo, hi] and s_r ~ U[0, n_total - n_r].
* ``contact_adjacency`` — the Hedetniemi input from a GC matrix:
  A_ij = -log(GC_ij) where the mean C-alpha distance is < 7.5 A, +inf
  elsewhere, 0 on the diagonal; plus the hop diameter of that graph.
"""
import numpy as np


def gaussian_nodes(M=64, n=1000, seed=0):
    """Returns (X[M, n] float32, rho[M, M] float64)."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.0, 0.95, size=M)
    z = rng.standard_normal(n)
    e = rng.standard_normal((M, n))
    X = s[:, None] * z[None, :] + np.sqrt(1.0 - s ** 2)[:, None] * e
    rho = np.outer(s, s)
    np.fill_diagonal(rho, 1.0)
    return X.astype(np.float32), rho


def _chain(N, rng, step=3.8):
    d = rng.standard_normal((N, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.cumsum(step * d, axis=0)


def calpha_trajectory(N=1000, n=2000, seed=1, modes=20, noise=0.5):
    """Returns (X[N, 3n] float32, mean[N, 3] float64).

    The mean is the seeded chain; X is the sampled trajectory in node-row
    layout. Memory is O(N * n) in float32 plus one modes x n block.
    """
    rng = np.random.default_rng(seed)
    mean = _chain(N, rng)
    idx = np.arange(N)
    centres = rng.uniform(0, N, size=modes)
    widths = rng.uniform(0.5, 1.5, size=modes) * max(N / 40.0, 2.0)
    dirs = rng.standard_normal((modes, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    sig = rng.uniform(0.4, 1.2, size=modes)
    env = np.exp(-0.5 * ((idx[None, :] - centres[:, None]) / widths[:, None]) ** 2)  # R x N
    amp = rng.standard_normal((modes, n)) * sig[:, None]                          # R x n
    X = np.empty((N, 3 * n), np.float32)
    for di in range(3):
        # displacement[node, frame] = sum_r env[r, node] * dirs[r, di] * amp[r, frame]
        disp = (env * dirs[:, di:di + 1]).T @ amp
        X[:, di * n:(di + 1) * n] = (mean[:, di:di + 1] + disp +
                                    noise * rng.standard_normal((N, n))).astype(np.float32)
    return X, mean


def all_pairs(N):
    """Unordered pairs (a < b) as an int32 [P, 2] array."""
    a, b = np.triu_indices(N, k=1)
    return np.stack([a, b], axis=1).astype(np.int32)


def bootstrap_windows(R=64, n_total=4000, lo=500, hi=4000, seed=3):
    """Returns int32 [R, 2] rows of (offset, length)."""
    rng = np.random.default_rng(seed)
    ln = rng.integers(lo, hi + 1, size=R)
    off = np.array([rng.integers(0, n_total - l + 1) for l in ln])
    return np.stack([off, ln], axis=1).astype(np.int32)


def mean_distances(mean):
    d = mean[:, None, :] - mean[None, :, :]
    return np.sqrt((d ** 2).sum(-1))


def contact_pairs(mean, cutoff=7.5):
    """Unordered pairs (a < b) whose mean distance is below cutoff."""
    D = mean_distances(mean)
    a, b = np.nonzero(np.triu(D < cutoff, k=1))
    return np.stack([a, b], axis=1).astype(np.int32)


def contact_adjacency(gc_pairs, pairs, N):
    """A[N, N] float32 from GC values of the contact pairs (symmetric)."""
    A = np.full((N, N), np.inf, np.float32)
    with np.errstate(divide="ignore"):
        w = (-np.log(gc_pairs.astype(np.float64))).astype(np.float32)
    A[pairs[:, 0], pairs[:, 1]] = w
    A[pairs[:, 1], pairs[:, 0]] = w
    np.fill_diagonal(A, 0.0)
    return A


def hop_diameter(A):
    """Largest finite hop count between any two nodes (edges: finite A_ij,
    i != j), by breadth-first search from every node."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path
    adj = np.isfinite(A)
    np.fill_diagonal(adj, False)
    hops = shortest_path(csr_matrix(adj.astype(np.int8)), method="D", directed=True,
                         unweighted=True)
    return int(hops[np.isfinite(hops)].max())
