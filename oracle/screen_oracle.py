"""Screener oracle — TEST INFRASTRUCTURE ONLY.

numpy restatements of the pair-feature map and of the 16 -> 64 -> 64 -> 1 MLP
(forward, backward, Adam) that netsci_amd/csrc/screener.hip runs on the GPU,
with the same bf16 rounding points, so the HIP kernels' features,
predictions and gradients can be checked. The screener has no counterpart in
the reference; these functions define it.

Pinning: `mlp_grad(..., bf16=False)` is checked against central finite
differences in float64 (tests/test_screen_oracle.py), and the features
against direct numpy statistics of the trajectory (Pearson r, Gaussian MI
from the 6x6 joint covariance via log-determinants, canonical correlations
via SVD).
"""
import numpy as np

F, H = 16, 64
O_W1, O_B1, O_W2, O_B2, O_W3, O_B3 = 0, 1024, 1088, 5184, 5248, 5312
N_PARAMS = 5313
SKIP = 5  # feature added to the output (the Gaussian-MI GC)


def bf16_round(x):
    """float32 -> bf16 (round to nearest even) -> float32, as the kernels do."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(x.shape)


def unpack(params):
    p = np.asarray(params, np.float64)
    return (p[O_W1:O_B1].reshape(F, H), p[O_B1:O_W2], p[O_W2:O_B2].reshape(H, H),
            p[O_B2:O_W3], p[O_W3:O_B3], p[O_B3])


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    p[O_W1:O_B1] = rng.standard_normal(F * H) * np.sqrt(2.0 / F)
    p[O_W2:O_B2] = rng.standard_normal(H * H) * np.sqrt(2.0 / H)
    p[O_W3:O_B3] = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return p


def _q(x, bf16):
    """bf16 operand (via float32, as the kernels round) or exact float64."""
    if bf16:
        return bf16_round(np.asarray(x, np.float32)).astype(np.float64)
    return np.asarray(x, np.float64)


def mlp_forward(params, x, bf16=True, keep=False):
    """out = relu(relu(x W1 + b1) W2 + b2) . w3 + b3 + x[5] with bf16
    operands for both matrix products (the kernels' MFMA inputs) and float
    accumulation."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    xq = _q(x, bf16)
    z1 = xq @ _q(W1, bf16) + b1
    h1 = np.maximum(z1, 0)
    z2 = _q(h1, bf16) @ _q(W2, bf16) + b2
    h2 = np.maximum(z2, 0)
    out = h2 @ w3 + b3 + xq[:, SKIP]
    if keep:
        return out, (xq, z1, h1, z2, h2)
    return out


def mlp_loss(params, x, y, w, bf16=True):
    out = mlp_forward(params, x, bf16)
    return float(np.sum(w * (out - y) ** 2) / len(y))


def mlp_grad(params, x, y, w, bf16=True):
    """Gradient of mean_i w_i (out_i - y_i)^2 in the parameter layout, with
    the kernels' bf16 rounding of dz2 and dz1 before their products."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    out, (xq, z1, h1, z2, h2) = mlp_forward(params, x, bf16, keep=True)
    g = 2.0 * w * (out - y) / len(y)
    dw3 = h2.T @ g
    db3 = g.sum()
    dz2 = np.outer(g, w3) * (z2 > 0)
    dW2 = _q(h1, bf16).T @ _q(dz2, bf16)
    db2 = dz2.sum(0)
    dh1 = _q(dz2, bf16) @ _q(W2, bf16).T
    dz1 = dh1 * (z1 > 0)
    dW1 = xq.T @ _q(dz1, bf16)
    db1 = dz1.sum(0)
    grad = np.zeros(N_PARAMS)
    grad[O_W1:O_B1] = dW1.ravel()
    grad[O_B1:O_W2] = db1
    grad[O_W2:O_B2] = dW2.ravel()
    grad[O_B2:O_W3] = db2
    grad[O_W3:O_B3] = dw3
    grad[O_B3] = db3
    return grad


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def node_rows(X, N, n, dstride=None):
    ds = dstride or n
    return np.stack([np.stack([X[a, c * ds: c * ds + n] for c in range(3)]) for a in range(N)])


def pair_features(X, ab, N, n, dstride=None):
    """The 16 screening features of each pair (float64), from the trajectory
    directly: see netsci_amd/csrc/screener.hip for the list."""
    T = node_rows(np.asarray(X, np.float64), N, n, dstride)      # N x 3 x n
    mu = T.mean(axis=2)
    D = T - mu[:, :, None]
    out = np.zeros((len(ab), F))
    for p, (a, b) in enumerate(ab):
        Caa = D[a] @ D[a].T / n
        Cbb = D[b] @ D[b].T / n
        Cab = D[a] @ D[b].T / n
        ta, tb = np.trace(Caa), np.trace(Cbb)
        r = np.trace(Cab) / np.sqrt(ta * tb)
        La, Lb = np.linalg.cholesky(Caa), np.linalg.cholesky(Cbb)
        Tm = np.linalg.inv(La) @ Cab @ np.linalg.inv(Lb).T
        rho = np.sort(np.clip(np.linalg.svd(Tm, compute_uv=False), 0, 1 - 1e-6))[::-1]
        ig = min(-0.5 * np.sum(np.log(1 - rho ** 2)), 10.0)
        dist = np.linalg.norm(mu[a] - mu[b])
        ra, rb = np.sqrt(ta), np.sqrt(tb)
        corr = Cab / np.sqrt(np.outer(np.diag(Caa), np.diag(Cbb)))
        out[p] = [abs(r), rho[0], rho[1], rho[2], ig, np.sqrt(1 - np.exp(-2 * ig / 3)),
                  dist / 10.0, min(7.5 / max(dist, 1.0), 2.0), ra, rb,
                  min(ra, rb) / max(ra, rb), np.log1p(abs(a - b)) / np.log1p(N),
                  np.linalg.norm(Tm) / np.sqrt(3.0), np.abs(corr).max(), np.log(n) / 10.0, 1.0]
    return out
