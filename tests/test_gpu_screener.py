"""The learned screener's GPU kernels against the screener oracle, and the
screened all-pairs pipeline against the plain path (GPU required).

Tolerances: features are stored as bf16 (relative spacing 2^-8), so GPU and
oracle features agree to 1 bf16 step plus the FP32-vs-float64 difference of
the covariance (|diff| <= 0.01 + 2^-7 |value|). MLP forward and gradients use
the same bf16 operand rounding as the oracle; only the FP32 accumulation
order differs (relative 1e-4 on outputs, 2e-3 on gradients, scaled by the
gradient's magnitude).
"""
import numpy as np
import pytest

from oracle import screen_oracle as so
from netsci_amd import abi, screen, synth

pytestmark = pytest.mark.gpu


def _features(N=12, n=700, seed=31):
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=seed)
    ab = synth.all_pairs(N)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    f = screen.features(dX, X.shape[1], n, N, n, dab, len(ab)).numpy()
    return X, ab, f


def test_features_match_oracle():
    print("features", flush=True)
    X, ab, f = _features()
    got = screen.bf16_to_float(f).reshape(len(ab), 16)
    want = so.pair_features(X, ab, 12, 700)
    np.testing.assert_allclose(got, want, atol=0.01, rtol=2 ** -7)


def test_mlp_forward_matches_oracle():
    print("mlp forward", flush=True)
    rng = np.random.default_rng(1)
    P = 1000  # not a multiple of 16: the tail group is partial
    x = so.bf16_round(rng.standard_normal((P, 16)).astype(np.float32))
    feat = (x.view(np.uint32) >> 16).astype(np.uint16)
    p = so.init_params(2)
    p[so.O_B1:so.O_W2] = rng.standard_normal(64) * 0.1
    p[so.O_B2:so.O_W3] = rng.standard_normal(64) * 0.1
    p[so.O_B3] = 0.05
    s = screen.Screener(p)
    got = s.predict(feat)
    want = so.mlp_forward(p, x)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4 * np.abs(want).max())


def test_mlp_gradient_matches_oracle():
    print("mlp gradient", flush=True)
    rng = np.random.default_rng(3)
    B = 333
    x = so.bf16_round(rng.standard_normal((B, 16)).astype(np.float32))
    feat = (x.view(np.uint32) >> 16).astype(np.uint16)
    y = rng.random(B).astype(np.float32)
    w = (1 + rng.random(B)).astype(np.float32)
    p = so.init_params(4)
    p[so.O_B1:so.O_W2] = 0.05
    p[so.O_B2:so.O_W3] = 0.05
    dp, df = abi.DeviceArray.from_numpy(p), abi.DeviceArray.from_numpy(feat)
    dy, dw = abi.DeviceArray.from_numpy(y), abi.DeviceArray.from_numpy(w)
    dg, dl = abi.DeviceArray(p.shape, np.float32), abi.DeviceArray((1,), np.float32)
    abi.mlp_grad(dp, df, dy, dw, B, dg, dl)
    got = dg.numpy()
    want = so.mlp_grad(p, x, y, w)
    for lo, hi in [(so.O_W1, so.O_B1), (so.O_B1, so.O_W2), (so.O_W2, so.O_B2),
                   (so.O_B2, so.O_W3), (so.O_W3, so.O_B3), (so.O_B3, so.N_PARAMS)]:
        scale = np.abs(want[lo:hi]).max()
        np.testing.assert_allclose(got[lo:hi], want[lo:hi], atol=2e-3 * scale, rtol=2e-3)
    assert abs(float(dl.numpy()[0]) - so.mlp_loss(p, x, y, w)) < 1e-4 * so.mlp_loss(p, x, y, w)
    # one Adam step
    dm = abi.DeviceArray.from_numpy(np.zeros_like(p))
    dv = abi.DeviceArray.from_numpy(np.zeros_like(p))
    abi.mlp_adam(dp, dg, dm, dv, 1, 1e-3)
    q, _, _ = so.adam(p.astype(np.float64), got.astype(np.float64), 0, 0, 1, 1e-3)
    np.testing.assert_allclose(dp.numpy(), q, atol=2e-6)


def test_training_reduces_loss():
    print("training", flush=True)
    rng = np.random.default_rng(5)
    P = 8192
    x = rng.standard_normal((P, 16)).astype(np.float32)
    x[:, 15] = 1.0
    y = (np.tanh(x[:, 0] - 0.5 * x[:, 1]) * 0.5 + 0.5).astype(np.float32)
    feat = (so.bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)
    s = screen.Screener(seed=0)
    hist = s.fit(feat, y, epochs=12, batch=1024, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0], hist


def test_screened_all_pairs_small():
    """Train on three small trajectories, screen a fourth: recall of the
    GC > 0.3 pairs >= 99.5% against the plain path, computed values
    bit-identical to the plain path, and (measured on the same test pairs)
    what the trained network gains over the single-feature rules: it keeps
    fewer pairs than a |Pearson r| threshold and no more than a Gaussian-MI
    threshold at equal recall."""
    N, n, k = 300, 1500, 6
    print("screened: training data", flush=True)
    feat, y = screen.training_set(trajectories=3, N=N, n=n, pairs_per=44850, k=k, seed=300)
    s, info = screen.train_screener(feat, y, recall=0.999, epochs=30, seed=0)
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=399)
    ab = synth.all_pairs(N)
    res = screen.screened_gc(X, N, n, k, s, audit_rate=0.02, seed=1)
    dX = abi.DeviceArray.from_numpy(X)
    plain = screen.exact_gc(dX, X.shape[1], n, 3, n, ab, k)
    comp = res["computed"]
    np.testing.assert_array_equal(res["gc"][comp], plain[comp])
    pos = plain > screen.GC_THRESHOLD
    recall = float(np.sum(pos & comp) / max(pos.sum(), 1))
    stats = {k2: v for k2, v in res.items() if k2 not in ("gc", "computed")}
    print(f"recall {recall:.4f}, computed fraction {comp.mean():.4f}, stats {stats}", flush=True)
    assert recall >= 0.995
    assert comp.mean() < 0.25
    dab = abi.DeviceArray.from_numpy(ab)
    ft = screen.features(dX, X.shape[1], n, N, n, dab, len(ab)).numpy()
    fl = screen.bf16_to_float(ft).reshape(len(ab), 16)
    frac = {name: screen.min_fraction_at_recall(sc, plain, 0.995)
            for name, sc in [("mlp", s.predict(ft)), ("gaussian_mi_gc", fl[:, 5]),
                             ("abs_pearson_r", fl[:, 0])]}
    print(f"pairs kept at 99.5% recall: {frac}; positives {pos.mean():.4f}", flush=True)
    assert frac["mlp"] < frac["abs_pearson_r"]
    assert frac["mlp"] <= 1.02 * frac["gaussian_mi_gc"]
