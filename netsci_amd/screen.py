"""Screened all-pairs generalized correlation with a learned pair screener.

No reference counterpart: the reference computes KSG for every requested pair
(netcalc/src/generalized_correlation.cpp:104-173). Here a small MLP, trained
on exact GC values that the plain path produces, predicts from 16 cheap
per-pair features (covariance statistics; see screener.hip) which pairs may
exceed a GC threshold. Exact KSG then runs only on those candidates plus a
random audit sample of the rest:

* every GC value reported for a computed pair is exact KSG (bit-identical to
  the plain path);
* the audit sample estimates how many above-threshold pairs were screened
  out; if the estimated recall falls below `min_recall`, the plain path runs
  on every remaining pair (the fallback), so the result is never less than
  the plain method's;
* training runs on the GPU with the project's own bf16-MFMA forward and
  backward kernels and Adam.
"""
import time

import numpy as np

from . import abi, synth

GC_THRESHOLD = 0.3
N_FEATURES = 16


def exact_gc(dX, ld, dstride, d, n, ab_np, k, out=None):
    """Exact GC (KSG + host double transform) of the given pairs."""
    ab_np = np.ascontiguousarray(ab_np, np.int32).reshape(-1, 2)
    P = len(ab_np)
    if P == 0:
        return np.zeros(0, np.float32)
    dab = abi.DeviceArray.from_numpy(ab_np)
    dmi = abi.DeviceArray((P,), np.float32)
    abi.ksg_mi(dX, ld, dstride, d, n, dab, P, k, dmi)
    return abi.gc_from_mi(dmi.numpy(), d)


def features(dX, ld, dstride, N, n, dab, P):
    """Device bf16 features (P x 16, uint16 bits) of the pairs in dab."""
    dfeat = abi.DeviceArray((P, N_FEATURES), np.uint16)
    abi.screen_features(dX, ld, dstride, N, n, dab, P, dfeat)
    return dfeat


def bf16_to_float(u16):
    return (np.asarray(u16, np.uint16).astype(np.uint32) << 16).view(np.float32)


def sample_weights(y):
    """Heavier weight on the region that decides recall at the threshold."""
    return (1.0 + 4.0 * (y > GC_THRESHOLD - 0.1)).astype(np.float32)


class Screener:
    """The 16 -> 64 -> 64 -> 1 MLP: FP32 master parameters, GPU training."""

    def __init__(self, params=None, seed=0):
        if params is None:
            params = init_params(seed)
        self.params = np.ascontiguousarray(params, np.float32)
        self.threshold = None

    # -- training ---------------------------------------------------------
    def fit(self, feat, y, w=None, epochs=40, batch=4096, lr=2e-3, seed=0, log=None):
        """Mini-batch Adam on the GPU. feat: P x 16 uint16 (bf16), y: GC."""
        feat = np.ascontiguousarray(feat, np.uint16)
        y = np.ascontiguousarray(y, np.float32)
        w = sample_weights(y) if w is None else np.ascontiguousarray(w, np.float32)
        P = len(y)
        rng = np.random.default_rng(seed)
        dp = abi.DeviceArray.from_numpy(self.params)
        dm = abi.DeviceArray.from_numpy(np.zeros_like(self.params))
        dv = abi.DeviceArray.from_numpy(np.zeros_like(self.params))
        dg = abi.DeviceArray(self.params.shape, np.float32)
        dl = abi.DeviceArray((1,), np.float32)
        step = 0
        history = []
        for ep in range(epochs):
            perm = rng.permutation(P)
            df = abi.DeviceArray.from_numpy(feat[perm])
            dy = abi.DeviceArray.from_numpy(y[perm])
            dw = abi.DeviceArray.from_numpy(w[perm])
            lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))  # cosine decay
            losses = []
            for b0 in range(0, P - batch + 1, batch):
                step += 1
                abi.check(abi.lib().nsx_mlp_grad(dp.ptr, _ptr_at(df, b0 * N_FEATURES),
                                                 _ptr_at(dy, b0), _ptr_at(dw, b0), batch,
                                                 dg.ptr, dl.ptr, None), "mlp grad")
                abi.mlp_adam(dp, dg, dm, dv, step, float(lr_ep))
                if b0 == 0:
                    losses.append(float(dl.numpy()[0]))
            history.append(float(np.mean(losses)) if losses else float("nan"))
            if log:
                log(f"epoch {ep} loss {history[-1]:.5f}")
        self.params = dp.numpy()
        return history

    def predict_device(self, dfeat, P):
        dp = abi.DeviceArray.from_numpy(self.params)
        dout = abi.DeviceArray((P,), np.float32)
        abi.mlp_forward(dp, dfeat, P, dout)
        return dout

    def predict(self, feat):
        feat = np.ascontiguousarray(feat, np.uint16)
        df = abi.DeviceArray.from_numpy(feat)
        return self.predict_device(df, len(feat)).numpy()

    def calibrate(self, pred, y, recall=0.999, margin=0.01):
        """Threshold on the prediction that keeps `recall` of the held-out
        pairs with y > GC_THRESHOLD, lowered by `margin` (GC units) against
        the sampling noise of that quantile."""
        pos = np.sort(pred[y > GC_THRESHOLD])
        if len(pos) == 0:
            self.threshold = float(np.min(pred))
            return self.threshold
        i = int(np.floor((1.0 - recall) * len(pos)))
        self.threshold = float(pos[max(i, 0)]) - margin
        return self.threshold

    def save(self, path):
        np.savez(path, params=self.params, threshold=np.float32(self.threshold or 0.0))

    @classmethod
    def load(cls, path):
        z = np.load(path)
        s = cls(z["params"])
        s.threshold = float(z["threshold"])
        return s


def _ptr_at(darr, elem_offset):
    """Device pointer `elem_offset` elements into a DeviceArray."""
    import ctypes
    return ctypes.c_void_p(darr.ptr.value + darr.dtype.itemsize * elem_offset)


def init_params(seed=0):
    """He-style initialisation in the parameter layout of screener.hip."""
    F, H = N_FEATURES, 64
    rng = np.random.default_rng(seed)
    p = np.zeros(abi.mlp_param_count(), np.float32)
    p[0:F * H] = rng.standard_normal(F * H) * np.sqrt(2.0 / F)
    o = F * H + H
    p[o:o + H * H] = rng.standard_normal(H * H) * np.sqrt(2.0 / H)
    o += H * H + H
    p[o:o + H] = rng.standard_normal(H) * 0.01  # start close to the skip (feature 5)
    return p


def min_fraction_at_recall(score, y, recall=0.995):
    """Smallest fraction of pairs a score threshold must keep so that
    `recall` of the pairs with y > GC_THRESHOLD are kept (measured, not
    calibrated: compares screening rules on the same data)."""
    pos = np.sort(score[y > GC_THRESHOLD])
    if len(pos) == 0:
        return 0.0
    t = pos[int(np.floor((1.0 - recall) * len(pos)))]
    return float(np.mean(score >= t))


def training_set(trajectories=8, N=1000, n=5000, pairs_per=60000, k=6, seed=100, log=None):
    """Exact GC of random pairs from `trajectories` seeded C-alpha-like
    trajectories (the plain path is the teacher). Returns (feat, y)."""
    feats, ys = [], []
    d = 3
    for t in range(trajectories):
        X, _ = synth.calpha_trajectory(N=N, n=n, seed=seed + t)
        ab = synth.all_pairs(N)
        rng = np.random.default_rng(seed + 1000 + t)
        ab = ab[np.sort(rng.choice(len(ab), size=min(pairs_per, len(ab)), replace=False))]
        dX = abi.DeviceArray.from_numpy(X)
        dab = abi.DeviceArray.from_numpy(ab)
        t0 = time.perf_counter()
        y = exact_gc(dX, X.shape[1], n, d, n, ab, k)
        f = features(dX, X.shape[1], n, N, n, dab, len(ab)).numpy()
        feats.append(f)
        ys.append(y)
        if log:
            log(f"trajectory {t}: {len(ab)} pairs, {np.mean(y > GC_THRESHOLD):.3f} above "
                f"{GC_THRESHOLD}, {time.perf_counter() - t0:.1f} s")
    return np.concatenate(feats), np.concatenate(ys)


def train_screener(feat, y, recall=0.999, epochs=40, seed=0, log=None):
    """Fit on 90% of the pairs, calibrate the threshold on the other 10%."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(y))
    cut = int(0.9 * len(y))
    tr, va = idx[:cut], idx[cut:]
    s = Screener(seed=seed)
    hist = s.fit(feat[tr], y[tr], epochs=epochs, seed=seed, log=log)
    pv = s.predict(feat[va])
    s.calibrate(pv, y[va], recall=recall)
    return s, {"history": hist, "val_pred": pv, "val_y": y[va], "val_feat": feat[va]}


def baseline_threshold(feat, y, recall=0.999, column=5):
    """The rule without the network: a threshold on one feature (default:
    the Gaussian-MI GC, column 5) calibrated to the same recall."""
    v = bf16_to_float(feat[:, column])
    pos = np.sort(v[y > GC_THRESHOLD])
    i = int(np.floor((1.0 - recall) * len(pos)))
    return float(pos[max(i, 0)]) - 1e-6, v


def screened_gc(X, N, n, k, screener, ab=None, audit_rate=0.01, min_recall=0.995, seed=0,
                dX=None, log=None):
    """All-pairs GC through the screener. Returns a dict with `gc` (float32,
    exact for computed pairs, NaN for pairs screened out), `computed`
    (bool), timings and audit statistics."""
    d = 3
    ld = X.shape[1]
    ab = synth.all_pairs(N) if ab is None else np.ascontiguousarray(ab, np.int32)
    P = len(ab)
    if dX is None:
        dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    abi.sync()
    t0 = time.perf_counter()
    dfeat = features(dX, ld, n, N, n, dab, P)
    dpred = screener.predict_device(dfeat, P)
    dsel = abi.DeviceArray((P, 2), np.int32)
    didx = abi.DeviceArray((P,), np.int64)
    dflag = abi.DeviceArray((P,), np.uint8)
    n_sel, n_audit = abi.screen_select(dpred, P, screener.threshold, audit_rate, seed, dab, dsel,
                                       didx, dflag)
    abi.sync()
    t_screen = time.perf_counter() - t0
    dmi = abi.DeviceArray((max(n_sel, 1),), np.float32)
    if n_sel:
        abi.ksg_mi(dX, ld, n, d, n, dsel, n_sel, k, dmi)
    mi = dmi.numpy()[:n_sel]
    gc_sel = abi.gc_from_mi(mi, d)
    idx = didx.numpy()[:n_sel]
    flag = dflag.numpy()
    t_exact = time.perf_counter() - t0 - t_screen
    gc = np.full(P, np.nan, np.float32)
    gc[idx] = gc_sel
    is_audit = flag[idx] == 2
    pos_found = int(np.sum((gc_sel > GC_THRESHOLD) & ~is_audit))
    audit_pos = int(np.sum((gc_sel > GC_THRESHOLD) & is_audit))
    n_cand = n_sel - n_audit
    non_cand = P - n_cand
    missed_est = audit_pos * (non_cand / max(n_audit, 1))
    recall_est = pos_found / max(pos_found + missed_est, 1e-9) if pos_found else 1.0
    fallback = recall_est < min_recall
    t_fallback = 0.0
    if fallback:
        rest = np.nonzero(np.isnan(gc))[0]
        t1 = time.perf_counter()
        gc[rest] = exact_gc(dX, ld, n, d, n, ab[rest], k)
        t_fallback = time.perf_counter() - t1
        if log:
            log(f"fallback: estimated recall {recall_est:.4f} < {min_recall}; "
                f"{len(rest)} more pairs computed")
    return {"gc": gc, "computed": ~np.isnan(gc), "pairs": P, "candidates": n_cand,
            "audit": n_audit, "audit_positives": audit_pos, "positives_found": pos_found,
            "recall_estimate": recall_est, "fallback": bool(fallback),
            "t_screen_s": t_screen, "t_exact_s": t_exact, "t_fallback_s": t_fallback,
            "t_total_s": t_screen + t_exact + t_fallback}
