"""Benchmark: all-pairs generalized correlation on one MI355X.

Workload (the configuration the headline metric is quoted on): N=1000
C-alpha-like nodes, d=3, n=2000 frames, k=6, float32, all 499,500 unordered
pairs (netsci_amd.synth.calpha_trajectory, seed 1). One step = one pass of
the KSG GC pair loop over all pairs with X and the pair list already resident
in HBM: the persistent KSG launch, the copy of the MI values to the host and
the host double-precision GC transform (what netcalc.generalizedCorrelation
does after its uploads).

Prints one JSON line with the driver's fields plus:
  roofline     — the KSG kernel's algorithmic FP32 VALU rate, (8d+2)*n*(n-1)
                 lane-ops per pair over the kernel's event-timed duration,
                 against MI355X's non-packed VALU peak (256 CUs x 64 lanes x
                 2.4 GHz = 39.32 T lane-ops/s); traffic from the committed
                 rocprofv3 counter summary when present;
  cpu_baseline — the oracle (the reference CPU path restated: std::sort per
                 frame, single thread) timed on a sample of the same pairs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("GC node-pairs/s at N=1000,d=3,n=2000; screened vs plain at "
          "≥99.5% recall")
VALU_PEAK = 256 * 64 * 2.4e9  # lane-ops/s, non-packed


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nodes", type=int, default=1000)
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--k", type=int, default=6)
    ap.add_argument("--cpu-pairs", type=int, default=24,
                    help="pairs timed for the CPU baseline (0 to skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def traffic_from_profile(name):
    """HBM bytes per launch recorded by rocprofv3 --pmc (see profiles/)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(pairs shard across GPUs with no exchange, but only the one-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)

    from netsci_amd import abi, synth

    abi.init(0)
    if args.copy_calibration_mb:
        nbytes = args.copy_calibration_mb << 20
        src = abi.DeviceArray((nbytes // 4,), np.float32)
        dst = abi.DeviceArray((nbytes // 4,), np.float32)
        abi.check(abi.lib().nsx_memcpy_d2d(dst.ptr, src.ptr, nbytes, None), "copy")
        abi.sync()
        print(json.dumps({"copy_bytes": nbytes}), flush=True)
        return
    dev_name, cus, _ = abi.device_info()
    N, n, d, k = args.nodes, args.frames, 3, args.k
    X, _ = synth.calpha_trajectory(N=N, n=n, seed=1)
    ab = synth.all_pairs(N)
    P = len(ab)

    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((P,), np.float32)
    mi_host = np.empty(P, np.float32)
    timer = abi.Timer()

    def step():
        timer.start()
        abi.ksg_mi(dX, X.shape[1], n, d, n, dab, P, k, dmi)
        timer.stop()
        mi = dmi.numpy()
        gc = abi.gc_from_mi(mi, d)
        return mi, gc, timer.ms()

    print(f"bench: {dev_name}, {cus} CUs; {P} pairs, n={n}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        step()
        print(f"bench: warmup {w} done", file=sys.stderr, flush=True)
    abi.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mi, gc, kms = step()
        kernel_ms.append(kms)
        print(f"bench: step kernel {kms:.1f} ms", file=sys.stderr, flush=True)
    abi.sync()
    wall = time.perf_counter() - t0
    ties = abi.last_tie_frames()

    ms_per_step = 1e3 * wall / args.steps
    value = P * args.steps / wall
    work_per_pair = (8 * d + 2) * n * (n - 1)
    kms = float(np.mean(kernel_ms))
    achieved = work_per_pair * P / (kms * 1e-3)

    cpu = None
    if args.cpu_pairs > 0:
        import oracle
        rng = np.random.default_rng(0)
        sample = ab[rng.choice(P, size=min(args.cpu_pairs, P), replace=False)]
        t1 = time.perf_counter()
        ref = oracle.pairs(X, sample, k, n, d, gc=True)
        cpu_s = time.perf_counter() - t1
        idx = {tuple(r): i for i, r in enumerate(ab)}
        match = all(ref[i] == gc[idx[tuple(r)]] for i, r in enumerate(sample))
        cpu = {"value": len(sample) / cpu_s, "unit": "GC node-pairs/s", "cores": 1,
               "kind": "port",
               "sample": f"{len(sample)} seeded random pairs of the same workload, "
                         f"{cpu_s:.1f} s single-thread; GPU values bit-identical: {match}"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GC node-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"all-pairs KSG generalized correlation, N={N} C-alpha nodes, "
                               f"d={d}, n={n} frames, k={k}, {P} pairs (plain path)",
                   "nodes": N, "frames": n, "d": d, "k": k, "pairs": P,
                   "device": dev_name, "cus": cus},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "T lane-ops/s", "frac": achieved / VALU_PEAK,
                     "traffic": traffic_from_profile("ksg_c1_traffic.json"),
                     "kernel_ms": kms,
                     "work": f"(8d+2)*n*(n-1) = {work_per_pair} FP32 VALU lane-ops per pair"},
        "cpu_baseline": cpu,
        "tie_frames": ties,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
