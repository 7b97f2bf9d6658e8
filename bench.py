"""Benchmark: all-pairs generalized correlation on one MI355X.

Workload (the configuration the headline metric is quoted on): N=1000
C-alpha-like nodes, d=3, n=2000 frames, k=6, float32, all 499,500 unordered
pairs (netsci_amd.synth.calpha_trajectory, seed 1). One step = one pass of
the KSG GC pair loop over all pairs with X and the pair list already resident
in HBM: the persistent KSG launch, the copy of the MI values to the host and
the host double-precision GC transform (what netcalc.generalizedCorrelation
does after its uploads).

Prints one JSON line with the driver's fields plus:
  roofline     — the dominant kernel's algorithmic FP32 VALU rate against
                 MI355X's non-packed VALU peak (256 CUs x 64 lanes x 2.4 GHz =
                 39.32 T lane-ops/s): on the sorted-list path (the default)
                 the walk kernel, 4d+k+5 lane-ops per list entry walked
                 over its hipEvent-timed duration; on the all-frames path
                 (8d+2)*n*(n-1) lane-ops per pair; traffic from the committed
                 rocprofv3 counter summary when present;
  all_frames_equivalent — the reference algorithm's work per pair over the
                 whole call, reported beside the roofline, not folded in;
  cpu_baseline — the oracle (the reference CPU path restated: std::sort per
                 frame, single thread) timed on a sample of the same pairs;
  screened     — the screened all-pairs pass on the same workload (screener
                 trained on 8 other trajectories of the same size), with its
                 speedup over the plain pass and its recall of GC > 0.3.
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
HBM_PEAK = 8.0e12  # B/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nodes", type=int, default=1000)
    ap.add_argument("--frames", type=int, default=2000)
    ap.add_argument("--k", type=int, default=6)
    ap.add_argument("--cpu-pairs", type=int, default=128,
                    help="pairs timed for the CPU baseline, about 0.16 s each on one core "
                         "(0 to skip)")
    ap.add_argument("--screen-nodes", type=int, default=4000)
    ap.add_argument("--boot-nodes", type=int, default=512,
                    help="nodes of the bootstrap workload (all their pairs in each replicate)")
    ap.add_argument("--workload", default="gc", choices=["gc", "bootstrap", "apsp", "screened"],
                    help="gc: the headline all-pairs GC pass (default); bootstrap: 64 uneven "
                         "frame windows of a 512-node, 4000-frame trajectory, one graph-captured "
                         "batched launch vs one launch per replicate; apsp: Hedetniemi on the "
                         "-log(GC) contact graph of the gc workload")
    ap.add_argument("--checkpoint-freq", type=int, default=10000,
                    help="gc workload: also time netcalc.generalizedCorrelationWithCheckpointing "
                         "at this frequency against the plain netcalc call (0 to skip)")
    ap.add_argument("--no-screen", action="store_true",
                    help="gc workload: skip the screened pass (screener trained on 8 other "
                         "trajectories of the same size, then all pairs through it)")
    ap.add_argument("--method", choices=["auto", "lists", "all-frames"], default="auto",
                    help="KSG algorithm of the plain pass (identical results)")
    ap.add_argument("--walk-waves", type=int, default=0,
                    help="list path: waves per walk workgroup (0 = default)")
    ap.add_argument("--walk-batch", type=int, default=0,
                    help="list path: finished lanes per refill round (0 = default)")
    ap.add_argument("--pass2", choices=["default", "plain", "symmetric"], default="default",
                    help="KSG pass 2 for the gc workload: the library default, each ordered "
                         "frame pair, or each unordered pair once (both frames credited)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only time the VALU throughput probe (v_sub, v_max3, v_fma, KSG mix)")
    ap.add_argument("--apsp-reps", type=int, default=7,
                    help="timed calls of each Hedetniemi entry point (median reported)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


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
    if args.valu_probe:
        # The probe is a separate measurement library (libnsx_probe.so),
        # not part of libnsx's C ABI.
        import ctypes
        probe = ctypes.CDLL(os.path.join(ROOT, "netsci_amd", "libnsx_probe.so"))
        probe.nsx_valu_probe.argtypes = [ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_longlong)]
        res = {}
        for kind, name in enumerate(["v_sub_f32", "v_max3_f32+v_sub_f32", "v_fma_f32", "ksg_mix"]):
            ms, lanes = ctypes.c_double(0), ctypes.c_longlong(0)
            rc = probe.nsx_valu_probe(kind, 20000, ctypes.byref(ms), ctypes.byref(lanes))
            if rc:
                raise RuntimeError(f"nsx_valu_probe failed: {rc}")
            ms, lanes = ms.value, lanes.value
            res[name] = {"ms": ms, "lanes": lanes}
            log(f"probe {name}: {ms:.2f} ms")
        print(json.dumps({"valu_probe": res, "iters": 20000, "ops_per_iter": 32}), flush=True)
        return
    if args.workload == "bootstrap":
        return bench_bootstrap(args, abi, synth)
    if args.workload == "apsp":
        return bench_apsp(args, abi, synth)
    if args.workload == "screened":
        return bench_screened(args, abi, synth)
    dev_name, cus, _ = abi.device_info()
    if args.pass2 != "default":
        abi.set_config(2, 128, 4, 0, 1 if args.pass2 == "symmetric" else 0)
    abi.set_method({"auto": abi.METHOD_AUTO, "lists": abi.METHOD_LISTS,
                    "all-frames": abi.METHOD_ALL_FRAMES}[args.method],
                   args.walk_waves, args.walk_batch)
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
        t = time.perf_counter()
        timer.start()
        abi.ksg_mi(dX, X.shape[1], n, d, n, dab, P, k, dmi)
        timer.stop()
        mi = dmi.numpy()
        gc = abi.gc_from_mi(mi, d)
        return mi, gc, timer.ms(), 1e3 * (time.perf_counter() - t)

    def phase_line(i, kms, sms):
        ph = abi.last_phase_ms() if abi.last_method()[0] == 2 else {}
        parts = " ".join(f"{key} {v:.1f}" for key, v in ph.items())
        log(f"step {i}: {sms:.1f} ms (call {abi.last_call_ms():.1f}, stream {kms:.1f}; "
            f"{parts})")
        return ph

    print(f"bench: {dev_name}, {cus} CUs; {P} pairs, n={n}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        _, _, kms, sms = step()
        phase_line(f"warmup {w}", kms, sms)
    abi.sync()
    kernel_ms, step_ms, step_phases, call_ms = [], [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        mi, gc, kms, sms = step()
        kernel_ms.append(kms)
        step_ms.append(sms)
        call_ms.append(abi.last_call_ms())
        step_phases.append(phase_line(i, kms, sms))
    abi.sync()
    wall = time.perf_counter() - t0
    ties = abi.last_tie_frames()
    method, list_entries = abi.last_method()

    ms_per_step = 1e3 * wall / args.steps
    value = P * args.steps / wall
    kms = float(np.median(kernel_ms))
    # The reference algorithm's work (every ordered frame pair, both passes):
    # reported beside the roofline, never folded into it.
    work_per_pair = (8 * d + 2) * n * (n - 1)
    equivalent = work_per_pair * P / (kms * 1e-3)
    # Median over the timed steps of every phase (each step's phases are in
    # steps_ms below).
    phases = ({key: float(np.median([p[key] for p in step_phases])) for key in step_phases[0]}
              if method == 2 else None)
    if method == 2:
        # Sorted-list path: the walk kernel dominates. Its algorithmic work
        # per list entry: both marginal distances (2d subtractions, 2(d-1)
        # maxima; absolute values are operand modifiers), the joint max, the
        # seen test, the distance code (saturating subtract, shift,
        # shift-or), the select of the entry or the no-op value, and the
        # (k+1)-slot insertion of 32-bit (code, frame) entries (one med3 per
        # slot): 4d+k+5.
        per_entry = 4 * d + k + 5
        walk_ms = phases["walk"]
        achieved = per_entry * list_entries / (walk_ms * 1e-3)
        roofline = {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                    "unit": "T lane-ops/s", "frac": achieved / VALU_PEAK,
                    "traffic": traffic_from_profile("ksg_walk_c1_traffic.json"),
                    "kernel": f"walk_kernel<{d}, {k}>", "kernel_ms": walk_ms,
                    "work": f"4d+k+5 = {per_entry} VALU lane-ops per list entry walked, "
                            f"{list_entries} entries"}
    else:
        roofline = {"bound": "valu", "achieved": equivalent / 1e12, "peak": VALU_PEAK / 1e12,
                    "unit": "T lane-ops/s", "frac": equivalent / VALU_PEAK,
                    "traffic": traffic_from_profile("ksg_c1_traffic.json"),
                    "kernel": f"ksg_kernel<{d}, {k}, 2>", "kernel_ms": kms,
                    "work": f"(8d+2)*n*(n-1) = {work_per_pair} FP32 VALU lane-ops per pair"}

    cpu = None
    if args.cpu_pairs > 0:
        import oracle
        rng = np.random.default_rng(0)
        sample = ab[rng.choice(P, size=min(args.cpu_pairs, P), replace=False)]
        ref = np.empty(len(sample), np.float32)
        per_pair = np.empty(len(sample))
        for i in range(len(sample)):
            t1 = time.perf_counter()
            ref[i] = oracle.pairs(X, sample[i:i + 1], k, n, d, gc=True)[0]
            per_pair[i] = time.perf_counter() - t1
        cpu_s = float(per_pair.sum())
        idx = {tuple(r): i for i, r in enumerate(ab)}
        match = all(ref[i] == gc[idx[tuple(r)]] for i, r in enumerate(sample))
        cpu = {"value": len(sample) / cpu_s, "unit": "GC node-pairs/s", "cores": 1,
               "kind": "port",
               "sample": f"{len(sample)} seeded random pairs of the same workload, {cpu_s:.1f} s "
                         f"on one core of {host_cpu_model()} ({os.cpu_count()} logical CPUs "
                         f"visible); per pair {1e3 * per_pair.min():.0f}-"
                         f"{1e3 * per_pair.max():.0f} ms; extrapolated all {P} pairs "
                         f"{P * cpu_s / len(sample) / 3600:.1f} h; GPU values bit-identical: "
                         f"{match}"}

    checkpointed = None
    if args.checkpoint_freq > 0:
        checkpointed = checkpointed_at_workload(X, ab, n, d, k, gc, args.checkpoint_freq)

    screened = None
    if not args.no_screen:
        screened = screened_at_workload(abi, X, ab, N, n, k, dX, gc, float(np.median(step_ms)))

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
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tie_frames": ties,
        "ksg_method": {1: "all-frames", 2: "sorted lists"}.get(method, "none"),
        "list_entries": list_entries,
        "phases_ms": phases,
        "call_ms": kms,
        "steps_ms": {"step": [round(v, 2) for v in step_ms],
                     "call": [round(v, 2) for v in call_ms],
                     "median": float(np.median(step_ms)),
                     "max_over_median": float(np.max(step_ms) / np.median(step_ms))},
        "all_frames_equivalent": {
            "rate": equivalent / 1e12, "unit": "T lane-ops/s",
            "multiple_of_valu_peak": equivalent / VALU_PEAK,
            "work": f"(8d+2)*n*(n-1) = {work_per_pair} lane-ops per pair, the reference "
                    "algorithm's comparisons; a rate above the peak is the sorted lists' saving"},
        "screened": screened,
        "checkpointed": checkpointed,
    }
    print(json.dumps(out), flush=True)


def checkpointed_at_workload(X, ab, n, d, k, plain_gc, freq):
    """The checkpointed pair loop (generalized_correlation.cpp:8-101: R
    saved as <base>_<i>.npy after every freq-th pair) against the plain one,
    both through netcalc from host arrays (uploads included), outside the
    headline's timed region; one warm-up each, then one timed call."""
    import tempfile
    from netsci_amd import cuarray, netcalc
    Xc, abc = cuarray.FloatCuArray(), cuarray.IntCuArray()
    Xc.fromNumpy2D(X)
    abc.fromNumpy2D(ab)
    times = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name in ("plain", "checkpointed", "plain", "checkpointed"):
            R = cuarray.FloatCuArray()
            t0 = time.perf_counter()
            if name == "plain":
                netcalc.generalizedCorrelation(X=Xc, R=R, ab=abc, k=k, n=n, xd=2, d=d,
                                               platform=netcalc.GPU_PLATFORM)
            else:
                netcalc.generalizedCorrelationWithCheckpointing(
                    X=Xc, R=R, ab=abc, k=k, n=n, xd=2, d=d, platform=netcalc.GPU_PLATFORM,
                    checkpointFrequency=freq, checkpointFileName=os.path.join(tmp, "ck.npy"))
            times[name] = 1e3 * (time.perf_counter() - t0)
            same = bool(np.array_equal(R.toNumpy1D(), plain_gc))
            if not same:
                log(f"{name}: values differ from the plain pass")
        files = sorted(os.listdir(tmp))
    log(f"checkpointed {times['checkpointed']:.1f} ms vs plain {times['plain']:.1f} ms "
        f"(frequency {freq}, {len(ab) // freq + 1} checkpoints)")
    return {"frequency": freq, "plain_ms": times["plain"], "checkpointed_ms": times["checkpointed"],
            "ratio": times["checkpointed"] / times["plain"], "files_left": files,
            "values_equal_plain": same}


def screened_at_workload(abi, X, ab, N, n, k, dX, plain_gc, plain_ms):
    """The screened all-pairs pass on the headline workload itself: a
    screener trained on exact GC of 60,000 random pairs from each of 8 other
    seeded trajectories of the same size (seeds 100-107), then every pair of
    this X through it (features, MLP, selection, exact KSG on candidates plus
    a 1% audit). Outside the plain pass's timed region; one warm-up pass,
    then one timed pass. Recall is against the plain pass's exact GC; the
    speedup divides the median plain step by the screened pass."""
    from netsci_amd import screen
    t0 = time.perf_counter()
    feat, y = screen.training_set(trajectories=8, N=N, n=n, pairs_per=60000, k=k, seed=100)
    model, _ = screen.train_screener(feat, y, recall=0.999, epochs=40, seed=0)
    t_train = time.perf_counter() - t0
    log(f"screener trained in {t_train:.1f} s, threshold {model.threshold:.4f}")
    screen.screened_gc(X, N, n, k, model, ab=ab, audit_rate=0.01, seed=1, dX=dX)
    abi.sync()
    t1 = time.perf_counter()
    res = screen.screened_gc(X, N, n, k, model, ab=ab, audit_rate=0.01, seed=1, dX=dX)
    abi.sync()
    t_scr = time.perf_counter() - t1
    comp = res["computed"]
    pos = plain_gc > screen.GC_THRESHOLD
    recall = float(np.sum(pos & comp) / max(int(pos.sum()), 1))
    log(f"screened pass {1e3 * t_scr:.1f} ms, recall {recall:.5f}, computed {comp.mean():.4f}")
    return {"pairs_per_s": len(ab) / t_scr, "ms": 1e3 * t_scr,
            "speedup_vs_plain": plain_ms / (1e3 * t_scr),
            "recall_gc_gt_0.3": recall, "positives": int(pos.sum()),
            "computed_fraction": float(comp.mean()),
            "computed_values_bit_identical": bool(np.array_equal(res["gc"][comp],
                                                                 plain_gc[comp])),
            "fallback": bool(res.get("fallback", False)),
            "training": f"8 trajectories x 60000 pairs, N={N}, n={n}, seeds 100-107",
            "training_s": t_train}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def bench_bootstrap(args, abi, synth):
    """64 bootstrap replicates as contiguous windows (n_r ~ U[500, 4000]) of
    a 512-node (--boot-nodes), 4000-frame C-alpha-like trajectory, all pairs
    (130,816 at 512 nodes) in each. Timed: the hipGraph-captured plan on the
    sorted-list path (one graph launch builds every window's tables and
    walks all its pairs; no host synchronisation inside), the eager windowed
    call on the same path, one call per replicate (the reference runs its
    pair loop once per replicate), and the all-frames graph plan (autotuned
    persistent launch). All four must give identical outputs."""
    N, nT, d, k, R = args.boot_nodes, 4000, 3, args.k, 64
    X, _ = synth.calpha_trajectory(N=N, n=nT, seed=7)
    ab = synth.all_pairs(N)
    P = len(ab)
    win = synth.bootstrap_windows(R=R, n_total=nT, lo=500, hi=4000, seed=3)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    timer = abi.Timer()
    outs = {}

    def timed(name, fn, reps):
        out = []
        for _ in range(reps):
            timer.start()
            fn()
            timer.stop()
            out.append(timer.ms())
            log(f"{name} {out[-1]:.1f} ms")
        return float(np.min(out)), out

    # List-path graph plan.
    outs["graph_lists"] = abi.DeviceArray((R, P), np.float32)
    t0 = time.perf_counter()
    plan = abi.KsgPlan(dX, X.shape[1], nT, d, win, dab, P, k, outs["graph_lists"],
                       autotune=False)
    create_s = time.perf_counter() - t0
    assert plan.method() == 2
    plan.launch()
    abi.sync()
    t_graph, graph_ms = timed("list-path graph replay", plan.launch, args.steps)
    ties = plan.tie_frames()
    plan.close()

    # Eager windows on the list path.
    outs["eager_lists"] = abi.DeviceArray((R, P), np.float32)
    run_w = lambda: abi.ksg_mi_windows(dX, X.shape[1], nT, d, win, dab, P, k,
                                       outs["eager_lists"])
    run_w()
    t_eager, eager_ms = timed("eager windows (list path)", run_w, args.steps)
    lmethod, lentries = abi.last_method()
    phases = abi.last_phase_ms()

    # One call per replicate (list path).
    outs["per_replicate"] = abi.DeviceArray((R, P), np.float32)

    def per_rep():
        for r, (off, ln) in enumerate(win):
            abi.ksg_mi(dX, X.shape[1], nT, d, int(ln), dab, P, k, outs["per_replicate"],
                       frame_offset=int(off), out_offset=r * P)
    t_rep, _ = timed("one call per replicate", per_rep, 1)

    # All-frames graph plan (autotuned).
    outs["graph_all_frames"] = abi.DeviceArray((R, P), np.float32)
    abi.set_method(abi.METHOD_ALL_FRAMES)
    t0 = time.perf_counter()
    aplan = abi.KsgPlan(dX, X.shape[1], nT, d, win, dab, P, k, outs["graph_all_frames"],
                        autotune=True)
    tune_s = time.perf_counter() - t0
    abi.set_method(abi.METHOD_AUTO)
    cfg = aplan.config()
    t_all, _ = timed("all-frames graph replay", aplan.launch, 1)
    aplan.close()

    ref = outs["graph_lists"].numpy()
    same = all(np.array_equal(ref, o.numpy()) for o in outs.values())
    work = float(sum(int(l) * (int(l) - 1) for _, l in win)) * P * (8 * d + 2)
    out = {
        "metric": "bootstrap replicate node-pairs/s",
        "value": R * P / (t_graph * 1e-3), "unit": "replicate node-pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": t_graph, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"bootstrap GC: {N} nodes, d={d}, k={k}, {R} windows "
                               f"n_r ~ U[500,4000] of a {nT}-frame trajectory, {P} pairs each"},
        "graph_lists_ms": t_graph, "graph_lists_steps_ms": graph_ms,
        "graph_lists_create_s": create_s,
        "eager_lists_ms": t_eager, "eager_lists_steps_ms": eager_ms,
        "eager_lists_phases_ms": phases, "lists_entries": lentries,
        "per_replicate_lists_ms": t_rep,
        "graph_all_frames_ms": t_all, "graph_all_frames_plan": cfg, "tune_s": tune_s,
        "graph_vs_per_replicate": t_rep / t_graph, "graph_vs_eager_windows": t_eager / t_graph,
        "lists_vs_all_frames_graph": t_all / t_graph,
        "outputs_identical": same, "tie_frames": ties,
        "all_frames_equivalent": {"rate": work / (t_graph * 1e-3) / 1e12,
                                  "unit": "T lane-ops/s",
                                  "multiple_of_valu_peak": work / (t_graph * 1e-3) / VALU_PEAK},
    }
    print(json.dumps(out), flush=True)


def bench_screened(args, abi, synth):
    """Screened vs plain all-pairs GC at N=4000, d=3, n=5000, k=6
    (7,998,000 pairs). The screener is trained first on exact GC of 60,000
    random pairs from each of 8 seeded 1000-node, 5000-frame trajectories of
    the same generator (held-out 10% calibrate the threshold for 99.9%
    recall). Then, on a fresh 4000-node trajectory: the plain pass on the
    sorted-list path over every pair (the fastest plain GPU path, and the
    recall reference; its tables, 400 GB for all nodes, run on the blocked
    schedule; the faster of two calls), the all-frames kernel timed on a 200,000-pair sample and
    extrapolated, and the screened pass (features, MLP, selection, exact KSG
    on candidates plus a 1% audit). The speedup divides the fastest plain
    pass by the screened pass."""
    from netsci_amd import screen
    N, n, d, k = args.screen_nodes, 5000, 3, args.k
    t0 = time.perf_counter()
    feat, y = screen.training_set(trajectories=8, N=1000, n=n, pairs_per=60000, k=k, seed=100,
                                  log=log)
    t_data = time.perf_counter() - t0
    s, info = screen.train_screener(feat, y, recall=0.999, epochs=40, seed=0)
    t_train = time.perf_counter() - t0 - t_data
    log(f"training: {len(y)} pairs ({t_data:.1f} s exact GC + features), fit {t_train:.1f} s, "
        f"loss {info['history'][0]:.5f} -> {info['history'][-1]:.5f}, threshold {s.threshold:.4f}")

    X, _ = synth.calpha_trajectory(N=N, n=n, seed=9001)
    ab = synth.all_pairs(N)
    P = len(ab)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(ab)
    dmi = abi.DeviceArray((P,), np.float32)
    timer = abi.Timer()
    # Plain pass on the sorted-list path, every pair, one call, twice: the
    # first call also grows the library's table buffer to its budget (a
    # fresh allocation of about 240 GB), the second runs as the screened
    # pass later does, on the buffer already held.
    plain_calls = []
    for call in range(2):
        log(f"plain pass (sorted lists), {P} pairs, call {call + 1}")
        timer.start()
        abi.ksg_mi(dX, X.shape[1], n, d, n, dab, P, k, dmi)
        timer.stop()
        plain_calls.append(timer.ms() * 1e-3)
        plain_method = abi.last_method()[0]
        plain_phases = abi.last_phase_ms()
        blocks, slots, builds = abi.last_list_info()
        log(f"plain {plain_calls[-1]:.2f} s (method {plain_method}, {blocks} node blocks, "
            f"{builds} node tables built, phases {plain_phases})")
    t_plain = min(plain_calls)
    plain = abi.gc_from_mi(dmi.numpy(), d)
    # All-frames kernel on a seeded sample, extrapolated.
    rng = np.random.default_rng(5)
    sample = ab[np.sort(rng.choice(P, size=200000, replace=False))]
    dsm = abi.DeviceArray.from_numpy(sample)
    dsmi = abi.DeviceArray((len(sample),), np.float32)
    abi.set_method(abi.METHOD_ALL_FRAMES)
    timer.start()
    abi.ksg_mi(dX, X.shape[1], n, d, n, dsm, len(sample), k, dsmi)
    timer.stop()
    abi.set_method(abi.METHOD_AUTO)
    t_af = timer.ms() * 1e-3 * P / len(sample)
    idx = np.searchsorted(ab[:, 0].astype(np.int64) * N + ab[:, 1],
                          sample[:, 0].astype(np.int64) * N + sample[:, 1])
    af_same = bool(np.array_equal(dsmi.numpy(), dmi.numpy()[idx]))
    log(f"all-frames plain, extrapolated from {len(sample)} pairs: {t_af:.1f} s "
        f"(values equal the list path: {af_same})")
    # Screened pass (one warm-up on a different seed for the audit sample).
    t1 = time.perf_counter()
    res = screen.screened_gc(X, N, n, k, s, ab=ab, audit_rate=0.01, seed=1, dX=dX, log=log)
    t_scr = time.perf_counter() - t1
    scr_phases = abi.last_phase_ms()
    scr_blocks, _, scr_builds = abi.last_list_info()
    comp = res["computed"]
    pos = plain > screen.GC_THRESHOLD
    recall = float(np.sum(pos & comp) / max(pos.sum(), 1))
    exact = bool(np.array_equal(res["gc"][comp], plain[comp]))
    log(f"screened {t_scr:.2f} s (screen {res['t_screen_s']:.2f} s, exact {res['t_exact_s']:.2f} s),"
        f" recall {recall:.5f}, computed {comp.mean():.4f}")
    fastest = min(t_plain, t_af)
    out = {
        "metric": "screened vs plain all-pairs GC speedup at >=99.5% recall",
        "value": fastest / t_scr, "unit": "x", "higher_is_better": True, "n_gpus": 1,
        "dtype": "f32 (KSG), bf16 MFMA (screener)", "data": "synthetic",
        "config": {"workload": f"all-pairs GC N={N}, d={d}, n={n}, k={k}, {P} pairs",
                   "training": "8 trajectories x 60000 pairs, N=1000, n=5000"},
        "plain_lists_s": t_plain, "plain_lists_calls_s": plain_calls,
        "plain_lists_phases_ms": plain_phases,
        "plain_lists_node_blocks": blocks, "plain_lists_node_tables_built": builds,
        "plain_all_frames_s_extrapolated": t_af,
        "plain_all_frames_sample": f"{len(sample)} seeded pairs, equal to the list path: {af_same}",
        "screened_s": t_scr, "speedup_vs_lists_plain": t_plain / t_scr,
        "speedup_vs_all_frames_plain": t_af / t_scr,
        "plain_pairs_per_s": P / fastest, "screened_pairs_per_s": P / t_scr,
        "screened_exact_phases_ms": scr_phases, "screened_exact_node_blocks": scr_blocks,
        "screened_exact_node_tables_built": scr_builds,
        "recall_gc_gt_0.3": recall, "positives": int(pos.sum()),
        "computed_fraction": float(comp.mean()), "computed_values_bit_identical": exact,
        "screen": {k2: v for k2, v in res.items() if k2 not in ("gc", "computed")},
        "training_s": t_data + t_train,
    }
    print(json.dumps(out), flush=True)


def bench_apsp(args, abi, synth):
    """Hedetniemi APSP with paths on the contact graph of the gc workload:
    A_ij = -log(GC_ij) where the mean C-alpha distance is below 7.5 A,
    maxPathLength = hop diameter + 1."""
    N, n, d, k = args.nodes, args.frames, 3, args.k
    X, mean = synth.calpha_trajectory(N=N, n=n, seed=1)
    pairs = synth.contact_pairs(mean)
    dX = abi.DeviceArray.from_numpy(X)
    dab = abi.DeviceArray.from_numpy(pairs)
    dmi = abi.DeviceArray((len(pairs),), np.float32)
    abi.ksg_mi(dX, X.shape[1], n, d, n, dab, len(pairs), k, dmi)
    gc = abi.gc_from_mi(dmi.numpy(), d)
    A = synth.contact_adjacency(gc, pairs, N)
    L = synth.hop_diameter(A) + 1
    log(f"apsp: N={N}, {len(pairs)} contact pairs, L={L}")
    dA = abi.DeviceArray.from_numpy(A)
    dH = abi.DeviceArray((N, N), np.float32)
    dHi = abi.DeviceArray((L, N, N), np.float32)
    dP = abi.DeviceArray((N * L * N,), np.int32)
    timer = abi.Timer()
    abi.hedetniemi(dA, N, L, 1e-4, dH, dHi, dP)
    abi.sync()
    # Median of args.apsp_reps timed calls of each entry point (each call is
    # short, so single calls vary with the box's clocks).
    len_ms, all_ms = [], []
    for _ in range(args.apsp_reps):
        timer.start()
        abi.hedetniemi_lengths(dA, N, L, dH)
        timer.stop()
        len_ms.append(timer.ms())
    for _ in range(args.apsp_reps):
        timer.start()
        abi.hedetniemi(dA, N, L, 1e-4, dH, dHi, dP)
        timer.stop()
        all_ms.append(timer.ms())
    t_len, t_all = float(np.median(len_ms)), float(np.median(all_ms))
    log(f"lengths {t_len:.3f} ms, with paths {t_all:.3f} ms (medians of {args.apsp_reps})")
    # The library takes the neighbour-list step when the rows of A hold at
    # most 128 entries < +inf on average (a contact graph); its floor is HBM:
    # each step reads H_p and writes H_{p+1}. Dense graphs take the tiled
    # N^3 min-plus kernel, bound by FP32 VALU.
    per_row = (A < np.inf).sum(axis=1)
    degree = float(((per_row + 3) // 4 * 4).sum()) / N  # padded, as the library counts
    if degree <= 128:
        step_bytes = 2.0 * N * N * 4 * L
        roof = {"bound": "hbm", "achieved": step_bytes / (t_len * 1e-3) / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": step_bytes / (t_len * 1e-3) / HBM_PEAK,
                "work": f"sparse min-plus steps (mean degree {degree:.1f}): read H_p and "
                        f"write H_p+1, 2 N^2 x 4 B per step"}
    else:
        ops = 2.0 * N ** 3 * L
        roof = {"bound": "valu", "achieved": ops / (t_len * 1e-3) / 1e12,
                "peak": VALU_PEAK / 1e12, "unit": "T lane-ops/s",
                "frac": ops / (t_len * 1e-3) / VALU_PEAK,
                "work": "dense min-plus: 2 N^3 lane-ops (add, min) per step"}
    out = {"metric": "Hedetniemi APSP with paths", "value": t_all, "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"N={N} contact graph (-log GC, mean distance < 7.5 A), "
                                  f"L={L}", "contact_pairs": int(len(pairs))},
           "lengths_ms": t_len, "backtrack_ms": t_all - t_len,
           "calls_ms": {"lengths": [round(x, 4) for x in len_ms],
                        "with_paths": [round(x, 4) for x in all_ms]},
           "roofline": roof}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
