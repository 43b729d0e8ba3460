"""Benchmark: points/s to the exact mutual-reachability MST + labels on one
MI355X.

Workload: the 2-D scale configuration -- n = 2,000,000 float32 points, 50
anisotropic Gaussians (Zipf 1.1 sizes, random rotations, axis std
log-uniform 0.1-2.0) + 10% uniform noise, seed 1; min_samples = 10.

One step = one pass over the batch with the points already resident in HBM:
KD-tree, core distances, kNN-seeded forest and all Boruvka rounds on the
device (edge list and core distances in HBM), the edge sort on the device,
the copy of the sorted edges to the host and the host label stage (single
linkage, condensed tree, EOM selection, labels and membership strengths,
min_cluster_size = 100).

Prints one JSON line (rank 0).  The roofline figure is for the dominant
kernel, timed with HIP events on the library's stream; the CPU baseline is
the oracle (a C restatement of the reference: its prange loops threaded, its
label stage compiled) on the same 2M-point workload, calibrated against the
reference's published timings.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per dispatch from two rocprofv3 --pmc runs (FETCH_SIZE, then
# WRITE_SIZE, each scaled by what the counter reads for a plain 256 MiB
# copy) on this workload; command lines and raw counts in
# profiles/pmc_2Mx2D_traffic.md.
PMC_TRAFFIC_BYTES = {
    "boruvka_wave_kernel<2, 2>": 77.62e6,
    "knn_kernel<2, false>": 260.93e6,
}


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=2_000_000)
    p.add_argument("--min-samples", type=int, default=10)
    p.add_argument("--min-cluster-size", type=int, default=100)
    p.add_argument("--cpu-sample", type=int, default=2_000_000,
                   help="points of the same workload timed on the oracle (0: skip)")
    p.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                   help="host threads for the oracle's kNN and round searches (the reference's prange loops)")
    p.add_argument("--copy-probe", action="store_true",
                   help="only run a plain 256 MiB device copy (calibrates HBM byte counters) and exit")
    p.add_argument("--no-forms", action="store_true",
                   help="skip the learned / high-dimensional / 10M x 3-D / PLSCAN-sweep / batch side measurements")
    p.add_argument("--forms-only", default=None,
                   help="comma-separated side measurements to run alone (" + ",".join(FORMS) +
                        "), printing their JSON object instead of the bench line (for profiling one form)")
    p.add_argument("--batch-one-call", type=int, default=None,
                   help="with --forms-only batched: one call on the 512 datasets with this many workers")
    return p.parse_args()


def copy_probe():
    """A plain device-to-device copy of known size: 256 MiB read, 256 MiB
    written per dispatch, repeated 5 times."""
    import torch

    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": src.numel() * 4, "dispatches": 5}))


def main():
    args = _parse()
    if args.forms_only:
        names = [f for f in args.forms_only.split(",") if f]
        bad = [f for f in names if f not in FORMS]
        if bad:
            raise SystemExit(f"unknown forms {bad}; choose from {list(FORMS)}")
        import torch

        from fast_hdbscan_amd import workloads

        Xd = None
        if "learned" in names or "serving" in names:
            Xd = torch.from_numpy(workloads.scale_2d(args.n, seed=1)).cuda()
        print(json.dumps(side_forms(Xd, args.min_samples, names, args.batch_one_call)))
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import fast_hdbscan_amd as F
    from fast_hdbscan_amd import workloads

    n, ms = args.n, args.min_samples
    X = workloads.scale_2d(n, seed=1)
    d = X.shape[1]
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    kn = ms if ms > 1 else 1
    out = {
        "src": torch.empty(n - 1, dtype=torch.int32, device=dev),
        "dst": torch.empty(n - 1, dtype=torch.int32, device=dev),
        "w_sq": torch.empty(n - 1, dtype=torch.float32, device=dev),
        "core_sq": torch.empty(n, dtype=torch.float32, device=dev),
        "neighbors": torch.empty((n, kn), dtype=torch.int32, device=dev),
    }
    torch.cuda.synchronize()

    from fast_hdbscan_amd.estimator import hdbscan_device

    mcs = args.min_cluster_size
    work = {"mst": out}
    stats = {}
    for _ in range(args.warmup):
        hdbscan_device(Xd, ms, mcs, stats=stats, work=work)
    torch.cuda.synchronize()

    per_step = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = {}
        labels, strengths = hdbscan_device(Xd, ms, mcs, stats=s, work=work)
        per_step.append(s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = n * args.steps / elapsed
    n_clusters = int(labels.max()) + 1
    noise = int((labels < 0).sum())

    # Dominant kernel: kNN vs the per-round main-pass nearest-edge search
    # (HIP events on the library stream).  Algorithmic bytes per launch:
    #   knn   : n*(4d) read + n*(k-1)*8 neighbour positions and distances
    #           + 4n core, all in tree order             (k = min_samples+1)
    #   query : n*(4d + 4 core + 4 label + 8 best)        (main pass, per round)
    k = ms + 1 if ms > 1 else 2
    knn_ms = float(np.mean([s["t_knn_ms"] for s in per_step]))
    rounds = per_step[-1]["rounds"]
    q_ms_total = float(np.mean([s["t_query_main_ms"] for s in per_step]))
    q_ms = q_ms_total / max(rounds, 1)
    knn_bytes = n * (4 * d) + n * (k - 1) * 8 + 4 * n
    q_bytes = n * (4 * d + 16)
    if knn_ms >= q_ms_total:
        dom, dur_ms, alg = "knn_kernel<2, false>", knn_ms, knn_bytes
    else:
        dom, dur_ms, alg = "boruvka_wave_kernel<2, 2>", q_ms, q_bytes
    achieved = alg / (dur_ms * 1e-3) / 1e9

    cpu = None
    if args.cpu_sample > 0:
        cpu = cpu_baseline(X, args.cpu_sample, ms, mcs, args.cpu_threads)

    pcie = pcie_inclusive(X, ms, mcs, work, max(2, args.steps // 2))

    forms = None if args.no_forms else side_forms(Xd, ms)

    line = {
        "metric": "points/s to exact MR-MST+labels at 2Mx2D",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "2-D anisotropic 50-Gaussian mixture + 10% noise, seed 1",
            "n": n,
            "d": d,
            "min_samples": ms,
            "min_cluster_size": mcs,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": PMC_TRAFFIC_BYTES.get(dom) if (n == 2_000_000 and ms == 10) else None,
            "avg_launch_ms": dur_ms,
            "alg_bytes_per_launch": alg,
        },
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "phases_ms": {key: float(np.mean([s[key] for s in per_step]))
                      for key in ("t_tree_ms", "t_knn_ms", "t_init_ms", "t_rounds_ms", "t_query_ms", "t_query_main_ms",
                                  "t_total_ms",
                                  "t_sort_copy_ms", "t_labels_ms", "t_linkage_ms", "t_condense_ms", "t_select_ms",
                                  "t_label_pass_ms")},
        "mst_points_per_s": n / (1e-3 * float(np.mean([s["t_total_ms"] for s in per_step]))),
        "rounds": rounds,
        "clusters": n_clusters,
        "noise_points": noise,
        "forms": forms,
    }
    print(json.dumps(line))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_fit(X, ms, mcs, threads):
    """The reference's HDBSCAN fit as the oracle restates it on a multi-core
    host: serial KD-tree build, kNN rows and each Boruvka round's searches
    over `threads` threads with shared component bounds (the reference's
    prange loops), serial merges, then the label stage compiled, one thread
    (as the reference's numba-compiled cluster_trees).  Returns seconds
    (tree, mst, labels)."""
    from oracle import labels as OL
    from oracle import oracle as O

    t0 = time.perf_counter()
    tree = O.build_kdtree(X)
    t1 = time.perf_counter()
    r = O.boruvka_racy(tree, ms, nthreads=threads)
    t2 = time.perf_counter()
    OL.clusters_from_spanning_tree_c(r.edges(), min_cluster_size=mcs)
    t3 = time.perf_counter()
    return t1 - t0, t2 - t1, t3 - t2


def cpu_baseline(X, sample, ms, mcs, threads):
    """The oracle on the same workload (median of 3 runs), and on the
    reference's published configurations for calibration."""
    import numpy as np

    from fast_hdbscan_amd import workloads

    Xs = X if sample == X.shape[0] else workloads.scale_2d(sample, seed=1)
    runs = [_oracle_fit(Xs, ms, mcs, threads) for _ in range(3)]
    tot = sorted(runs, key=sum)[1]
    # calibration: the reference's own published timings:
    # benchmarking/results.csv:58 (20k x 2-D blobs, 3 centres std 0.5,
    # standardised, min_samples 5, min_cluster_size 15: 0.0796 s) and
    # doc/benchmarks.ipynb cells 3-7 (1.28M x 2-D blobs, 100 centres,
    # std 1/9, min_samples 10, min_cluster_size 100: ~9.6 s on 8 cores)
    from sklearn.datasets import make_blobs
    from sklearn.preprocessing import StandardScaler

    cal = []
    X20 = StandardScaler().fit_transform(make_blobs(20_000, 2, centers=3, cluster_std=0.5, random_state=0)[0])
    X20 = X20.astype(np.float32)
    t20 = sorted(sum(_oracle_fit(X20, 5, 15, threads)) for _ in range(5))[2]
    cal.append({"config": "20k x 2-D blobs (results.csv:58), ms 5, mcs 15", "stand_in_s": t20,
                "published_s": 0.0796, "ratio": t20 / 0.0796})
    X128 = make_blobs(1_280_000, 2, centers=100, cluster_std=100 / 900, random_state=0)[0].astype(np.float32)
    t128 = sum(_oracle_fit(X128, 10, 100, threads))
    cal.append({"config": "1.28M x 2-D blobs, 100 centres (doc/benchmarks.ipynb), ms 10, mcs 100",
                "stand_in_s": t128, "published_s": 9.6, "published_hardware": "8 cores", "ratio": t128 / 9.6})
    return {
        "value": sample / sum(tot),
        "unit": "points/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "mst_only_value": sample / (tot[0] + tot[1]),
        "sample": f"the same workload at n={sample} (seed 1), min_samples={ms}, min_cluster_size={mcs}, median "
                  f"of 3: the oracle as the reference runs on a multi-core host -- serial KD-tree build "
                  f"({tot[0]:.2f} s), kNN and per-round searches over {threads} threads with shared component "
                  f"bounds, serial merges ({tot[1]:.2f} s), then the compiled label stage on one thread "
                  f"({tot[2]:.2f} s)",
        "calibration": cal,
    }


def pcie_inclusive(X, ms, mcs, work, steps):
    """Points/s when each step also copies the float32 points from pinned
    host memory to HBM (not the headline: the headline starts with the
    points resident in HBM)."""
    import torch

    from fast_hdbscan_amd.estimator import hdbscan_device

    Xh = torch.from_numpy(X).pin_memory()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Xd = Xh.to("cuda:0", non_blocking=False)
        hdbscan_device(Xd, ms, mcs, work=work)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"value": X.shape[0] * steps / el, "ms_per_step": 1e3 * el / steps, "steps": steps}


FORMS = ("learned", "high_dim", "scale_3d", "plscan_sweep", "batched", "serving")


def _mst_ms(X, m, reps=3, **kw):
    """Best wall time (ms) of `reps` device MSTs, with that run's stats."""
    import torch

    from fast_hdbscan_amd.hdbscan import mst_device

    best, st_best = None, None
    for _ in range(reps):
        st = {}
        torch.cuda.synchronize()
        t = time.perf_counter()
        mst_device(X, m, stats=st, **kw)
        torch.cuda.synchronize()
        el = 1e3 * (time.perf_counter() - t)
        if best is None or el < best:
            best, st_best = el, st
    return best, st_best


def form_learned(Xd, ms):
    """Learned search radii on the bench's own batch: plain, trained on
    every call, trained once and reused."""
    from fast_hdbscan_amd import _lib

    ctx = _lib.context(0)
    L = _lib.load()
    plain_ms, pst = _mst_ms(Xd, ms)
    res = {"plain": {"mst_ms": plain_ms, "knn_ms": pst["t_knn_ms"]}}
    for mode, name in ((1, "train_each_call"), (2, "trained_once")):
        _lib.check(ctx, L.fh_set_learned_radius(ctx, mode))
        if mode == 2:
            _mst_ms(Xd, ms, reps=1)  # the one training
        t, st = _mst_ms(Xd, ms)
        res[name] = {"mst_ms": t, "knn_ms": st["t_knn_ms"], "train_ms": st["t_learn_ms"],
                     "predict_ms": st["t_predict_ms"],
                     "ball_hit_rate": 1.0 - st["ball_misses"] / Xd.shape[0],
                     "speedup_vs_plain": plain_ms / t}
    _lib.check(ctx, L.fh_set_learned_radius(ctx, 0))
    res["note"] = ("exact by construction (ball misses and ties re-run the plain query); the plain kNN already "
                   "prunes after its first leaf, so the ball bound does not pay for itself here")
    return res


def form_high_dim():
    """The MFMA high-dimensional path: 1M x 64, 100 Gaussians, min_samples
    10."""
    import torch

    from fast_hdbscan_amd import workloads

    n5, d5 = 1_000_000, 64
    X5 = torch.from_numpy(workloads.gaussian_highdim(n5, d5, 100, seed=3)).cuda()
    t5, st5 = _mst_ms(X5, 10, reps=1)
    # algorithmic flops of the round scan: 2*d per scored pair, over the
    # 128 x 128 tile pairs the scans actually scored (pairs of tiles inside
    # one component are skipped), per launch
    # per launch of the round scan (the first round may come from the kNN
    # candidate lists, without a scan: scan_rounds counts the launches)
    nscan = max(st5["scan_rounds"], 1)
    scan_ms = st5["t_query_main_ms"] / nscan
    flops = 2.0 * d5 * 128.0 * 128.0 * st5["scored_tile_pairs"] / nscan
    achieved = flops / (scan_ms * 1e-3) / 1e12
    split = os.environ.get("FH_MFMA_SPLIT", "1") != "0"
    if split:
        # FP16 hi/lo split: three 16x16x32 FP16 products (hi.lo, lo.hi,
        # hi.hi) per product of the algorithm, against the FP16 dense peak
        roof = {"bound": "mfma", "kernel": "brute_scan_kernel<16, 1, true>", "dtype": "fp16 hi/lo split, fp32 accumulate",
                "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s", "frac": achieved / 2500.0,
                "executed_flops_per_launch": 3.0 * flops, "frac_executed": 3.0 * achieved / 2500.0,
                "fp32_peak_frac": achieved / 157.3}
    else:
        roof = {"bound": "mfma", "kernel": "brute_scan_kernel<16, 1, false>", "dtype": "fp32",
                "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s", "frac": achieved / 157.3}
    roof.update({"avg_launch_ms": scan_ms, "alg_flops_per_launch": flops})
    return {
        "workload": "1M x 64, 100 Gaussians (means N(0,4I), within N(0,I)), seed 3, min_samples 10",
        "mst_ms": t5, "points_per_s": n5 / (t5 * 1e-3), "rounds": st5["rounds"],
        "knn_ms": st5["t_knn_ms"], "knn_tree_rows": st5["knn_tree_rows"], "exact_rows": st5["exact_rows"],
        "scored_tile_pairs": st5["scored_tile_pairs"], "scan_rounds": st5["scan_rounds"], "roofline": roof,
    }


def form_scale_3d():
    """10M x 3-D mixture, min_samples 15: plain, then the learned-radius
    form (trained once on a 1% sample, then reused)."""
    import torch

    from fast_hdbscan_amd import _lib, workloads

    ctx = _lib.context(0)
    L = _lib.load()
    X3 = torch.from_numpy(workloads.scale_3d()).cuda()
    t3, st3 = _mst_ms(X3, 15, reps=2)
    out = {"workload": "10M x 3-D, 200 anisotropic Gaussians + 5% noise, seed 2, min_samples 15",
           "mst_ms": t3, "points_per_s": X3.shape[0] / (t3 * 1e-3), "rounds": st3["rounds"],
           "knn_ms": st3["t_knn_ms"], "rounds_ms": st3["t_rounds_ms"], "tree_ms": st3["t_tree_ms"]}
    _lib.check(ctx, L.fh_set_learned_radius(ctx, 2))
    _mst_ms(X3, 15, reps=1)  # the one training
    t3l, st3l = _mst_ms(X3, 15, reps=2)
    _lib.check(ctx, L.fh_set_learned_radius(ctx, 0))
    out["learned_trained_once"] = {
        "mst_ms": t3l, "knn_ms": st3l["t_knn_ms"], "predict_ms": st3l["t_predict_ms"],
        "sample_and_features_ms": st3l["t_learn_ms"],
        "ball_hit_rate": 1.0 - st3l["ball_misses"] / X3.shape[0], "speedup_vs_plain": t3 / t3l}
    return out


def form_plscan_sweep():
    """PLSCAN min_samples 2..64 on 500k x 2-D, layers for each: one at a
    time, the shared-kNN sweep of all MSTs, then pipelined (device MSTs back
    to back, layer stages in 8 host threads)."""
    import numpy as np
    import torch

    from fast_hdbscan_amd import workloads
    from fast_hdbscan_amd.hdbscan import _edges_f64, mst_device
    from fast_hdbscan_amd.layers import layers_from_mst, mst_sweep_device, plscan_sweep

    Xs = torch.from_numpy(workloads.anisotropic_mixture(500_000, 2, 50, 0.1, seed=4)).cuda()
    t0 = time.perf_counter()
    t_mst = 0.0
    n_layers = []
    for m in range(2, 65):
        a = time.perf_counter()
        src, dst, w, core, _ = mst_device(Xs, m)
        e = _edges_f64(src.cpu().numpy(), dst.cpu().numpy(), w.cpu().numpy())
        c = np.sqrt(core.cpu().numpy())
        t_mst += time.perf_counter() - a
        n_layers.append(len(layers_from_mst(e, c, Xs.shape[0], base_min_cluster_size=5)[0]))
    total = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mst_sweep_device(Xs, list(range(2, 65)))
    torch.cuda.synchronize()
    sweep_mst_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sw = plscan_sweep(Xs, range(2, 65), base_min_cluster_size=5, workers=8)
    total_p = time.perf_counter() - t0
    same = [len(sw[m][0]) for m in range(2, 65)] == n_layers
    return {"workload": "500k x 2-D, 50 anisotropic Gaussians + 10% noise, seed 4; min_samples 2..64, "
                        "base_min_cluster_size 5",
            "serial_total_s": total, "serial_mst_and_copy_s": t_mst, "serial_layer_stage_s": total - t_mst,
            "shared_knn_all_63_msts_s": sweep_mst_s, "pipelined_total_s": total_p, "pipelined_workers": 8,
            "pipelined_same_layer_counts": same, "layers_per_min_samples": n_layers}


def form_batched(one_call_workers=None):
    """Many independent small datasets through the native batch scheduler:
    datasets up to one work-group's tree in one persistent launch per
    dimension, the others on worker streams (with one_call_workers: a single
    call on the 512 datasets, for profiling)."""
    import os

    from fast_hdbscan_amd import workloads
    from fast_hdbscan_amd.batch import hdbscan_batch, mst_batch

    data = workloads.batch_datasets(512)
    pts = sum(x.shape[0] for x in data)
    res = {"workload": "512 datasets, n log-uniform 500..50000, d in {2,3,5,8}, 3-20 Gaussians + 5% noise, "
                       "seeds 1000..1511; min_samples 5, min_cluster_size 20", "points": pts}

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    if one_call_workers:
        w = int(one_call_workers)
        el = timed(lambda: hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=w))
        res[f"workers_{w}"] = {"s": el, "datasets_per_s": len(data) / el, "points_per_s": pts / el}
        return res
    hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=8)  # contexts and buffers sized
    for w in (1, 8, 16):
        el = timed(lambda: hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=w))
        res[f"workers_{w}"] = {"s": el, "datasets_per_s": len(data) / el, "points_per_s": pts / el}
    # the same with every dataset on the worker streams (no persistent launch)
    os.environ["FH_BATCH_PERSISTENT"] = "0"
    try:
        hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=8)
        el = timed(lambda: hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=8))
    finally:
        del os.environ["FH_BATCH_PERSISTENT"]
    res["workers_8_streams_only"] = {"s": el, "datasets_per_s": len(data) / el, "points_per_s": pts / el}
    # MSTs alone of the datasets the persistent launch takes, one call
    out = mst_batch(data, min_samples=5)
    small = [x for x, o in zip(data, out) if o[4]]
    sp = sum(x.shape[0] for x in small)
    el = min(timed(lambda: mst_batch(small, min_samples=5)) for _ in range(3))
    os.environ["FH_BATCH_PERSISTENT"] = "0"
    try:
        el1 = timed(lambda: mst_batch(small, min_samples=5))
    finally:
        del os.environ["FH_BATCH_PERSISTENT"]
    res["persistent_msts"] = {"datasets": len(small), "points": sp, "s": el, "datasets_per_s": len(small) / el,
                              "points_per_s": sp / el, "one_at_a_time_s": el1,
                              "speedup_vs_one_at_a_time": el1 / el}
    # the whole 4,096-dataset workload once, 8 and 16 workers
    data = workloads.batch_datasets(4096)
    pts = sum(x.shape[0] for x in data)
    for w in (8, 16):
        el = timed(lambda: hdbscan_batch(data, min_samples=5, min_cluster_size=20, workers=w))
        res[f"all_4096_workers_{w}"] = {"points": pts, "s": el, "datasets_per_s": len(data) / el,
                                        "points_per_s": pts / el}
    return res


def form_serving(Xd, ms, mcs=100, requests=10):
    """A stream of requests on the bench's 2M points, as a server would run
    them: each request's host label stage overlaps the next request's device
    stage (two sets of buffers, one host thread for the label stages).  Same
    work per request as a timed step; reported beside `value`, which runs
    the requests one after the other."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import torch

    from fast_hdbscan_amd.estimator import device_stage, host_stage, hdbscan_device

    works = [{}, {}]
    for wk in works:
        host_stage(device_stage(Xd, ms, work=wk), mcs, work=wk)
    torch.cuda.synchronize()
    ex = ThreadPoolExecutor(1)
    pending, last = None, None
    t0 = time.perf_counter()
    for i in range(requests):
        wk = works[i % 2]
        staged = device_stage(Xd, ms, work=wk)
        if pending is not None:
            pending.result()
        pending = ex.submit(host_stage, staged, mcs, work=wk)
    last = pending.result()
    el = time.perf_counter() - t0
    ex.shutdown()
    ref = hdbscan_device(Xd, ms, mcs)
    same = bool(np.array_equal(last[0], ref[0]) and np.array_equal(last[1], ref[1]))
    n = Xd.shape[0]
    return {"requests": requests, "ms_per_request": 1e3 * el / requests, "points_per_s": n * requests / el,
            "labels_identical": same,
            "note": "request i's host label stage runs while request i+1's MST runs on the device"}


def side_forms(Xd, ms, names=FORMS, batch_one_call=None):
    """The other forms of the path, each run once after the timed steps
    (not part of `value`): learned search radii on the same 2M x 2-D batch,
    the MFMA high-dimensional path at 1M x 64, the 10M x 3-D mixture, a
    PLSCAN min_samples sweep on 500k x 2-D, a batch of 512 small
    independent datasets, and a stream of requests on the same 2M points
    served with their host stages overlapped."""
    out = {}
    for name in names:
        if name == "learned":
            out[name] = form_learned(Xd, ms)
        elif name == "high_dim":
            out[name] = form_high_dim()
        elif name == "scale_3d":
            out[name] = form_scale_3d()
        elif name == "plscan_sweep":
            out[name] = form_plscan_sweep()
        elif name == "batched":
            out[name] = form_batched(batch_one_call)
        elif name == "serving":
            out[name] = form_serving(Xd, ms)
    return out


if __name__ == "__main__":
    main()
