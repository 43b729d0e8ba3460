"""Run the MST path once on a named synthetic workload and print its
statistics -- the driver used for kernel profiles:

    python -m fast_hdbscan_amd.run_workload --workload highdim --n 1000000
    rocprofv3 --kernel-trace --stats -d out -- python3 -m fast_hdbscan_amd.run_workload --workload scale2d

Workloads (fast_hdbscan_amd.workloads): scale2d (2-D mixture, seed 1),
scale3d (3-D mixture, seed 2), highdim (d = 64 Gaussians, seed 3), blobs,
batch (--n independent small datasets, default 512, through the batch
scheduler with --workers threads; min_samples 5, min_cluster_size 20).
"""
from __future__ import annotations

import argparse
import json
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="scale2d", choices=["scale2d", "scale3d", "highdim", "blobs", "batch"])
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--min-samples", type=int, default=None)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--learned", type=int, default=0, help="learned-radius mode (0 off, 1, 2)")
    args = ap.parse_args(argv)

    import torch

    from . import _lib, workloads
    from .hdbscan import mst_device

    if args.workload == "batch":
        from .batch import hdbscan_batch

        data = workloads.batch_datasets(args.n or 512)
        pts = sum(x.shape[0] for x in data)
        for r in range(args.repeat):
            t = time.perf_counter()
            hdbscan_batch(data, min_samples=args.min_samples or 5, min_cluster_size=20, workers=args.workers)
            el = time.perf_counter() - t
            print(json.dumps({"workload": "batch", "datasets": len(data), "points": pts, "workers": args.workers,
                              "wall_ms": 1e3 * el, "datasets_per_s": len(data) / el}), flush=True)
        return

    if args.workload == "scale2d":
        X, ms = workloads.scale_2d(args.n or 2_000_000, seed=1), 10
    elif args.workload == "scale3d":
        X, ms = workloads.scale_3d(args.n or 10_000_000, seed=2), 15
    elif args.workload == "highdim":
        X, ms = workloads.gaussian_highdim(args.n or 1_000_000, 64, 100, seed=3), 10
    else:
        X, ms = workloads.blobs_small(args.n or 10_000, seed=0), 10
    ms = args.min_samples or ms
    Xd = torch.from_numpy(X).cuda()
    ctx = _lib.context(0)
    _lib.check(ctx, _lib.load().fh_set_learned_radius(ctx, args.learned))
    for r in range(args.repeat):
        st = {}
        torch.cuda.synchronize()
        t = time.perf_counter()
        mst_device(Xd, ms, stats=st)
        torch.cuda.synchronize()
        st["wall_ms"] = 1e3 * (time.perf_counter() - t)
        st["workload"] = args.workload
        st["n"] = X.shape[0]
        st["d"] = X.shape[1]
        st["min_samples"] = ms
        print(json.dumps(st), flush=True)


if __name__ == "__main__":
    main()
