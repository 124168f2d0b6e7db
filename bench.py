#!/usr/bin/env python3
"""Voronoi cells/sec on one MI355X.

One step = one full tessellation of the workload with the points already in
HBM: bounding box, device counting sort into the uniform grid, the cell
kernel (every cell's volume, face count, face areas and sorted neighbour
list), face offsets and CSR compaction.  Prints one JSON line.

Workloads (--workload):
  uniform1m    1,000,000 points i.i.d. uniform in [0,1]^3, seed 1, grid of
               about 5 points per cell (cell_density 5), plain mode
  clustered4m  4,000,000 points, 80% in 64 Gaussian clusters (sigma 0.02),
               20% uniform, seed 25, cell_density 5
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "the-tessellator_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import tess  # noqa: E402

# MI355X published FP64 peak (vector and matrix): 78.6 TFLOP/s with an FMA
# counted as two operations.  The cell kernel issues no FMA (contraction is
# off for bit-exact parity), so every counted op is one instruction lane-op.
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "uniform1m": dict(desc="plain kernel, 1,000,000 points i.i.d. uniform in [0,1]^3, f64, box walls",
                      points=1_000_000, seed=1, density=5.0, kind="uniform"),
    "clustered4m": dict(desc="plain kernel, 4,000,000 clustered points (64 Gaussians sigma 0.02 + 20% uniform)",
                        points=4_000_000, seed=25, density=5.0, kind="clustered"),
}


def make_points(w):
    if w["kind"] == "uniform":
        return tess.uniform(w["points"], w["seed"])
    return tess.clustered(w["points"], w["seed"])


def cpu_baseline(pts, density, sample, threads):
    """Time the CPU restatement (oracle) on a random sample of the same cells."""
    from oracle import oracle as orc

    rng = np.random.default_rng(0)
    cells = np.sort(rng.choice(len(pts), min(sample, len(pts)), replace=False)).astype(np.uint32)
    r = orc.compute(pts, np.zeros(3), np.ones(3), density=density, cells=cells, threads=threads)
    return len(cells) / r["seconds"], r["seconds"], len(cells)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="uniform1m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--vcap", type=int, default=0, help="vertex capacity of the cell kernel (0: automatic)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    pts = make_points(w)
    n = len(pts)
    ctx = tess.Context(0)
    d = tess.Diagram(ctx, pts, np.zeros(3), np.ones(3), density=w["density"], vertex_capacity=args.vcap)
    for _ in range(args.warmup):
        d.compute()
    ctx.synchronize()
    cells_ms, bin_ms, compact_ms, flops = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d.compute()
        t = d.timings()
        cells_ms.append(t["cells_ms"])
        bin_ms.append(t["bin_ms"])
        compact_ms.append(t["compact_ms"])
        flops.append(t["flops"])
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / (ms_per_step / 1e3)

    kernel_s = float(np.mean(cells_ms)) / 1e3
    achieved = float(np.mean(flops)) / kernel_s / 1e12
    faces = int(d.cells()["face_count"].sum())
    # algorithmic HBM bytes of the cell kernel per launch: site position + id
    # (28 B), volume + face count + status (16 B), 12 B per staged face
    alg_bytes = 44 * n + 12 * faces

    line = {
        "metric": "Voronoi cells/sec (f64, 1 MI355X): plain vs learned-radius batched kernel",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded xoshiro256**)",
        "config": {"workload": w["desc"], "points": n, "seed": w["seed"], "cell_density": w["density"],
                   "mode": "plain", "box": [[0, 0, 0], [1, 1, 1]], "vertex_capacity": args.vcap or "auto"},
        "phases_ms": {"bin": float(np.mean(bin_ms)), "cells": float(np.mean(cells_ms)),
                      "compact": float(np.mean(compact_ms))},
        "roofline": {
            "bound": "mfma",
            "ceiling": "FP64 (vector = matrix peak) 78.6 TFLOP/s, FMA counted as 2; the geometry issues "
                       "only single-rounded add/sub/mul/div/sqrt, no MFMA",
            "kernel": "k_cells (one cell per wavefront)",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "flops_per_launch": float(np.mean(flops)),
            "kernel_ms": float(np.mean(cells_ms)),
            "alg_hbm_gbs": alg_bytes / kernel_s / 1e9,
            "traffic": None,
        },
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        rate, sec, cnt = cpu_baseline(pts, w["density"], args.cpu_sample, threads)
        line["cpu_baseline"] = {
            "value": rate, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"{cnt} random cells of the same {n}-point set (seed 0 choice), grid build included, "
                      f"{sec:.1f} s on {threads} threads of {os.cpu_count()} visible",
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
