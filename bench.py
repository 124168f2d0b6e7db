#!/usr/bin/env python3
"""Voronoi cells/sec on one MI355X: plain vs learned-radius batched kernel.

One step = one full tessellation of the workload with the points already in
HBM: bounding box, device counting sort into the uniform grid, the cell
passes (every cell's volume, face count, face areas and sorted neighbour
list), face offsets and CSR compaction.  Prints one JSON line.

Workloads (--workload):
  clustered4m  (default) 4,000,000 points in [0,1]^3, 80% from 64 Gaussian
               clusters (sigma 0.02, resampled inside the box) + 20% uniform,
               seed 25, grid of about 5 points per cell.  Timed with the
               learned-radius batched path (tess_batch: graph-captured
               binning and cell passes, learned cost order and capacity,
               autotuned); the plain kernel (tess_diagram) is timed on the
               same points in the same run.  The model is trained first on
               the GPU from the plain kernel's own counters on five
               independently seeded sets of this kind (seeds 20-24).
  uniform1m    1,000,000 points i.i.d. uniform in [0,1]^3, seed 1, plain
               kernel only.
  batch512     512 independent sets, N log-uniform on [1e3, 2e5], box sides
               uniform on [0.5, 2], uniform points (seeds 3000 + i); batched
               path vs the plain kernel set by set.
  geometry1m   vertices and face rings of every cell of the 1,000,000 uniform
               points (seed 1): the geometry call after a compute (a cell
               pass that stages the geometry, offsets, compaction), results
               left in HBM; reports cells/s of the export, and beside it
               cells + geometry by one call on a diagram not yet computed
               (its compute pass stages the geometry) against compute then
               geometry.
  sites1m      cells of 1,000,000 uniform sites that are not points of the
               diagram of the 1,000,000 uniform points (get_cell_at_particle
               for every site, tess_diagram_cells_at_particles); the sites
               arrive as a host array, so cells/s is taken over the device
               span of the call (cell passes, offsets, compaction) and the
               whole call, upload included, is reported beside it.
  queries1m    radius queries (find_neighbors_in_real_radius, r = 0.01) at
               the 1,000,000 points of the clustered set (seed 25, grid of
               about 5 points per cell), positions and radii resident in
               HBM, results left in HBM; reports queries/s (its own metric).
  fcc16m       16 perturbed FCC lattices of 63^3 cells (1,000,188 points
               each; cubic lattice constant a = 1, nearest-neighbour distance
               a / sqrt 2; Gaussian jitter sigma 1e-3; seeds 4000-4015) in
               one batch with the learned order (model trained on two other
               lattices, seeds 4100-4101); the plain batch is timed beside
               it and the learned path's fallback rate is reported.

The roofline of the cell passes is FP64 VALU issue: the geometry is single
rounded add/sub/mul/div/sqrt with FMA contraction off, so the no-FMA
ceiling (39.3 Tops/s, half the 78.6 TFLOP/s FMA-counted peak) is reported
beside the published peak.
"""
import argparse
import json
import os
import sys
import tempfile
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
FP64_NOFMA_TOPS = 39.3
HBM_PEAK_GBS = 8000.0
METRIC = "Voronoi cells/sec (f64, 1 MI355X): plain vs learned-radius batched kernel"

WORKLOADS = {
    "clustered4m": dict(desc="4,000,000 clustered points (64 Gaussians sigma 0.02 + 20% uniform) in [0,1]^3, f64, "
                             "box walls: batched path with the autotuned order (learned or list), plain kernel "
                             "timed beside",
                        points=4_000_000, seed=25, density=5.0, kind="clustered", train_seeds=(20, 21, 22, 23, 24),
                        cpu_frac=0.01),
    "uniform1m": dict(desc="plain kernel, 1,000,000 points i.i.d. uniform in [0,1]^3, f64, box walls",
                      points=1_000_000, seed=1, density=5.0, kind="uniform"),
    "batch512": dict(desc="batched kernel, 512 uniform sets, N log-uniform [1e3, 2e5], box sides U[0.5, 2]",
                     sets=512, seed=3000, density=5.0, kind="batch"),
    "geometry1m": dict(desc="geometry export (vertices + face rings, CSR) of every cell, 1,000,000 points i.i.d. "
                            "uniform in [0,1]^3, f64, box walls", points=1_000_000, seed=1, density=5.0,
                       kind="geometry"),
    "sites1m": dict(desc="cells of 1,000,000 sites that are not points of the diagram (get_cell_at_particle), "
                         "sites i.i.d. uniform (seed 2) against 1,000,000 points i.i.d. uniform (seed 1) in "
                         "[0,1]^3, f64, box walls", points=1_000_000, seed=1, sites=1_000_000, site_seed=2,
                    density=5.0, kind="sites"),
    "queries1m": dict(desc="real-radius neighbour queries (r = 0.01) at the 1,000,000 points of the clustered set "
                           "(64 Gaussians sigma 0.02 + 20% uniform, seed 25), f64",
                      points=1_000_000, seed=25, density=5.0, kind="queries", radius=0.01),
    "fcc16m": dict(desc="learned + batched kernel, 16 perturbed FCC lattices of 63^3 cells (1,000,188 points each, "
                        "cubic constant a = 1, jitter sigma 1e-3) in [0,63]^3, f64, box walls",
                   lattices=16, ncell=63, seed=4000, train_seeds=(4100, 4101), density=5.0, kind="fcc",
                   cpu_frac=0.002),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_points(w):
    if w["kind"] == "uniform":
        return tess.uniform(w["points"], w["seed"])
    return tess.clustered(w["points"], w["seed"])


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def roofline_fp64(achieved, kernel, flops, kernel_ms):
    """FP64 VALU roofline fields of a cell pass (achieved in Tops/s)."""
    return {
        "bound": "valu",
        "ceiling": "FP64 VALU: published peak 78.6 TFLOP/s counts an FMA as 2; the geometry issues only "
                   "single-rounded add/sub/mul/div/sqrt (contraction off, no MFMA), so its own ceiling is "
                   "39.3 Tops/s (frac_no_fma)",
        "kernel": kernel,
        "achieved": achieved,
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / FP64_PEAK_TFLOPS,
        "frac_no_fma": achieved / FP64_NOFMA_TOPS,
        "flops_per_launch": flops,
        "kernel_ms": kernel_ms,
        "traffic": None,
    }


def cpu_baseline(pts, lo, hi, density, frac, threads):
    """The CPU restatement (oracle, C++ -O3, OpenMP dynamic over cells) on a
    seeded random sample of the same cells, grid build included: best of 3
    on `threads` host threads and best of 3 on one."""
    from oracle import oracle as orc

    rng = np.random.default_rng(0)
    k = max(1, int(round(frac * len(pts))))
    cells = np.sort(rng.choice(len(pts), k, replace=False)).astype(np.uint32)
    best = {}
    for t in (threads, 1):
        secs = []
        for _ in range(3):
            r = orc.compute(pts, lo, hi, density=density, cells=cells, threads=t)
            secs.append(r["seconds"])
        best[t] = min(secs)
    return {
        "value": k / best[threads], "unit": "cells/s", "cores": threads, "kind": "port",
        "single_core": k / best[1],
        "sample": f"{k} random cells ({100 * frac:.0f}% of the {len(pts)}-point set, seed 0 choice), grid build "
                  f"included, best of 3: {best[threads]:.2f} s on {threads} threads, {best[1]:.2f} s on 1",
        "nproc": os.cpu_count(), "cpu_model": cpu_model(),
    }


def fixed_config(text):
    """'64,0,4,1[,1]' -> first_capacity, xcd_chunk, redo_chunk, learned[, eager]."""
    v = [int(x) for x in text.split(",")]
    return dict(zip(("first_capacity", "xcd_chunk", "redo_chunk", "learned", "eager"), v))


def timed(fn, steps, warmup, ctx):
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    t0 = time.perf_counter()
    tims = []
    for _ in range(steps):
        tims.append(fn())
    ctx.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, tims


def mean(tims, k):
    return float(np.mean([t[k] for t in tims]))


def train_model(ctx, w, path):
    """Train the radius model on the plain kernel's counters of the training sets."""
    m = tess.Model(ctx, seed=7)
    t0 = time.perf_counter()
    for s in w["train_seeds"]:
        pts = tess.clustered(w["points"], s)
        d = tess.Diagram(ctx, pts, np.zeros(3), np.ones(3), density=w["density"], collect_stats=True)
        d.compute()
        m.add_samples(d)
        d.close()
        log(f"training samples from seed {s}: {m.sample_count()}")
    t1 = time.perf_counter()
    loss = m.train(epochs=3, batch=8192, lr=2e-3, seed=1)
    t2 = time.perf_counter()
    m.save(path)
    log(f"model trained: loss per epoch {list(np.round(loss, 4))}")
    return m, dict(samples=m.sample_count(), sample_s=t1 - t0, train_s=t2 - t1,
                   loss=[float(x) for x in loss])


def bench_clustered(ctx, w, args):
    pts = make_points(w)
    n = len(pts)
    lo, hi = np.zeros(3), np.ones(3)
    if args.plain_only:
        # the plain kernel alone (profiling runs): no model, no batch
        d = tess.Diagram(ctx, pts, lo, hi, density=w["density"], xcd_chunk=args.xcd_chunk)

        def plain_only_step():
            d.compute()
            return d.timings()

        ms, tims = timed(plain_only_step, args.steps, args.warmup, ctx)
        return pts, lo, hi, n, ms, tims, d.cells(), {}, {"mode": "plain"}
    path = os.path.join(tempfile.gettempdir(), "tess_clustered4m_model.bin")
    if args.model:
        m = tess.Model(ctx)
        m.load(args.model)
        train = {"loaded": args.model}
    else:
        m, train = train_model(ctx, w, path)
    # plain kernel on the same points
    d = tess.Diagram(ctx, pts, lo, hi, density=w["density"], xcd_chunk=args.xcd_chunk)

    def plain_step():
        d.compute()
        return d.timings()

    plain_ms, plain_t = timed(plain_step, max(2, args.steps // 2), 1, ctx)
    plain_cells = d.cells()
    log(f"plain: {plain_ms:.1f} ms/step")
    # learned-radius batched path
    b = tess.Batch(ctx, [(pts, lo, hi)], density=w["density"])
    cache = os.path.join(tempfile.gettempdir(), "tess_tune_cache.txt")
    if args.batch_config:
        b.set_config(**fixed_config(args.batch_config))
        cfg = b.config()
    else:
        cfg = b.tune(m, cache_path=cache)
    log(f"config: {cfg}")
    sweep = []
    for k in ((dict(first_capacity=64, xcd_chunk=0, learned=0), dict(first_capacity=64, xcd_chunk=256, learned=0),
               dict(first_capacity=64, xcd_chunk=4096, learned=0), dict(first_capacity=128, xcd_chunk=0, learned=0),
               dict(first_capacity=64, xcd_chunk=0, learned=1))
              if args.sweep else ()):
        b.set_config(**k)
        b.compute(m)
        t = []
        for _ in range(2):
            b.compute(m)
            t.append(b.timings())
        sweep.append(dict(k, total_ms=mean(t, "total_ms"), cells_ms=mean(t, "cells_ms"), plan_ms=mean(t, "plan_ms"),
                          reruns=int(t[-1]["capacity_reruns"])))
        log(f"sweep {sweep[-1]}")
    b.set_config(**cfg)

    def step():
        b.compute(m)
        return b.timings()

    ms, tims = timed(step, args.steps, args.warmup, ctx)
    got = b.cells(0)
    identical = all(np.array_equal(got[k], plain_cells[k]) for k in plain_cells)
    log(f"learned batched: {ms:.1f} ms/step, bitwise equal to plain: {identical}")
    extra = {
        "plain": {"value": n / (plain_ms / 1e3), "ms_per_step": plain_ms,
                  "phases_ms": {k: mean(plain_t, k) for k in ("bin_ms", "cells_ms", "compact_ms")},
                  "capacity_reruns": int(plain_t[-1]["capacity_reruns"])},
        "speedup_vs_plain": plain_ms / ms,
        "bitwise_equal_to_plain": bool(identical),
        "fallback_cells": int(tims[-1]["fallback_cells"]),
        "fallback_rate": tims[-1]["fallback_cells"] / n,
        "tuned_config": cfg,
        "config_sweep": sweep,
        "model": train,
    }
    # host memory to host memory on the path just timed (untimed for the value)
    extra["end_to_end"] = end_to_end(ctx, pts, lo, hi, w["density"], batch_config=cfg, model=m)
    mode = "learned-batched" if cfg["learned"] else "batched (the tuner kept the plain list order)"
    if not cfg["learned"]:
        extra["fallback_cells"] = extra["fallback_rate"] = None
    return pts, lo, hi, n, ms, tims, got, extra, {"mode": mode}


def bench_uniform(ctx, w, args):
    pts = make_points(w)
    n = len(pts)
    lo, hi = np.zeros(3), np.ones(3)
    d = tess.Diagram(ctx, pts, lo, hi, density=w["density"], vertex_capacity=args.vcap, xcd_chunk=args.xcd_chunk)

    def step():
        d.compute()
        return d.timings()

    ms, tims = timed(step, args.steps, args.warmup, ctx)
    return pts, lo, hi, n, ms, tims, d.cells(), {}, {"mode": "plain", "vertex_capacity": args.vcap or "auto"}


def bench_batch(ctx, w, args):
    sets = []
    for i in range(w["sets"]):
        p, side = tess.batch_set(w["seed"] + i)
        sets.append((p, np.zeros(3), side))
    n = sum(len(p) for p, _, _ in sets)
    log(f"batch: {len(sets)} sets, {n} points")
    b = tess.Batch(ctx, sets, density=w["density"])
    cache = os.path.join(tempfile.gettempdir(), "tess_tune_cache.txt")
    if args.batch_config:
        b.set_config(**fixed_config(args.batch_config))
        cfg = b.config()
    else:
        cfg = b.tune(None, cache_path=cache)

    def step():
        b.compute()
        return b.timings()

    ms, tims = timed(step, args.steps, args.warmup, ctx)
    # plain kernel set by set on the same sets (diagrams created beforehand;
    # one pass of compute calls, wall clock)
    ds = [tess.Diagram(ctx, p, lo, hi, density=w["density"]) for p, lo, hi in sets]
    ctx.synchronize()
    t0 = time.perf_counter()
    for d in ds:
        d.compute()
    ctx.synchronize()
    plain_s = time.perf_counter() - t0
    for d in ds:
        d.close()
    extra = {"plain_set_by_set": {"value": n / plain_s, "seconds": plain_s}, "speedup_vs_plain": plain_s * 1e3 / ms,
             "tuned_config": cfg}
    lo, hi = sets[0][1], sets[0][2]
    return sets[0][0], lo, hi, n, ms, tims, None, extra, {"mode": "batched", "sets": len(sets)}


def end_to_end(ctx, pts, lo, hi, density, reps=3, batch_config=None, model=None):
    """The timed path from host memory to host memory: upload (create), the
    compute, and the download of every cell's results (PCIe both ways); the
    batched path with the given configuration when one is given, else the
    plain one."""
    best = None
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        if batch_config is not None:
            d = tess.Batch(ctx, [(pts, lo, hi)], density=density)
            d.set_config(**batch_config)
            d.compute(model)
            d.cells(0)
        else:
            d = tess.Diagram(ctx, pts, lo, hi, density=density)
            d.compute()
            d.cells()
        t1 = time.perf_counter()
        d.close()
        best = t1 - t0 if best is None else min(best, t1 - t0)
    path = "batched path with the timed configuration" if batch_config is not None else "plain path"
    return {"ms": best * 1e3, "cells_per_s": len(pts) / best,
            "includes": f"host points -> HBM, compute ({path}), every cell's results HBM -> host (best of 3)"}


def bench_fcc(ctx, w, args):
    """16 jittered FCC lattices in one batch, learned order (model trained on
    two other lattices), the plain batch timed beside it."""
    box = np.full(3, float(w["ncell"]))
    lat = [tess.fcc(w["seed"] + i, w["ncell"]) for i in range(w["lattices"])]
    n = sum(len(p) for p in lat)
    m = tess.Model(ctx, seed=3)
    t0 = time.perf_counter()
    for s in w["train_seeds"]:
        d = tess.Diagram(ctx, tess.fcc(s, w["ncell"]), np.zeros(3), box, density=w["density"], collect_stats=True)
        d.compute()
        m.add_samples(d)
        d.close()
    loss = m.train(epochs=2, batch=8192, lr=2e-3, seed=1)
    train = dict(samples=m.sample_count(), seconds=time.perf_counter() - t0, loss=[float(x) for x in loss])
    sets = [(p, np.zeros(3), box) for p in lat]
    plain = tess.Batch(ctx, sets, density=w["density"])

    def plain_step():
        plain.compute()
        return plain.timings()

    plain_ms, plain_t = timed(plain_step, max(2, args.steps // 2), 1, ctx)
    b = tess.Batch(ctx, sets, density=w["density"])
    b.set_config(**(fixed_config(args.batch_config) if args.batch_config else dict(learned=1)))

    def step():
        b.compute(m)
        return b.timings()

    ms, tims = timed(step, args.steps, args.warmup, ctx)
    identical = all(np.array_equal(b.cells(s)[k], v) for s in range(len(sets)) for k, v in plain.cells(s).items())
    extra = {
        "plain": {"value": n / (plain_ms / 1e3), "ms_per_step": plain_ms,
                  "phases_ms": {k: mean(plain_t, k) for k in ("bin_ms", "cells_ms", "compact_ms")}},
        "speedup_vs_plain": plain_ms / ms,
        "bitwise_equal_to_plain": bool(identical),
        "fallback_cells": int(tims[-1]["fallback_cells"]),
        "fallback_rate": tims[-1]["fallback_cells"] / n,
        "capacity_reruns": int(tims[-1]["capacity_reruns"]),
        "model": train,
    }
    plain.close()
    cells = {"face_count": np.concatenate([b.cells(s)["face_count"] for s in range(len(sets))])}
    b.close()
    return lat[0], np.zeros(3), box, n, ms, tims, cells, extra, {"mode": "learned-batched", "lattices": len(sets)}


def bench_geometry(ctx, w, args):
    """tess_diagram_compute_geometry after a compute: the staging cell pass,
    the two scans, sizing and the compaction; results left in HBM."""
    import ctypes as C

    pts = tess.uniform(w["points"], w["seed"])
    n = len(pts)
    d = tess.Diagram(ctx, pts, np.zeros(3), np.ones(3), density=w["density"])
    d.compute()
    flops = d.timings()["flops"]
    lib = d.lib
    ms2 = np.zeros(2)

    def step():
        d.ctx.check(lib.tess_diagram_compute_geometry(d.h))
        d.ctx.check(lib.tess_diagram_geometry_timings(d.h, ms2.ctypes.data_as(C.c_void_p)))
        return dict(stage_ms=float(ms2[0]), compact_ms=float(ms2[1]))

    # each step: a fresh compute (untimed), then the timed geometry call
    ms_list, tims = [], []
    for i in range(args.warmup + args.steps):
        d.compute()
        ctx.synchronize()
        t0 = time.perf_counter()
        t = step()
        ctx.synchronize()
        if i >= args.warmup:
            ms_list.append((time.perf_counter() - t0) * 1e3)
            tims.append(t)
    ms = float(np.mean(ms_list))
    nv, nr = C.c_uint64(), C.c_uint64()
    d.ctx.check(lib.tess_diagram_geometry_sizes(d.h, C.byref(nv), C.byref(nr)))
    # cells and geometry wanted together: compute then geometry (two cell
    # passes) against one geometry call on a diagram not yet computed, whose
    # compute pass stages the geometry itself (fresh diagrams built untimed)
    sep, one = [], []
    for i in range(1 + min(args.steps, 3)):
        ctx.synchronize()
        t0 = time.perf_counter()
        d.compute()
        step()
        ctx.synchronize()
        t1 = time.perf_counter()
        f = tess.Diagram(ctx, pts, np.zeros(3), np.ones(3), density=w["density"])
        ctx.synchronize()
        t2 = time.perf_counter()
        f.ctx.check(lib.tess_diagram_compute_geometry(f.h))
        ctx.synchronize()
        t3 = time.perf_counter()
        f.close()
        if i:
            sep.append((t1 - t0) * 1e3)
            one.append((t3 - t2) * 1e3)
    stage_ms = mean(tims, "stage_ms")
    achieved = flops / (stage_ms / 1e3) / 1e12
    line = {
        "metric": "Voronoi cells with vertices and face rings/sec (f64, 1 MI355X)",
        "value": n / (ms / 1e3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded xoshiro256**)",
        "config": {"workload": w["desc"], "points": n, "cell_density": w["density"]},
        "phases_ms": {"stage": stage_ms, "compact": mean(tims, "compact_ms")},
        "cells_and_geometry_ms": {"compute_then_geometry": float(np.mean(sep)),
                                  "one_call_staged_by_compute": float(np.mean(one))},
        "sizes": {"vertices": nv.value, "ring_entries": nr.value, "faces": int(d.cells()["face_count"].sum())},
        "roofline": roofline_fp64(achieved, "staging pass k_cells<1, 2, 1, 1> (+ its capacity rerun; it "
                                            "recomputes every cell), HIP events on the context stream", flops,
                                  stage_ms),
    }
    if not args.no_cpu:
        from oracle import oracle as orc

        threads = min(16, os.cpu_count() or 1)
        k = max(1, int(args.cpu_frac * n))
        cells = np.sort(np.random.default_rng(0).choice(n, k, replace=False)).astype(np.uint32)
        t0 = time.perf_counter()
        orc.compute_geometry(pts, np.zeros(3), np.ones(3), density=w["density"], cells=cells, threads=threads)
        sec = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": k / sec, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"{k} random cells of the {n} (seed 0 choice) with their rings, cells and grid build "
                      f"included, {sec:.1f} s on {threads} threads of {os.cpu_count()} visible",
        }
    d.close()
    print(json.dumps(line), flush=True)


def bench_sites(ctx, w, args):
    """tess_diagram_cells_at_particles for 1M sites: one step is one call
    (upload, first cell pass, capacity rerun, offsets, compaction)."""
    import ctypes as C

    pts = tess.uniform(w["points"], w["seed"])
    sites = np.ascontiguousarray(tess.uniform(w["sites"], w["site_seed"]))
    n, m = len(pts), len(sites)
    lo, hi = np.zeros(3), np.ones(3)
    # exact FP64 operation count of the cells (untimed, counters on)
    ds = tess.Diagram(ctx, pts, lo, hi, density=w["density"], collect_stats=True)
    ref = ds.cells_at_particles(sites, stats=True)
    flops = float(ref["stats"]["flops"].astype(np.float64).sum())
    faces = int(ref["face_count"].sum())
    ds.close()
    d = tess.Diagram(ctx, pts, lo, hi, density=w["density"])
    d.compute()
    lib = d.lib
    ms2, rr = np.zeros(2), C.c_uint32()

    def step():
        h = C.c_void_p()
        d.ctx.check(lib.tess_diagram_cells_at_particles(d.h, sites.ctypes.data_as(C.c_void_p), m, float("nan"), -1,
                                                        C.byref(h)))
        try:
            d.ctx.check(lib.tess_site_cells_timings(h, ms2.ctypes.data_as(C.c_void_p), C.byref(rr)))
        finally:
            lib.tess_site_cells_destroy(h)
        return dict(cells_ms=float(ms2[0]), compact_ms=float(ms2[1]), reruns=rr.value)

    wall_ms, tims = timed(step, args.steps, args.warmup, ctx)
    cells_ms = mean(tims, "cells_ms")
    span_ms = cells_ms + mean(tims, "compact_ms")
    achieved = flops / (cells_ms / 1e3) / 1e12
    line = {
        "metric": "Voronoi cells of outside sites/sec (get_cell_at_particle, f64, 1 MI355X)",
        "value": m / (span_ms / 1e3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": span_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded xoshiro256**)",
        "config": {"workload": w["desc"], "points": n, "sites": m, "cell_density": w["density"]},
        "phases_ms": {"cells": cells_ms, "offsets_compact": mean(tims, "compact_ms")},
        "whole_call": {"ms": wall_ms, "cells_per_s": m / (wall_ms / 1e3),
                       "note": "host sites uploaded over PCIe and result buffers allocated inside the call"},
        "capacity_reruns": tims[-1]["reruns"],
        "faces": faces,
        "roofline": roofline_fp64(achieved, "cell passes k_cells<1, 2, 1, 0> + capacity rerun, HIP events on "
                                            "the context stream", flops, cells_ms),
    }
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        # HBM bytes of the first cell pass over the sites, from the rocprofv3
        # FETCH_SIZE / WRITE_SIZE runs of this workload, copy-kernel scaled
        rec = json.load(open(tpath)).get(args.workload)
        if rec:
            line["roofline"]["traffic"] = rec["bytes_per_launch"]
            line["roofline"]["traffic_kernel"] = rec["kernel"]
    # algorithmic HBM bytes of the cell passes: site position (24 B), volume +
    # face count + status (16 B), 12 B per staged face
    line["roofline"]["alg_hbm_gbs"] = (40 * m + 12 * faces) / (cells_ms / 1e3) / 1e9
    if not args.no_cpu:
        from oracle import oracle as orc

        threads = min(16, os.cpu_count() or 1)
        k = max(1, int(args.cpu_frac * m))
        sel = np.random.default_rng(0).choice(m, k, replace=False)
        t0 = time.perf_counter()
        orc.compute_at_points(pts, sites[sel], lo, hi, density=w["density"], threads=threads)
        sec = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": k / sec, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"{k} random sites of the {m} (seed 0 choice), grid build included, {sec:.1f} s on "
                      f"{threads} threads of {os.cpu_count()} visible",
        }
    d.close()
    print(json.dumps(line), flush=True)


def bench_queries(ctx, w, args):
    """tess_diagram_query_device over HBM-resident query positions; one step
    is one call (count pass, scan, sizing, fill pass), results left in HBM."""
    pts = tess.clustered(w["points"], w["seed"])
    n = len(pts)
    d = tess.Diagram(ctx, pts, np.zeros(3), np.ones(3), density=w["density"])
    r = w["radius"]
    qt = tess.DeviceArray(ctx, pts)
    rt = tess.DeviceArray(ctx, np.array([r]))
    # exact work counts of the walk (untimed): in-range grid cells, candidate
    # points of those cells, kept points
    sizes = {}
    for kind in (tess.QUERY_CELLS, tess.QUERY_CELL_RADIUS, tess.QUERY_REAL_RADIUS):
        res = d.query_device(kind, qt.ptr, n, rt.ptr, 1)
        sizes[kind] = int(res.total)
        res.close()
    cells, cands, kept = sizes[tess.QUERY_CELLS], sizes[tess.QUERY_CELL_RADIUS], sizes[tess.QUERY_REAL_RADIUS]

    def step():
        res = d.query_device(tess.QUERY_REAL_RADIUS, qt.ptr, n, rt.ptr, 1)
        t = res.timings()
        res.close()
        return t

    ms, tims = timed(step, args.steps, args.warmup, ctx)
    fill_ms = mean(tims, "fill_ms")
    count_ms = mean(tims, "count_ms")
    # algorithmic (compulsory) bytes of the fill pass, every datum once: per
    # query its position (24 B), its slot offset (8 B) and its kept count
    # (4 B); per point its sorted index and position (28 B); the grid's
    # delimiters (4 B per cell); per kept point its id (4 B)
    cpd = d.grid()[2]
    fill_bytes = 36 * n + 28 * n + 4 * (cpd ** 3 + 1) + 4 * kept
    achieved = fill_bytes / (fill_ms / 1e3) / 1e9
    # the same reads counted per query (every in-range grid cell's two
    # delimiters, every candidate's index and position): mostly served by the
    # caches, since neighbouring queries share grid cells
    per_query_bytes = 40 * n + 8 * cells + 28 * cands + 4 * kept
    line = {
        "metric": "radius queries/sec (find_neighbors_in_real_radius, f64, 1 MI355X)",
        "value": n / (ms / 1e3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded xoshiro256**)",
        "config": {"workload": w["desc"], "points": n, "queries": n, "radius": r, "cell_density": w["density"]},
        "phases_ms": {"count": count_ms, "fill": fill_ms, "call": mean(tims, "total_ms")},
        "work": {"grid_cells": cells, "candidates": cands, "kept": kept},
        "roofline": {
            "bound": "hbm",
            "kernel": "fill pass k_query_box<2, true>, HIP events on the context stream",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "bytes_per_launch": fill_bytes,
            "kernel_ms": fill_ms,
            "traffic": None,
            "per_query_bytes": per_query_bytes,
            "per_query_gbs": per_query_bytes / (fill_ms / 1e3) / 1e9,
        },
    }
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        rec = json.load(open(tpath)).get(args.workload)
        if rec:
            # measured HBM bytes: above the compulsory ones by the kept ids'
            # partial-line writes and the re-reads of shared grid cells
            line["roofline"]["traffic"] = rec["bytes_per_launch"]
            line["roofline"]["traffic_kernel"] = rec["kernel"]
            line["roofline"]["hbm_gbs"] = rec["bytes_per_launch"] / (fill_ms / 1e3) / 1e9
            line["roofline"]["hbm_frac"] = line["roofline"]["hbm_gbs"] / HBM_PEAK_GBS
    if not args.no_cpu:
        from oracle import oracle as orc

        m = max(1, int(args.cpu_frac * n))
        sel = np.random.default_rng(0).choice(n, m, replace=False)
        t0 = time.perf_counter()
        off, _ = orc.queries(tess.QUERY_REAL_RADIUS, pts, pts[sel], r, density=w["density"])
        sec = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": m / sec, "unit": "queries/s", "cores": 1, "kind": "port",
            "sample": f"{m} random queries of the same set (seed 0 choice), grid build included, {sec:.1f} s "
                      f"on 1 thread",
        }
    d.close()
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="clustered4m", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-frac", type=float, default=None,
                    help="fraction of the cells the CPU baseline times (seeded random sample; default per "
                         "workload, 1-5%%)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--vcap", type=int, default=0, help="vertex capacity of the plain kernel (0: automatic)")
    ap.add_argument("--xcd-chunk", type=int, default=0,
                    help="plain kernel: cells each XCD takes in a row (0: automatic)")
    ap.add_argument("--model", default="", help="clustered4m: load this trained model instead of training one")
    ap.add_argument("--batch-config", default="", help="skip the autotuner: first_capacity,xcd_chunk,redo_chunk,"
                                                       "learned (e.g. 64,0,4,0)")
    ap.add_argument("--sweep", action="store_true", help="clustered4m: also time a fixed set of batch configurations")
    ap.add_argument("--plain-only", action="store_true",
                    help="clustered4m: time the plain kernel only (no model, no batch), e.g. under a profiler")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path; --gpus must be 1.", file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    if args.cpu_frac is None:
        args.cpu_frac = w.get("cpu_frac", 0.05)
    ctx = tess.Context(0)
    if w["kind"] == "queries":
        bench_queries(ctx, w, args)
        return
    if w["kind"] == "geometry":
        bench_geometry(ctx, w, args)
        return
    if w["kind"] == "sites":
        bench_sites(ctx, w, args)
        return
    fn = {"clustered": bench_clustered, "uniform": bench_uniform, "batch": bench_batch, "fcc": bench_fcc}[w["kind"]]
    pts, lo, hi, n, ms, tims, cells, extra, cfg_extra = fn(ctx, w, args)
    value = n / (ms / 1e3)
    kernel_ms = mean(tims, "cells_ms")
    flops = mean(tims, "flops")
    achieved = flops / (kernel_ms / 1e3) / 1e12
    faces = int(cells["face_count"].sum()) if cells is not None else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded xoshiro256**)",
        "config": dict({"workload": w["desc"], "points": n, "cell_density": w["density"]}, **cfg_extra),
        "phases_ms": {"bin": mean(tims, "bin_ms"), "plan": mean(tims, "plan_ms"), "cells": kernel_ms,
                      "compact": mean(tims, "compact_ms")},
        "roofline": roofline_fp64(achieved, "cell passes (first pass + capacity rerun), HIP events on the "
                                            "context stream", flops, kernel_ms),
    }
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        # HBM bytes per launch of the dominant cell kernel, from the rocprofv3
        # FETCH_SIZE / WRITE_SIZE runs of this workload, copy-kernel scaled
        rec = json.load(open(tpath)).get(args.workload)
        if rec:
            traffic = rec["bytes_per_launch"]
            byx = rec.get("by_xcd_chunk")
            tuned = extra.get("tuned_config")
            if byx and tuned:
                # the figure measured under the XCD mapping this run was timed
                # with (none if that mapping was not measured)
                traffic = byx.get(str(tuned["xcd_chunk"]))
                line["roofline"]["traffic_config"] = f"xcd_chunk {tuned['xcd_chunk']}"
            line["roofline"]["traffic"] = traffic
            line["roofline"]["traffic_kernel"] = rec["kernel"]
    if faces is not None:
        # algorithmic HBM bytes of the cell passes: site position + id (28 B),
        # volume + face count + status (16 B), 12 B per staged face
        line["roofline"]["alg_hbm_gbs"] = (44 * n + 12 * faces) / (kernel_ms / 1e3) / 1e9
    line.update(extra)
    if w["kind"] in ("clustered", "uniform") and not args.plain_only:
        if "end_to_end" not in line:
            line["end_to_end"] = end_to_end(ctx, pts, lo, hi, w["density"])
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(pts, lo, hi, w["density"], args.cpu_frac, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
