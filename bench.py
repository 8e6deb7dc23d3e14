"""Benchmark: SPH density+force pair-interactions/s on one MI355X.

Workload: one colony of 2,000,000 cells (radius 1.0, cutoff h = 2.0) drawn
from a seeded Gaussian spheroid (sigma = 0.35 * 146.6, z axis * 0.7), held in
HBM.  Untimed: Morton sort, radix tree, neighbour lists.  Timed: K launches of
the SPH density+force stage (density, v and forces over the neighbour lists),
mixed precision and plain FP64, each bracketed by HIP events on the library's
stream and synchronised on both sides.  `value` is the mixed form's
pair-interactions/s (P / t_sph, P = ordered in-cutoff pairs); the FP64 drop-in,
the full-step time (sort + tree + neighbour lists + SPH + CIC deposit to a 256^3
FP64 grid) and the mixed-vs-FP64 errors are reported beside it.

roofline: algorithmic bytes of one SPH launch, B = 96 N + 4 P (per cell:
x,y,z,size 32, ctype 4, list offset 8, young 4 read; density 8, force 24, v 8,
mindist 8 written; per pair one u32 index), over the mean launch time, against
MI355X's 8 TB/s.  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE /
WRITE_SIZE counters when a committed summary exists (profiles/), else null.

cpu_baseline: the CPU restatement of the reference path (pointer PR-octree,
per-cell DFS walks with OpenMP dynamic,64) on this host's cores, same
workload, the two walks timed (median of 3 after one warm-up).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0, help="cell-count scale (1.0 = the named workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads for the CPU baseline (0 = all)")
    ap.add_argument("--only", choices=["mixed", "fp64", "both"], default="both")
    ap.add_argument("--autotune", action="store_true")
    return ap.parse_args()


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "sph_mixed_c1_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(col, threads):
    from oracle import oracle as O

    if threads > 0:
        O.set_threads(threads)
    nthr = O.max_threads()
    oc = O.Colony.from_dict(col)
    t0 = O.wtime()
    tree = O.Octree(oc, col["types"])
    t_build = O.wtime() - t0
    tree.computepotential()  # warm-up
    tree.computegradient()
    walks = []
    for _ in range(3):
        t0 = O.wtime()
        tree.computepotential()
        t1 = O.wtime()
        tree.computegradient()
        t2 = O.wtime()
        walks.append((t1 - t0, t2 - t1))
    walks.sort(key=lambda w: w[0] + w[1])
    tp, tg = walks[1]
    return nthr, t_build, tp, tg


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import timothy_amd as tm
    from timothy_amd import generators as gen

    col = gen.config(1, scale=a.scale)
    n = col["n"]
    ctx = tm.Context(0)
    ctx.upload_colony(col)
    ctx.step(tm.FP64)
    P, maxnbr = ctx.pair_count()
    ref = ctx.results()  # FP64 drop-in, bit-exact with the CPU path
    if a.autotune:
        print(ctx.autotune(tm.MIXED), file=sys.stderr)

    B = 96.0 * n + 4.0 * P
    out = {}
    for name, prec in (("mixed", tm.MIXED), ("fp64", tm.FP64)):
        if a.only not in ("both", name):
            continue
        ctx.sph(prec)
        ctx.sync()
        t0 = time.perf_counter()
        ms = ctx.time_stage(tm.STAGE_SPH, prec, warmup=a.warmup, iters=a.steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        out[name] = dict(ms=ms, pairs_per_s=P / (ms * 1e-3), gbs=B / (ms * 1e-3) / 1e9, wall_s=wall)
    # mixed-vs-FP64 error (FP64 GPU == CPU path bit for bit, see tests)
    err = None
    if "mixed" in out:
        ctx.sph(tm.MIXED)
        mix = ctx.results()
        off, nbr = ctx.neighbor_lists()
        rows = np.repeat(np.arange(n), np.diff(off))
        X = np.stack([col["x"], col["y"], col["z"]], 1)
        d = np.sqrt(((X[rows] - X[nbr]) ** 2).sum(1))
        u = d / 2.0
        w = np.where(u <= 0.5, -2 * u + 3 * u * u, -(1 - u) ** 2)
        sabs = np.zeros(n)
        np.add.at(sabs, rows, np.abs(w) * 48 / (np.pi * 16.0))
        has = sabs > 0
        fe = np.abs(mix["force"] - ref["force"]).max(1)[has] / sabs[has]
        pos = ref["density"] > 0
        de = np.abs(mix["density"] - ref["density"])[pos] / ref["density"][pos]
        err = dict(force_max=float(fe.max()), force_rms=float(np.sqrt((fe ** 2).mean())),
                   density_max=float(de.max()), density_rms=float(np.sqrt((de ** 2).mean())),
                   v_max_abs_over_max=float(np.abs(mix["v"] - ref["v"]).max() / np.abs(ref["v"]).max()))
        ctx.sph(tm.FP64)
    # full step: sort + tree + neighbour lists + SPH (mixed) + CIC deposit 256^3 FP64
    grid = col["grid"]
    prod = tm.DeviceArray(ctx, int(grid["G"]) ** 3, np.float64)
    ctx.cic_deposit(grid, prod, 0)
    step_ms = ctx.time_stage(tm.STAGE_STEP, tm.MIXED, warmup=1, iters=3)
    build_ms = ctx.time_stage(tm.STAGE_BUILD, tm.MIXED, warmup=1, iters=3)
    nbr_ms = ctx.time_stage(tm.STAGE_NEIGHBORS, tm.MIXED, warmup=1, iters=3)
    cic_ms = ctx.time_stage(tm.STAGE_CIC, tm.MIXED, warmup=1, iters=3)
    full_ms = step_ms + cic_ms

    cpu = None
    if not a.no_cpu:
        nthr, t_build, tp, tg = cpu_baseline(col, a.cpu_threads)
        cpu = {"value": P / (tp + tg), "unit": "pair-interactions/s", "cores": nthr, "kind": "port",
               "sample": f"full workload ({n} cells, {P} pairs): computepotential + computegradient walks, "
                         f"median of 3 after a warm-up; serial octbuild {t_build:.2f} s not included",
               "t_potential_s": tp, "t_gradient_s": tg, "t_octbuild_s": t_build}

    mixed = out.get("mixed") or out.get("fp64")
    traffic = traffic_from_profiles()
    line = {
        "metric": "SPH density+force pair-interactions/s, mixed vs FP64 plain, 1 MI355X",
        "value": mixed["pairs_per_s"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair evaluation, fp64 predicate and per-cell sums" if "mixed" in out else "f64",
        "data": "synthetic (seeded Gaussian spheroid)",
        "config": {"workload": "single colony, 2,000,000 cells, Gaussian spheroid W=146.6, r=1, h=2, mixed SPH"
                   if a.scale == 1.0 else f"scaled {a.scale}", "cells": n, "pairs": P, "max_neighbours": maxnbr,
                   "parallelism": "1 GPU"},
        "fp64": out.get("fp64"),
        "mixed": out.get("mixed"),
        "speedup_mixed_vs_fp64": (out["fp64"]["ms"] / out["mixed"]["ms"]) if len(out) == 2 else None,
        "mixed_vs_fp64_error": err,
        "full_step_ms": {"total": full_ms, "build": build_ms, "neighbors": nbr_ms,
                         "sph_mixed": out["mixed"]["ms"] if "mixed" in out else None, "cic_deposit_256^3": cic_ms},
        "roofline": {"bound": "hbm", "achieved": B / (mixed["ms"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": B / (mixed["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": B},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (mixed["pairs_per_s"] / cpu["value"]) if cpu else None,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
