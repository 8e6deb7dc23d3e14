"""Benchmark: SPH density+force pair-interactions/s on one MI355X.

Default workload (the metric's): one colony of 2,000,000 cells (radius 1.0,
cutoff h = 2.0) drawn from a seeded Gaussian spheroid (sigma = 0.35 * 146.6,
z axis * 0.7), held in HBM.  Untimed: Morton sort, radix tree, neighbour
lists (with v and mindist, which the neighbour walk produces), LDS tiles,
and first tm_autotune (mixed; --no-autotune keeps the default shapes).
Timed: K launches of the SPH density+force stage (density and forces over
the neighbour lists), mixed precision and plain FP64, each bracketed by HIP
events on the library's stream and synchronised on both sides.  `value` is the mixed form's pair-interactions/s (P / t_sph, P =
ordered in-cutoff pairs); the FP64 drop-in, the full-step time (sort + tree +
neighbour lists + SPH + CIC deposit to a 256^3 FP64 grid), the resident
mechanics loop (a captured graph whose every replay sorts, builds, lists,
sums and moves the cells) and the mixed-vs-FP64 errors are reported beside
it.

--workload batched: 256 independent colonies of 1,000..200,000 cells
(log-uniform, ~9.6M cells), one launch per stage over all colonies, the step
captured in a HIP graph and replayed; per-colony mixed-vs-FP64 errors.
--workload large: 50,000,000 cells, 5% of radius 1.5; mixed SPH, CIC deposit
of cell volume onto 512^3 FP64, FP32 field gradient and interpolation of a
seeded smooth field on 512^3.

roofline: algorithmic bytes of one SPH launch, B = 76 N + 4 P (per cell:
x,y,z,size 32, ctype 4, list offset 8 read; density 8, force 24 written; per
pair one u32 index; v and mindist are produced by the neighbour walk), over
the mean launch time, against MI355X's 8 TB/s.  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE /
WRITE_SIZE counters when a committed summary exists (profiles/), else null.

cpu_baseline: the CPU restatement of the reference path (pointer PR-octree,
per-cell DFS walks with OpenMP dynamic,64) on this host's cores, the two
walks timed (median of 3 after one warm-up); for the default workload on the
full colony, for the others on a bounded sample named in the line.
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
METRIC = "SPH density+force pair-interactions/s, mixed vs FP64 plain, 1 MI355X"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", choices=["single", "batched", "large"], default="single")
    ap.add_argument("--scale", type=float, default=1.0, help="cell-count scale (1.0 = the named workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads for the CPU baseline (0 = all)")
    ap.add_argument("--only", choices=["mixed", "fp64", "both"], default="both")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default shapes instead of running tm_autotune (mixed) first")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def autotune(tm, ctx, a):
    """tm_autotune (mixed) before timing: the shapes it keeps and its trial count"""
    if not a.autotune:
        return None
    t0 = time.perf_counter()
    r = ctx.autotune(tm.MIXED)
    log(r["text"].splitlines()[-1])
    # the tile width x unroll trials are kept: the tuner minimises tile build
    # + pair kernel (both run once per step), which can cost the pair kernel
    # alone some time where a wider tile builds faster
    ts = [{k: t[k] for k in ("tile_width", "unroll", "tiles_ms", "sph_ms")} for t in r["trials"]
          if t["stage"] == "tiles_sph"]
    return {"chosen": r["chosen"], "trials": len(r["trials"]), "tiles_sph_trials": ts,
            "seconds": round(time.perf_counter() - t0, 3)}


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "k_sph_tile_mixed_gaussian2M_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_walks(col, threads, reps=3, deposit=False):
    """oracle octbuild + the two per-cell walks (median of `reps` after a
    warm-up); with deposit, also the serial CIC deposit onto the colony's grid
    (median of 3), the reference's src/patches.c:156-234"""
    from oracle import oracle as O

    prev = O.max_threads()
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
    for _ in range(reps):
        t0 = O.wtime()
        tree.computepotential()
        t1 = O.wtime()
        tree.computegradient()
        t2 = O.wtime()
        walks.append((t1 - t0, t2 - t1))
    walks.sort(key=lambda w: w[0] + w[1])
    tp, tg = walks[len(walks) // 2]
    t_dep = None
    if deposit:
        ts = []
        for _ in range(3):
            t0 = O.wtime()
            O.cic_deposit(oc, col["grid"], 0)
            ts.append(O.wtime() - t0)
        t_dep = sorted(ts)[1]
    O.set_threads(prev)
    return nthr, t_build, tp, tg, t_dep


def compat_step(col, reps=3):
    """one step through the reference's C API (libtimothy_compat.so) from a
    host caller, as a reference build linking it would run: wall time per
    call, PCIe copies included (the compat step timer, a child process)"""
    import subprocess
    import tempfile
    from timothy_amd import build as B
    if not os.path.exists(B.STEP_TIMER):
        return None
    grid = col["grid"]
    types = col["types"]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cells.bin")
        with open(path, "wb") as f:
            np.array([col["n"], len(types), grid["G"]], np.int64).tofile(f)
            np.array([[t["size"], t["h"], t["h2"], t["h3"], t["h4"]] for t in types], np.float64).tofile(f)
            np.array(list(grid["lower"]) + [float(grid["resolution"])], np.float64).tofile(f)
            np.stack([col["x"], col["y"], col["z"], col["size"]], 1).astype(np.float64).tofile(f)
            np.asarray(col["young"], np.float32).tofile(f)
            np.asarray(col["ctype"], np.int32).tofile(f)
            np.asarray(col["phase"], np.int32).tofile(f)
        r = subprocess.run([B.STEP_TIMER, path, str(reps)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        log(f"compat step timer failed ({r.returncode}): {r.stderr[-500:]}")
        return None
    out = json.loads(r.stdout)
    out["note"] = ("octbuild .. fieldgradient in src/step.c order through libtimothy_compat.so from host "
                   "celldata_t arrays, one 256^3 FP64 field, FP64 sums; median of the reps after a warm-up")
    return out


def direct_path(tm, ctx, a, build_ms, cic_ms):
    """the same step without LDS tiles (sph_tiles = 0): neighbour lists
    without the tile build, the direct pair kernels at their best shape
    (workgroup x records per trip x lanes per cell); the context is left as
    it was"""
    keep = {k: ctx.get_param(k) for k in ("sph_tiles", "sph_block", "sph_lanes", "sph_unroll_mixed",
                                          "sph_unroll_fp64")}
    ctx.set_param("sph_tiles", 0)
    ctx.neighbors()
    nbr = ctx.time_stage(tm.STAGE_NEIGHBORS, tm.MIXED, warmup=1, iters=3)
    best = {}
    for prec, name, key in ((tm.MIXED, "mixed", "sph_unroll_mixed"), (tm.FP64, "fp64", "sph_unroll_fp64")):
        for lanes in ((1, 2, 4, 8) if prec == tm.MIXED else (1,)):
            for bs in (64, 128, 256):
                for u in (1, 2, 4, 8):
                    ctx.set_param("sph_lanes", lanes)
                    ctx.set_param("sph_block", bs)
                    ctx.set_param(key, u)
                    ms = ctx.time_stage(tm.STAGE_SPH, prec, warmup=3, iters=20)
                    if name not in best or ms < best[name]["ms"]:
                        best[name] = {"ms": ms, "lanes": lanes, "sph_block": bs, "unroll": u}
    for k, v in keep.items():
        ctx.set_param(k, v)
    ctx.neighbors()
    return {"neighbors_ms": nbr, "sph_mixed": best["mixed"], "sph_fp64": best["fp64"],
            "full_step_ms_mixed": build_ms + nbr + best["mixed"]["ms"] + cic_ms,
            "note": "sph_tiles = 0: the tiles make the pair kernel faster but cost more to build than that saves "
                    "in a step at this size"}


def gpu_pairs(tm, col):
    with tm.Context(0) as c:
        c.upload_colony(col)
        c.step(tm.FP64)
        return c.pair_count()[0]


def copy_bandwidth(tm, ctx, mib=2048, iters=10):
    """device-to-device copy rate (read + write bytes / s): the achievable HBM figure"""
    nb = mib << 20
    src = tm.DeviceArray(ctx, nb, np.uint8)
    dst = tm.DeviceArray(ctx, nb, np.uint8)
    src.zero()
    f = tm.lib().tm_memcpy_dtod
    ctx._check(f(ctx.h, dst.ptr, src.ptr, nb))
    t0 = time.perf_counter()
    for _ in range(iters):
        ctx._check(f(ctx.h, dst.ptr, src.ptr, nb))
    dt = time.perf_counter() - t0
    src.free()
    dst.free()
    return 2.0 * nb * iters / dt / 1e9


def pair_abs_sum(col, off, nbr, chunk=1 << 21):
    """per cell sum_q |f_pq| of the FP64 pair force (normalises the force error)"""
    n = col["n"]
    h = np.array([t["h"] for t in col["types"]])
    h4 = np.array([t["h4"] for t in col["types"]])
    X = np.stack([col["x"], col["y"], col["z"]], 1)
    out = np.zeros(n)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        cnt = np.diff(off[s:e + 1])
        rows = np.repeat(np.arange(s, e), cnt)
        q = nbr[off[s]:off[e]]
        d = np.sqrt(((X[rows] - X[q]) ** 2).sum(1))
        ct = col["ctype"][rows]
        u = d / h[ct]
        w = np.where(u <= 0.5, -2 * u + 3 * u * u, -(1 - u) ** 2)
        out[s:e] = np.bincount(rows - s, np.abs(w) * 48 / (np.pi * h4[ct]), minlength=e - s)
    return out


def sph_errors(col, off, nbr, mix, ref):
    sabs = pair_abs_sum(col, off, nbr)
    has = sabs > 0
    fe = np.abs(mix["force"] - ref["force"]).max(1)[has] / sabs[has]
    fr = np.abs(mix["force"] - ref["force"]).max() / np.abs(ref["force"]).max()
    pos = ref["density"] > 0
    de = np.abs(mix["density"] - ref["density"])[pos] / ref["density"][pos]
    return dict(force_max=float(fe.max()), force_rms=float(np.sqrt((fe ** 2).mean())),
                force_raw_max_over_max=float(fr),
                density_max=float(de.max()), density_rms=float(np.sqrt((de ** 2).mean())),
                mindist_equal=bool(np.array_equal(mix["mindist"], ref["mindist"])),
                v_max_abs_over_max=float(np.abs(mix["v"] - ref["v"]).max() / max(np.abs(ref["v"]).max(), 1e-300)))


def interp_errors(tm, gen, ctx, col, field_seed=5):
    """FP32 field + FP32 gradient interpolated, against FP64 field + FP64
    gradient of the same (upcast) field"""
    grid = col["grid"]
    G = int(grid["G"])
    f32 = tm.DeviceArray.from_host(ctx, gen.smooth_field(grid, field_seed, np.float32))
    g32 = tm.DeviceArray(ctx, 3 * G ** 3, np.float32)
    ctx.field_gradient(grid, f32, g32)
    o32 = tm.DeviceArray(ctx, 4 * col["n"], np.float64)
    ctx.cic_interp(grid, f32, g32, o32)
    a = o32.get().reshape(-1, 4)
    f64 = tm.DeviceArray.from_host(ctx, f32.get().astype(np.float64))
    g64 = tm.DeviceArray(ctx, 3 * G ** 3, np.float64)
    ctx.field_gradient(grid, f64, g64)
    ctx.cic_interp(grid, f64, g64, o32)
    b = o32.get().reshape(-1, 4)
    for d in (f32, g32, o32, f64, g64):
        d.free()
    out = {}
    for k, name in enumerate(("value", "gx", "gy", "gz")):
        sc = max(np.abs(b[:, k]).max(), 1e-300)
        e = np.abs(a[:, k] - b[:, k]) / sc
        out[name] = dict(max=float(e.max()), rms=float(np.sqrt((e ** 2).mean())))
    out["normalised_by"] = "max |FP64 value| per component"
    return out


def environment_solve(tm, ctx, grid, prod):
    """one backward-Euler step of the field on the colony's grid with the
    deposit as its source (the reference's 7-point system and tolerance), the
    field starting at the boundary value; diffusion scaled so the step takes
    tens of CG iterations.  Algorithmic bytes of one CG iteration: 80 per
    node (new direction and A p: read r, p, write p, Ap; update: read x, p,
    r, Ap, write x, r)."""
    G = int(grid["G"])
    nodes = G ** 3
    field = tm.DeviceArray.from_host(ctx, np.full(nodes, 0.1575))
    res = float(grid["resolution"])
    dt = 16000.0
    D = 2.0 * res * res / dt  # z = 2
    par = tm.Diffusion.make(D, dt, 0.1575, 0.0, 1e-12, 200)
    ctx.diffusion_step(grid, par, field, prod)  # sizes the work arrays
    field.set(np.full(nodes, 0.1575))
    ctx.sync()
    t0 = time.perf_counter()
    it, rel = ctx.diffusion_step(grid, par, field, prod)
    ms = (time.perf_counter() - t0) * 1e3
    field.free()
    per = ms / max(it, 1)
    return {"z": D * dt / (res * res), "iterations": it, "relative_residual": rel, "ms": ms,
            "ms_per_iteration": per, "gbs_per_iteration": 80.0 * nodes / (per * 1e-3) / 1e9,
            "note": "wall time of the call, host checks of convergence every 8 iterations included"}


def time_sph(tm, ctx, a, n, P):
    B = 76.0 * n + 4.0 * P
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
        out[name] = dict(ms=ms, pairs_per_s=P / (ms * 1e-3), gbs=B / (ms * 1e-3) / 1e9,
                         roofline_frac=B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, wall_s=wall)
    return out, B


def roofline(B, ms, traffic):
    ach = B / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic, "algorithmic_bytes_per_launch": B, "kernel": "k_sph (mixed)"}


def line_common(a, out, n, P, maxnbr):
    mixed = out.get("mixed") or out.get("fp64")
    return {
        "metric": METRIC,
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
        "data": "synthetic (seeded)",
        "fp64": out.get("fp64"),
        "mixed": out.get("mixed"),
        "speedup_mixed_vs_fp64": (out["fp64"]["ms"] / out["mixed"]["ms"]) if len(out) == 2 else None,
    }


def run_single(a, tm, gen):
    col = gen.config(1, scale=a.scale)
    n = col["n"]
    ctx = tm.Context(0)
    ctx.upload_colony(col)
    ctx.step(tm.FP64)
    P, maxnbr = ctx.pair_count()
    ref = ctx.results()  # FP64 drop-in, bit-exact with the CPU path
    tune = autotune(tm, ctx, a)
    log(f"single colony n={n} P={P} max_nbr={maxnbr}")
    out, B = time_sph(tm, ctx, a, n, P)
    err = None
    if "mixed" in out:
        ctx.sph(tm.MIXED)
        mix = ctx.results()
        off, nbr = ctx.neighbor_lists()
        err = sph_errors(col, off, nbr, mix, ref)
        ctx.sph(tm.FP64)
    ierr = interp_errors(tm, gen, ctx, col)
    # full step: sort + tree + neighbour lists + SPH (mixed) + CIC deposit 256^3 FP64
    grid = col["grid"]
    prod = tm.DeviceArray(ctx, int(grid["G"]) ** 3, np.float64)
    ctx.cic_deposit(grid, prod, 0)
    step_ms = ctx.time_stage(tm.STAGE_STEP, tm.MIXED, warmup=1, iters=3)
    build_ms = ctx.time_stage(tm.STAGE_BUILD, tm.MIXED, warmup=1, iters=3)
    nbr_ms = ctx.time_stage(tm.STAGE_NEIGHBORS, tm.MIXED, warmup=1, iters=3)
    cic_ms = ctx.time_stage(tm.STAGE_CIC, tm.MIXED, warmup=1, iters=3)
    tiles_ms = ctx.time_stage(tm.STAGE_TILES, tm.MIXED, warmup=2, iters=10)
    direct = direct_path(tm, ctx, a, build_ms, cic_ms)
    env = environment_solve(tm, ctx, grid, prod)
    prod.free()
    copy_gbs = copy_bandwidth(tm, ctx)
    # resident mechanics loop: every replay of the captured graph is one whole
    # step of the colony in HBM -- sort, tree, lists, mixed SPH, and the move
    # (updatepositions with counter-based draws); wall time per replay,
    # including the host's check of the overflow flags
    ctx.graph_capture_loop(tm.Move.make([0.05], seed=1, step=0), tm.MIXED, None)
    for _ in range(2):
        ctx.graph_launch()
    ctx.sync()
    t0 = time.perf_counter()
    nloop = max(3, a.steps // 20)
    for _ in range(nloop):
        ctx.graph_launch()
    ctx.sync()
    loop_ms = (time.perf_counter() - t0) * 1e3 / nloop
    ctx.graph_destroy()

    compat = compat_step(col)
    cpu = cpu1 = None
    if not a.no_cpu:
        nthr, t_build, tp, tg, t_dep = cpu_walks(col, a.cpu_threads, deposit=True)
        cpu = {"value": P / (tp + tg), "unit": "pair-interactions/s", "cores": nthr, "kind": "port",
               "sample": f"full workload ({n} cells, {P} pairs): computepotential + computegradient walks, "
                         f"median of 3 after a warm-up; serial octbuild {t_build:.2f} s and serial CIC deposit "
                         f"{t_dep:.3f} s timed beside it (full_step_s)",
               "t_potential_s": tp, "t_gradient_s": tg, "t_octbuild_s": t_build, "t_cic_deposit_s": t_dep,
               "full_step_s": t_build + tp + tg + t_dep}
        # one core, on a bounded sample: the same generator at a tenth of the
        # cells (same density, so the same neighbours per cell)
        smp = gen.config(1, scale=0.1 * a.scale)
        Ps = gpu_pairs(tm, smp)
        _, b1, p1, g1, _ = cpu_walks(smp, 1, reps=1)
        cpu1 = {"value": Ps / (p1 + g1), "unit": "pair-interactions/s", "cores": 1, "kind": "port",
                "sample": f"same generator at a tenth of the cells ({smp['n']} cells, {Ps} pairs, same density): "
                          f"computepotential + computegradient walks on one thread, one run after a warm-up",
                "t_potential_s": p1, "t_gradient_s": g1, "t_octbuild_s": b1}
    line = line_common(a, out, n, P, maxnbr)
    mixed = out.get("mixed") or out.get("fp64")
    sph_mixed = out["mixed"]["ms"] if "mixed" in out else None
    sph_fp64 = out["fp64"]["ms"] if "fp64" in out else None
    # like for like: the tile build exists only to feed the pair kernel, and
    # the CPU's two walks both find the neighbours and sum the pairs
    lfl = {"tiles_ms": tiles_ms, "walk_ms": nbr_ms - tiles_ms, "neighbors_ms": nbr_ms}
    for name, t in (("mixed", sph_mixed), ("fp64", sph_fp64)):
        if t is None:
            continue
        lfl[name] = {"sph_incl_tiles_ms": tiles_ms + t,
                     "pairs_per_s_incl_tiles": P / ((tiles_ms + t) * 1e-3),
                     "neighbors_plus_sph_ms": nbr_ms + t,
                     "pairs_per_s_neighbors_plus_sph": P / ((nbr_ms + t) * 1e-3)}
        if cpu:
            lfl[name]["speedup_vs_cpu_walks"] = (cpu["t_potential_s"] + cpu["t_gradient_s"]) / ((nbr_ms + t) * 1e-3)
    gpu_full = step_ms + cic_ms
    line.update({
        "data": "synthetic (seeded Gaussian spheroid)",
        "config": {"workload": "single colony, 2,000,000 cells, Gaussian spheroid W=146.6, r=1, h=2, mixed SPH"
                   if a.scale == 1.0 else f"single colony scaled {a.scale}", "cells": n, "pairs": P,
                   "max_neighbours": maxnbr, "parallelism": "1 GPU"},
        "mixed_vs_fp64_error": err,
        "interp_fp32_vs_fp64_error_256^3": ierr,
        "full_step_ms": {"total": gpu_full, "build": build_ms, "neighbors": nbr_ms, "tiles_in_neighbors": tiles_ms,
                         "sph_mixed": sph_mixed, "cic_deposit_256^3": cic_ms},
        "like_for_like": lfl,
        "without_tiles": direct,
        "mechanics_loop_ms_per_step": loop_ms,
        "compat_step_host_arrays": compat,
        "environment_solve_256^3": env,
        "copy_kernel_gbs": copy_gbs,
        "roofline": roofline(B, mixed["ms"], traffic_from_profiles()),
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "speedup_vs_cpu": (mixed["pairs_per_s"] / cpu["value"]) if cpu else None,
        "speedup_full_step_vs_cpu": (cpu["full_step_s"] / (gpu_full * 1e-3)) if cpu else None,
        "autotune": tune,
    })
    line["roofline"]["frac_of_copy"] = line["roofline"]["achieved"] / copy_gbs
    ctx.close()
    return line


def run_batched(a, tm, gen):
    cols = gen.colonies_config3(scale=a.scale)
    ctx = tm.Context(0)
    off_c = ctx.upload_colonies(cols)
    n = int(off_c[-1])
    ctx.step(tm.FP64)
    P, maxnbr = ctx.pair_count()
    ref = ctx.results()
    log(f"batched: {len(cols)} colonies, n={n} P={P} max_nbr={maxnbr}")
    tune = autotune(tm, ctx, a)
    out, B = time_sph(tm, ctx, a, n, P)
    # graph-captured step: build + neighbour lists + mixed SPH for all colonies
    ctx.graph_capture(tm.MIXED)
    for _ in range(a.warmup):
        ctx.graph_launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.graph_launch()
    ctx.sync()
    graph_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    eager_ms = ctx.time_stage(tm.STAGE_STEP, tm.MIXED, warmup=1, iters=max(3, a.steps // 4))
    ctx.sph(tm.MIXED)
    mix = ctx.results()
    off, nbr = ctx.neighbor_lists()
    # per-colony check against FP64
    sabs = pair_abs_sum(dict(n=n, x=np.concatenate([c["x"] for c in cols]),
                             y=np.concatenate([c["y"] for c in cols]), z=np.concatenate([c["z"] for c in cols]),
                             ctype=np.concatenate([c["ctype"] for c in cols]), types=cols[0]["types"]), off, nbr)
    worst, rms_acc, cnt = 0.0, 0.0, 0
    dworst = 0.0
    for i in range(len(cols)):
        s, e = int(off_c[i]), int(off_c[i + 1])
        has = sabs[s:e] > 0
        fe = np.abs(mix["force"][s:e] - ref["force"][s:e]).max(1)[has] / sabs[s:e][has]
        if fe.size:
            worst = max(worst, float(fe.max()))
            rms_acc += float((fe ** 2).sum())
            cnt += fe.size
        pos = ref["density"][s:e] > 0
        de = np.abs(mix["density"][s:e] - ref["density"][s:e])[pos] / ref["density"][s:e][pos]
        if de.size:
            dworst = max(dworst, float(de.max()))
    err = dict(force_max_worst_colony=worst, force_rms=float(np.sqrt(rms_acc / max(cnt, 1))),
               density_max_worst_colony=dworst, mindist_equal=bool(np.array_equal(mix["mindist"], ref["mindist"])))
    cpu = None
    if not a.no_cpu:
        # bounded sample: the first 32 colonies, one oracle octree and walk pair each
        k = 32
        tp = tg = tb = 0.0
        Pk = int(off[int(off_c[k])])
        nthr = 0
        for c in cols[:k]:
            nthr, t_build, p_, g_, _ = cpu_walks(c, a.cpu_threads)
            tp, tg, tb = tp + p_, tg + g_, tb + t_build
        cpu = {"value": Pk / (tp + tg), "unit": "pair-interactions/s", "cores": nthr, "kind": "port",
               "sample": f"first {k} of {len(cols)} colonies ({int(off_c[k])} cells, {Pk} pairs), "
                         f"computepotential + computegradient walks per colony, median of 3",
               "t_potential_s": tp, "t_gradient_s": tg, "t_octbuild_s": tb}
    line = line_common(a, out, n, P, maxnbr)
    mixed = out.get("mixed") or out.get("fp64")
    line.update({
        "data": "synthetic (seeded: uniform balls and Gaussian spheroids)",
        "config": {"workload": f"batched: {len(cols)} colonies of 1,000..200,000 cells (log-uniform), "
                               f"{n} cells, r=1, h=2, one launch per stage, graph-captured mixed step",
                   "colonies": len(cols), "cells": n, "pairs": P, "max_neighbours": maxnbr, "parallelism": "1 GPU"},
        "mixed_vs_fp64_error": err,
        "graph_step_ms": graph_ms,
        "eager_step_ms": eager_ms,
        "roofline": roofline(B, mixed["ms"], None),
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (mixed["pairs_per_s"] / cpu["value"]) if cpu else None,
        "autotune": tune,
    })
    ctx.close()
    return line


def time_host(ctx, fn, iters):
    fn()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(iters):
        fn()
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3 / iters


def run_large(a, tm, gen):
    col = gen.config(4, scale=a.scale)
    n = col["n"]
    grid = col["grid"]
    G = int(grid["G"])
    log(f"large: n={n} grid {G}^3")
    ctx = tm.Context(0)
    ctx.upload_colony(col)
    ctx.step(tm.FP64)
    P, maxnbr = ctx.pair_count()
    ref = ctx.results()
    log(f"large: P={P} max_nbr={maxnbr}")
    tune = autotune(tm, ctx, a)
    out, B = time_sph(tm, ctx, a, n, P)
    err = None
    if "mixed" in out:
        ctx.sph(tm.MIXED)
        mix = ctx.results()
        off, nbr = ctx.neighbor_lists()
        err = sph_errors(col, off, nbr, mix, ref)
        del off, nbr
    build_ms = ctx.time_stage(tm.STAGE_BUILD, tm.MIXED, warmup=1, iters=3)
    nbr_ms = ctx.time_stage(tm.STAGE_NEIGHBORS, tm.MIXED, warmup=1, iters=3)
    prod = tm.DeviceArray(ctx, G ** 3, np.float64)
    ctx.cic_deposit(grid, prod, 1)
    dep_ms = ctx.time_stage(tm.STAGE_CIC, tm.MIXED, warmup=1, iters=3)
    mass = float(prod.get().sum())
    vol = float((4.0 / 3.0 * np.pi * col["size"] ** 3).sum())
    prod.free()
    f32 = tm.DeviceArray.from_host(ctx, gen.smooth_field(grid, 5, np.float32))
    g32 = tm.DeviceArray(ctx, 3 * G ** 3, np.float32)
    o = tm.DeviceArray(ctx, 4 * n, np.float64)
    grad_ms = time_host(ctx, lambda: ctx.field_gradient(grid, f32, g32), 5)
    interp_ms = time_host(ctx, lambda: ctx.cic_interp(grid, f32, g32, o), 5)
    for d in (f32, g32, o):
        d.free()
    ierr = interp_errors(tm, gen, ctx, col)
    cpu = None
    if not a.no_cpu:
        smp = gen.config(4, scale=0.04 * a.scale)
        Ps = gpu_pairs(tm, smp)
        nthr, t_build, tp, tg, _ = cpu_walks(smp, a.cpu_threads)
        cpu = {"value": Ps / (tp + tg), "unit": "pair-interactions/s", "cores": nthr, "kind": "port",
               "sample": f"same generator at 4% scale ({smp['n']} cells, {Ps} pairs, same density and type mix): "
                         f"computepotential + computegradient walks, median of 3",
               "t_potential_s": tp, "t_gradient_s": tg, "t_octbuild_s": t_build}
    line = line_common(a, out, n, P, maxnbr)
    mixed = out.get("mixed") or out.get("fp64")
    G3 = G ** 3
    line.update({
        "data": "synthetic (seeded Gaussian spheroid, 5% type 1)",
        "config": {"workload": f"large single colony: {n} cells, 5% radius 1.5 (h 3), Gaussian W=428.7, "
                               f"mixed SPH + CIC volume deposit {G}^3 FP64 + FP32 field gradient/interpolation {G}^3",
                   "cells": n, "pairs": P, "max_neighbours": maxnbr, "grid": G, "parallelism": "1 GPU"},
        "mixed_vs_fp64_error": err,
        "interp_fp32_vs_fp64_error": ierr,
        "stage_ms": {"build": build_ms, "neighbors": nbr_ms, "sph_mixed": out.get("mixed", {}).get("ms"),
                     "sph_fp64": out.get("fp64", {}).get("ms"), "cic_deposit_volume": dep_ms,
                     "field_gradient_f32": grad_ms, "cic_interp_f32": interp_ms},
        # algorithmic bytes: deposit 32 N + 8 G, gradient (4 + 12) G, interpolation 56 N + 16 G
        "stage_gbs": {"cic_deposit_volume": (32.0 * n + 8.0 * G3) / (dep_ms * 1e-3) / 1e9,
                      "field_gradient_f32": 16.0 * G3 / (grad_ms * 1e-3) / 1e9,
                      "cic_interp_f32": (56.0 * n + 16.0 * G3) / (interp_ms * 1e-3) / 1e9},
        "deposit_mass_rel_err": abs(mass - vol) / vol,
        "roofline": roofline(B, mixed["ms"], None),
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (mixed["pairs_per_s"] / cpu["value"]) if cpu else None,
        "autotune": tune,
    })
    ctx.close()
    return line


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import timothy_amd as tm
    from timothy_amd import generators as gen

    run = {"single": run_single, "batched": run_batched, "large": run_large}[a.workload]
    line = run(a, tm, gen)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
