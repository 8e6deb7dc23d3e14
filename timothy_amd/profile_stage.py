"""Run one stage of the cell-mechanics path repeatedly, for profilers.

    python -m timothy_amd.profile_stage --stage sph --precision mixed --iters 20

Builds the 2,000,000-cell Gaussian-spheroid colony (the benchmark workload,
or a scaled copy with --scale), prepares the stages the chosen one depends on,
then launches the chosen stage --iters times on the library's stream.  Used
with rocprofv3 to read kernel durations and HBM counters of one kernel.
"""
import argparse
import json
import sys

import numpy as np

import timothy_amd as tm
from timothy_amd import generators as gen

STAGES = {"build": tm.STAGE_BUILD, "neighbors": tm.STAGE_NEIGHBORS, "sph": tm.STAGE_SPH, "cic": tm.STAGE_CIC,
          "step": tm.STAGE_STEP}
EXTRA = ("move", "diffusion")  # driven from here: step + move, or deposit + environment solve


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--stage", choices=sorted(STAGES) + list(EXTRA), default="sph")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--param", action="append", default=[], help="key=value tunables")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also copy this many MiB device-to-device once (counter calibration)")
    a = ap.parse_args(argv)
    prec = tm.MIXED if a.precision == "mixed" else tm.FP64
    col = gen.config(1, scale=a.scale)
    with tm.Context(0) as ctx:
        for kv in a.param:
            k, v = kv.split("=")
            ctx.set_param(k, int(v))
        ctx.upload_colony(col)
        ctx.step(prec)
        if a.stage == "cic":
            prod = tm.DeviceArray(ctx, int(col["grid"]["G"]) ** 3, np.float64)
            ctx.cic_deposit(col["grid"], prod, 0)
        P, _ = ctx.pair_count()
        if a.copy_mib:
            nb = a.copy_mib << 20
            src = tm.DeviceArray(ctx, nb, np.uint8)
            dst = tm.DeviceArray(ctx, nb, np.uint8)
            src.zero()
            ctx._check(tm.lib().tm_memcpy_dtod(ctx.h, dst.ptr, src.ptr, nb))
        if a.stage == "move":
            # each iteration: a mixed step, then the cells move (Philox draws)
            import time
            t0 = time.perf_counter()
            for k in range(a.iters):
                ctx.step(prec)
                ctx.update_positions(tm.Move.make([0.05], seed=1, step=k))
            ms = (time.perf_counter() - t0) * 1e3 / a.iters
        elif a.stage == "diffusion":
            # each iteration: one environment solve on the colony's 256^3 grid, the
            # deposit as its source, z = 2 (tens of CG iterations)
            import time
            grid = col["grid"]
            nodes = int(grid["G"]) ** 3
            prod = tm.DeviceArray(ctx, nodes, np.float64)
            ctx.cic_deposit(grid, prod, 0)
            field = tm.DeviceArray(ctx, nodes, np.float64)
            res = float(grid["resolution"])
            par = tm.Diffusion.make(2.0 * res * res / 16000.0, 16000.0, 0.1575, 0.0, 1e-12, 200)
            init = np.full(nodes, 0.1575)
            t0 = time.perf_counter()
            its = 0
            for _ in range(a.iters):
                field.set(init)
                its += ctx.diffusion_step(grid, par, field, prod)[0]
            ms = (time.perf_counter() - t0) * 1e3 / a.iters
            print(json.dumps({"cg_iterations_per_solve": its / a.iters}), flush=True)
        else:
            ms = ctx.time_stage(STAGES[a.stage], prec, warmup=0, iters=a.iters)
        print(json.dumps({"stage": a.stage, "precision": a.precision, "cells": col["n"], "pairs": P,
                          "iters": a.iters, "ms_per_iter": ms}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
