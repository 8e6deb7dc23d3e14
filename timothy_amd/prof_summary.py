"""Short text summaries of rocprofv3 output databases (kernel times, HBM bytes).

    python -m timothy_amd.prof_summary kernels <dir> [--cmd "..."]
        per-kernel calls / mean / total / share from a --kernel-trace --stats run
    python -m timothy_amd.prof_summary traffic <fetch_dir> <write_dir> --kernel k_sph
            --copy-bytes BYTES [--json out.json] [--cmd "..."]
        mean FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of one kernel, each
        scaled by what the same counter reads for the device-to-device copy
        kernel of BYTES bytes per dispatch in the same run (read factor =
        BYTES / FETCH_SIZE, write factor = BYTES / WRITE_SIZE, over the copy
        dispatches of at least half the largest one); prints HBM bytes
        per launch and, with --json, writes them for the benchmark to report.

Reads the SQLite databases rocprofv3 writes (*_results.db); nothing here
touches the GPU.
"""
from __future__ import annotations

import argparse
import collections
import glob
import json
import os
import sqlite3
import sys


def _db(d):
    dbs = sorted(glob.glob(os.path.join(d, "**", "*.db"), recursive=True))
    if not dbs:
        raise SystemExit(f"no rocprofv3 database under {d}")
    return sqlite3.connect(dbs[0])


def _short(name):
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def kernels(d, top=25):
    """per kernel: dispatches, mean / min duration (ns in the kernels view),
    share of all kernel time, registers and LDS of the launch"""
    con = _db(d)
    agg = collections.defaultdict(list)
    meta = {}
    for name, dur, vg, agpr, sg, lds, wg in con.execute(
            "select name, duration, vgpr_count, accum_vgpr_count, sgpr_count, lds_size, workgroup_x from kernels"):
        k = _short(name)
        agg[k].append(float(dur))
        meta[k] = (vg, agpr, sg, lds, wg)
    total = sum(sum(v) for v in agg.values())
    rows = sorted(agg.items(), key=lambda kv: -sum(kv[1]))
    out = ["| kernel | calls | mean us | min us | total ms | share % | vgpr/agpr/sgpr | lds B | wg |",
           "|---|---:|---:|---:|---:|---:|---|---:|---:|"]
    for k, v in rows[:top]:
        vg, agpr, sg, lds, wg = meta[k]
        out.append(f"| `{k}` | {len(v)} | {_mean(v) / 1e3:.2f} | {min(v) / 1e3:.2f} | {sum(v) / 1e6:.3f} | "
                   f"{100.0 * sum(v) / total:.1f} | {vg}/{agpr}/{sg} | {lds} | {wg} |")
    return "\n".join(out)


def counters(d, counter):
    con = _db(d)
    cols = [c[1] for c in con.execute("pragma table_info(counters_collection)")]
    agg = collections.defaultdict(list)
    for r in con.execute("select * from counters_collection"):
        rec = dict(zip(cols, r))
        if rec.get("counter_name") != counter:
            continue
        agg[_short(rec.get("kernel_name", ""))].append(float(rec.get("value")))
    return agg


def _mean(v):
    return sum(v) / len(v)


def traffic(fetch_dir, write_dir, kernel, copy_bytes):
    f = counters(fetch_dir, "FETCH_SIZE")
    w = counters(write_dir, "WRITE_SIZE")

    def pick(agg, key):
        ks = [k for k in agg if k.startswith(key)]
        if not ks:
            raise SystemExit(f"no dispatch of {key} in {sorted(agg)}")
        return ks[0], agg[ks[0]]

    kf, vf = pick(f, kernel)
    _, vw = pick(w, kernel)
    # full dispatches only (a solver's dispatches after convergence return at once)
    vf = [v for v in vf if v >= 0.5 * max(vf)]
    vw = [v for v in vw if v >= 0.5 * max(vw)]
    _, cf = pick(f, "__amd_rocclr_copyBuffer")
    _, cw = pick(w, "__amd_rocclr_copyBuffer")
    # the calibration copy's dispatches only (other small copies in the run are dropped)
    cf = [v for v in cf if v >= 0.5 * max(cf)]
    cw = [v for v in cw if v >= 0.5 * max(cw)]
    rf = copy_bytes / (_mean(cf) * 1024.0)
    wf = copy_bytes / (_mean(cw) * 1024.0)
    rb = _mean(vf) * 1024.0 * rf
    wb = _mean(vw) * 1024.0 * wf
    return dict(kernel=kf, dispatches=len(vf), copy_dispatches=len(cf), fetch_kib=_mean(vf), write_kib=_mean(vw),
                copy_fetch_kib=_mean(cf),
                copy_write_kib=_mean(cw), copy_bytes_per_dispatch=copy_bytes, read_factor=rf, write_factor=wf,
                read_bytes=rb, write_bytes=wb, hbm_bytes_per_launch=rb + wb)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--cmd", default="")
    k.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--json")
    t.add_argument("--cmd", default="")
    a = ap.parse_args(argv)
    if a.what == "kernels":
        if a.cmd:
            print(f"command: `{a.cmd}`\n")
        print(kernels(a.dir, a.top))
    else:
        r = traffic(a.fetch_dir, a.write_dir, a.kernel, a.copy_bytes)
        r["command"] = a.cmd
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(r, fh, indent=1)
        print(json.dumps(r, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
