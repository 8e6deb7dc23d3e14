"""Build the MI355X libraries in-tree.

  timothy_amd/libtimothy_hip.so     HIP kernels + native C ABI (include/timothy_hip.h)
  timothy_amd/libtimothy_compat.so  the reference's octree/potential/patches/
                                    fieldgradient C API over it (include/timothy_compat.h)

All device code is compiled for gfx950 only, without FMA contraction (the
FP64 path reproduces the reference's rounding) and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtimothy_hip.so")
COMPAT = os.path.join(HERE, "libtimothy_compat.so")

HIP_SOURCES = ["tm_api.hip", "tm_sort.hip", "tm_build.hip", "tm_neighbors.hip", "tm_tiles.hip", "tm_sph.hip",
               "tm_grid.hip", "tm_octree.hip", "tm_move.hip",
               "tm_diffusion.hip", "tm_vtk.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    noise = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def _obj(src):
    return os.path.join(OUT, os.path.splitext(src)[0] + ".o")


def build(verbose: bool = False) -> None:
    os.makedirs(OUT, exist_ok=True)
    jobs = min(8, len(HIP_SOURCES))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
                for s in HIP_SOURCES]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [_obj(s) for s in HIP_SOURCES])
    os.replace(tmp, LIB)
    # reference-API layer: plain C over the native ABI
    if not os.path.exists(os.path.join(CSRC, "compat", "tm_compat.c")):
        return
    tmp = COMPAT + ".tmp"
    _run(["gcc", "-std=gnu11", "-O2", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter",
          "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "compat", "tm_compat.c"), "-o", tmp,
          "-L" + HERE, "-ltimothy_hip", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(tmp, COMPAT)
    build_step_timer()
    build_step_harness()
    if verbose:
        print("built", LIB, COMPAT, HARNESS)


HARNESS = os.path.join(ROOT, "tests", "compat", "_build", "step_harness")
STEP_TIMER = os.path.join(OUT, "step_timer")


def build_step_timer() -> None:
    """the program bench.py times the compat step with (a reference-style
    host caller of libtimothy_compat.so)"""
    tmp = STEP_TIMER + ".tmp"
    _run(["gcc", "-std=gnu11", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "compat", "step_timer.c"), "-o", tmp, "-L" + HERE, "-ltimothy_compat", "-ltimothy_hip",
          "-Wl,-rpath,$ORIGIN/..", "-lm"])
    os.replace(tmp, STEP_TIMER)


ENV_HARNESS = os.path.join(ROOT, "tests", "compat", "_build", "env_harness")


def build_step_harness() -> None:
    """The C programs the drop-in tests run: they link libtimothy_compat.so the
    way the reference's own executable would link octree.o/potential.o/..."""
    for name, out in (("step_harness.c", HARNESS), ("env_harness.c", ENV_HARNESS)):
        src = os.path.join(ROOT, "tests", "compat", name)
        if not os.path.exists(src):
            continue
        os.makedirs(os.path.dirname(out), exist_ok=True)
        tmp = out + ".tmp"
        _run(["gcc", "-std=gnu11", "-O2", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"), src, "-o", tmp,
              "-L" + HERE, "-ltimothy_compat", "-ltimothy_hip", "-Wl,-rpath,$ORIGIN/../../../timothy_amd", "-lm"])
        os.replace(tmp, out)


if __name__ == "__main__":
    build(verbose=True)
