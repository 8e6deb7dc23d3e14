"""CPU checks of the shipped libraries (no device calls): both shared
libraries load and export every function their public header declares; the
ctypes mirror has a signature for every native entry point; the synthetic
workload generators are deterministic."""
import ctypes as C
import os
import re

import numpy as np

import timothy_amd as tm
from timothy_amd import build as B
from timothy_amd import generators as gen

INC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def _declared(header, pattern):
    txt = open(os.path.join(INC, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    txt = re.sub(r"typedef[^;]*;", "", txt, flags=re.S)
    txt = re.sub(r"^[ \t]*#[^\n]*(\\\n[^\n]*)*", "", txt, flags=re.M)  # preprocessor lines
    txt = re.sub(r"TM_LAYOUT_ASSERT\([^;]*\);", "", txt)  # compile-time layout checks
    return sorted(set(re.findall(pattern, txt)))


def test_native_abi_exports():
    names = tm.header_functions()
    assert len(names) >= 40
    L = tm.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_header():
    names = tm.header_functions()
    L = tm.lib()
    unset = [n for n in names if getattr(L, n).restype is C.c_int and not getattr(L, n).argtypes]
    assert not unset, f"functions without a ctypes signature: {unset}"


def test_compat_abi_exports():
    names = _declared("timothy_compat.h", r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;")
    names = [n for n in names if n not in ("if", "while", "for", "sizeof", "return")]
    for must in ("octbuild", "octfree", "octheapinit", "octcomputebox", "octlocateregion", "octheappush",
                 "octheappop", "octheapfree", "octcomputeboxr", "computepotential", "computeremotepotential",
                 "computegradient", "computeremotegradient", "patches_alloc", "patches_free",
                 "patches_env2cellsinit", "patches_env2cellswait", "patches_cells2envinit",
                 "patches_cells2envwait", "fieldgradient", "tm_compat_shutdown"):
        assert must in names, must
    tm.lib()  # the compat library resolves its native symbols from here
    L = C.CDLL(B.COMPAT)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    for var in ("affShift", "affScale"):  # globals the reference defines in src/global.h
        assert hasattr(L, var)


def test_step_harness_built():
    assert os.access(B.HARNESS, os.X_OK)


def test_generators_deterministic():
    a = gen.config(1, scale=0.005)
    b = gen.config(1, scale=0.005)
    for k in ("x", "y", "z", "size", "young", "ctype", "phase"):
        assert np.array_equal(a[k], b[k])
    assert a["grid"]["G"] == b["grid"]["G"]
    c = gen.config(1, scale=0.005)
    c2 = gen.gaussian_colony(a["n"], a["seed"] + 1, 1.0)
    assert not np.array_equal(c["x"], c2["x"])
    cols = gen.colonies_config3(ncol=4, scale=0.01)
    assert len(cols) == 4 and all(col["n"] > 0 for col in cols)
