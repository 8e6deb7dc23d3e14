"""ctypes wrapper of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg as the timed CPU restatement of
the reference path.  The product (timothy_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class CellType(C.Structure):
    _fields_ = [("size", C.c_double), ("h", C.c_double), ("h2", C.c_double), ("h3", C.c_double), ("h4", C.c_double)]


class Cells(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("size", C.c_void_p),
        ("young", C.c_void_p), ("ctype", C.c_void_p), ("phase", C.c_void_p),
    ]


class Node(C.Structure):
    _fields_ = [
        ("xcode", C.c_uint32), ("ycode", C.c_uint32), ("zcode", C.c_uint32),
        ("xlimit", C.c_uint32), ("ylimit", C.c_uint32), ("zlimit", C.c_uint32), ("level", C.c_uint32),
        ("father", C.c_int64), ("child", C.c_int64 * 8), ("data", C.c_int32),
    ]


class Tree(C.Structure):
    _fields_ = [
        ("shift", C.c_double * 3), ("scale", C.c_double),
        ("loc", C.POINTER(C.c_uint32)), ("nodes", C.POINTER(Node)),
        ("nnodes", C.c_int64), ("cap", C.c_int64),
    ]


def build():
    """Compile the oracle with its Makefile (gcc, no fast-math, no contraction)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_octbuild.restype = C.c_int
        L.or_octbuild.argtypes = [C.POINTER(Cells), C.POINTER(CellType), C.POINTER(Tree)]
        L.or_treefree.argtypes = [C.POINTER(Tree)]
        L.or_computebox.argtypes = [C.POINTER(Cells), C.POINTER(CellType), C.POINTER(Tree), C.c_int64,
                                    u32p, u32p]
        L.or_locateregion.restype = C.c_int64
        L.or_locateregion.argtypes = [C.POINTER(Tree), u32p, u32p]
        L.or_computepotential.argtypes = [C.POINTER(Cells), C.POINTER(CellType), C.POINTER(Tree), C.c_int,
                                          f64p, f64p, f64p]
        L.or_computegradient.argtypes = [C.POINTER(Cells), C.POINTER(CellType), C.POINTER(Tree), C.c_int,
                                         f64p, f64p, f64p]
        L.or_walk_candidates.restype = C.c_int64
        L.or_walk_candidates.argtypes = [C.POINTER(Cells), C.POINTER(CellType), C.POINTER(Tree), i64p,
                                         C.c_void_p, C.c_void_p]
        L.or_brute_neighbors.restype = C.c_int64
        L.or_brute_neighbors.argtypes = [C.POINTER(Cells), C.POINTER(CellType), i64p, C.c_void_p]
        L.or_leaf_levels.argtypes = [C.POINTER(Tree), C.c_int64, i32p]
        L.or_sphkernel.restype = C.c_float
        L.or_sphkernel.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_sphgradient.argtypes = [C.c_int] + [C.c_double] * 9 + [f64p]
        L.or_potential.restype = C.c_double
        L.or_potential.argtypes = ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_float,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_float,
                                    C.POINTER(CellType), C.POINTER(C.c_double), C.POINTER(C.c_double)])
        L.or_cic_deposit.argtypes = [C.POINTER(Cells), C.c_int, f64p, C.c_float, i64p, i64p, i64p, f64p]
        L.or_cic_interp.argtypes = [C.POINTER(Cells), f64p, C.c_float, i64p, i64p, i64p, f64p, f64p, f64p]
        L.or_fieldgradient.argtypes = [i64p, C.c_float, f64p, f64p]
        L.or_wtime.restype = C.c_double
        L.or_max_threads.restype = C.c_int
        L.or_set_threads.argtypes = [C.c_int]
        L.or_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.or_philox4x32_10.restype = None
        L.or_philox_draws.argtypes = [C.c_int64, C.c_int, C.c_uint64, C.c_uint64, f32p]
        L.or_philox_draws.restype = None
        L.or_updatepositions.argtypes = [C.c_int64, C.c_int, f64p, f64p, f64p, f64p, i32p, f32p, f64p, f64p, f64p,
                                         C.c_void_p, f64p, C.c_void_p]
        L.or_updatepositions.restype = C.c_int64
        L.or_diffusion_rhs.argtypes = [i64p, C.c_float, C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                       f64p, f64p]
        L.or_diffusion_rhs.restype = None
        L.or_diffusion_apply.argtypes = [i64p, C.c_float, C.c_double, C.c_double, C.c_double, f64p, f64p]
        L.or_diffusion_apply.restype = None
        L.or_diffusion_solve.argtypes = [i64p, C.c_float, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_void_p, f64p, C.c_double, C.c_int, C.c_int]
        L.or_diffusion_solve.restype = C.c_int
        L.or_writevtk.argtypes = [C.c_char_p, C.c_int64, f64p, f64p, f64p, i32p, f64p, f64p, f64p, C.c_int]
        L.or_writevtk.restype = C.c_int
        _lib = L
    return _lib


def celltypes_struct(types):
    """types: list of dicts/tuples (size, h, h2, h3, h4) -> ctypes array"""
    arr = (CellType * len(types))()
    for i, t in enumerate(types):
        if isinstance(t, dict):
            t = (t["size"], t["h"], t["h2"], t["h3"], t["h4"])
        arr[i] = CellType(*[float(v) for v in t])
    return arr


class Colony:
    """Holds numpy SoA arrays alive for the C struct that points into them."""

    def __init__(self, x, y, z, size, young, ctype, phase=None):
        n = len(x)
        self.x = np.ascontiguousarray(x, np.float64)
        self.y = np.ascontiguousarray(y, np.float64)
        self.z = np.ascontiguousarray(z, np.float64)
        self.size = np.ascontiguousarray(size, np.float64)
        self.young = np.ascontiguousarray(young, np.float32)
        self.ctype = np.ascontiguousarray(ctype, np.int32)
        self.phase = np.ascontiguousarray(phase if phase is not None else np.ones(n), np.int32)
        self.n = n
        self.c = Cells(n, self.x.ctypes.data, self.y.ctypes.data, self.z.ctypes.data, self.size.ctypes.data,
                       self.young.ctypes.data, self.ctype.ctypes.data, self.phase.ctypes.data)

    @classmethod
    def from_dict(cls, d):
        return cls(d["x"], d["y"], d["z"], d["size"], d["young"], d["ctype"], d.get("phase"))


class Octree:
    def __init__(self, colony: Colony, types):
        self.colony = colony
        self.types = celltypes_struct(types)
        self.t = Tree()
        rc = lib().or_octbuild(C.byref(colony.c), self.types, C.byref(self.t))
        if rc != 0:
            lib().or_treefree(C.byref(self.t))
            raise ValueError(f"oracle octbuild failed ({rc}): degenerate codes")

    def __del__(self):
        try:
            lib().or_treefree(C.byref(self.t))
        except Exception:
            pass

    @property
    def shift(self):
        return np.array(self.t.shift[:], np.float64)

    @property
    def scale(self):
        return float(self.t.scale)

    def loc(self):
        n = self.colony.n
        if n == 0:
            return np.zeros((0, 3), np.uint32)
        return np.ctypeslib.as_array(self.t.loc, shape=(n * 3,)).copy().reshape(n, 3)

    def nodes(self):
        """structured numpy copy of the node array"""
        k = self.t.nnodes
        out = np.zeros(k, dtype=[("code", np.uint32, 3), ("limit", np.uint32, 3), ("level", np.uint32),
                                 ("father", np.int64), ("child", np.int64, 8), ("data", np.int32)])
        for i in range(k):
            nd = self.t.nodes[i]
            out[i]["code"] = (nd.xcode, nd.ycode, nd.zcode)
            out[i]["limit"] = (nd.xlimit, nd.ylimit, nd.zlimit)
            out[i]["level"] = nd.level
            out[i]["father"] = nd.father
            out[i]["child"] = nd.child[:]
            out[i]["data"] = nd.data
        return out

    def leaf_levels(self):
        out = np.zeros(self.colony.n, np.int32)
        lib().or_leaf_levels(C.byref(self.t), self.colony.n, out)
        return out

    def computebox(self, p):
        mn = np.zeros(3, np.uint32)
        mx = np.zeros(3, np.uint32)
        lib().or_computebox(C.byref(self.colony.c), self.types, C.byref(self.t), p, mn, mx)
        return mn, mx

    def locateregion(self, mn, mx):
        return int(lib().or_locateregion(C.byref(self.t), np.ascontiguousarray(mn, np.uint32),
                                         np.ascontiguousarray(mx, np.uint32)))

    def computepotential(self, dimension=3, init=None):
        n = self.colony.n
        dens = np.zeros(n) if init is None else init[0].copy()
        v = np.zeros(n) if init is None else init[1].copy()
        md = np.zeros(n) if init is None else init[2].copy()
        lib().or_computepotential(C.byref(self.colony.c), self.types, C.byref(self.t), dimension, dens, v, md)
        return dens, v, md

    def computegradient(self, dimension=3):
        n = self.colony.n
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().or_computegradient(C.byref(self.colony.c), self.types, C.byref(self.t), dimension, fx, fy, fz)
        return np.stack([fx, fy, fz], axis=1)

    def candidates(self):
        """(offsets[n+1], cand[int32], in_h[uint8]) in the reference's DFS pop order"""
        n = self.colony.n
        off = np.zeros(n + 1, np.int64)
        tot = lib().or_walk_candidates(C.byref(self.colony.c), self.types, C.byref(self.t), off, None, None)
        cand = np.zeros(max(tot, 1), np.int32)
        inh = np.zeros(max(tot, 1), np.uint8)
        lib().or_walk_candidates(C.byref(self.colony.c), self.types, C.byref(self.t), off,
                                 cand.ctypes.data, inh.ctypes.data)
        return off, cand[:tot], inh[:tot]


def brute_neighbors(colony: Colony, types):
    ts = celltypes_struct(types)
    n = colony.n
    off = np.zeros(n + 1, np.int64)
    tot = lib().or_brute_neighbors(C.byref(colony.c), ts, off, None)
    nbr = np.zeros(max(tot, 1), np.int32)
    lib().or_brute_neighbors(C.byref(colony.c), ts, off, nbr.ctypes.data)
    return off, nbr[:tot]


def sphkernel(dist, h, h2, h3, dimension=3):
    return lib().or_sphkernel(dimension, dist, h, h2, h3)


def sphgradient(p, q, h, h3, h4, dimension=3):
    f = np.zeros(3)
    lib().or_sphgradient(dimension, *[float(v) for v in p], *[float(v) for v in q], h, h3, h4, f)
    return f


def potential(p, psize, pyoung, q, qsize, qyoung, ptype, dimension=3, density=0.0, mindist=1e308):
    ts = celltypes_struct([ptype])
    d = C.c_double(density)
    m = C.c_double(mindist)
    r = lib().or_potential(dimension, float(p[0]), float(p[1]), float(p[2]), psize, pyoung,
                           float(q[0]), float(q[1]), float(q[2]), qsize, qyoung, ts, C.byref(d), C.byref(m))
    return r, d.value, m.value


def _grid_args(grid):
    lower = np.ascontiguousarray(grid["lower"], np.float64)
    lo = np.ascontiguousarray(grid["loweridx"], np.int64)
    hi = np.ascontiguousarray(grid["upperidx"], np.int64)
    ls = np.ascontiguousarray(grid["localsize"], np.int64)
    return lower, lo, hi, ls


def cic_deposit(colony: Colony, grid, weight_mode=0):
    lower, lo, hi, ls = _grid_args(grid)
    out = np.zeros(int(ls[0] * ls[1] * ls[2]), np.float64)
    lib().or_cic_deposit(C.byref(colony.c), weight_mode, lower, np.float32(grid["resolution"]), lo, hi, ls, out)
    return out


def cic_interp(colony: Colony, grid, field, gradient):
    lower, lo, hi, ls = _grid_args(grid)
    out = np.zeros(colony.n * 4, np.float64)
    lib().or_cic_interp(C.byref(colony.c), lower, np.float32(grid["resolution"]), lo, hi, ls,
                        np.ascontiguousarray(field, np.float64), np.ascontiguousarray(gradient, np.float64), out)
    return out.reshape(-1, 4)


def fieldgradient(localsize, resolution, field):
    ls = np.ascontiguousarray(localsize, np.int64)
    g = np.zeros(int(ls[0] * ls[1] * ls[2]) * 3, np.float64)
    lib().or_fieldgradient(ls, np.float32(resolution), np.ascontiguousarray(field, np.float64), g)
    return g


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


def philox_draws(n, dimension, seed, step):
    d = np.zeros(max(1, n * dimension), np.float32)
    lib().or_philox_draws(n, dimension, seed, step, d)
    return d[: n * dimension]


def updatepositions(x, y, z, size, ctype, randommove, fx, fy, fz, draws=None, box=(2048.0, 2048.0, 2048.0),
                    removecell=None, dimension=3):
    """src/cells.c:342-365 on copies of x, y, z (and removecell); draws None
    takes them from the C library's rand() like the reference.  Returns
    (x, y, z, removecell, newly_removed)."""
    x, y, z = (np.array(a, np.float64) for a in (x, y, z))
    n = len(x)
    rc = np.zeros(n, np.uint8) if removecell is None else np.array(removecell, np.uint8)
    dr = None if draws is None else np.ascontiguousarray(draws, np.float32)
    k = lib().or_updatepositions(n, dimension, x, y, z, np.ascontiguousarray(size, np.float64),
                                 np.ascontiguousarray(ctype, np.int32), np.ascontiguousarray(randommove, np.float32),
                                 np.ascontiguousarray(fx, np.float64), np.ascontiguousarray(fy, np.float64),
                                 np.ascontiguousarray(fz, np.float64),
                                 None if dr is None else dr.ctypes.data_as(C.c_void_p),
                                 np.ascontiguousarray(box, np.float64), rc.ctypes.data_as(C.c_void_p))
    return x, y, z, rc, int(k)


def _dparams(p):
    return (float(p["diffusion"]), float(p.get("lambda", 0.0)), float(p.get("boundary", 0.0)), float(p["dt"]))


def diffusion_rhs(ls, res, p, uold, src=None):
    ls = np.ascontiguousarray(ls, np.int64)
    D, lam, bc, dt = _dparams(p)
    b = np.zeros(int(np.prod(ls)))
    s = None if src is None else np.ascontiguousarray(src, np.float64)
    lib().or_diffusion_rhs(ls, np.float32(res), D, lam, bc, dt, None if s is None else s.ctypes.data_as(C.c_void_p),
                           np.ascontiguousarray(uold, np.float64).ravel(), b)
    return b


def diffusion_apply(ls, res, p, u):
    ls = np.ascontiguousarray(ls, np.int64)
    D, lam, bc, dt = _dparams(p)
    out = np.zeros(int(np.prod(ls)))
    lib().or_diffusion_apply(ls, np.float32(res), D, lam, dt, np.ascontiguousarray(u, np.float64).ravel(), out)
    return out


def diffusion_solve(ls, res, p, field, src=None, tol=1e-12, maxiter=50, substeps=1):
    """src/environment.c's backward-Euler step(s) solved by CG; returns (field, iterations)"""
    ls = np.ascontiguousarray(ls, np.int64)
    D, lam, bc, dt = _dparams(p)
    x = np.array(field, np.float64).ravel().copy()
    s = None if src is None else np.ascontiguousarray(src, np.float64)
    it = lib().or_diffusion_solve(ls, np.float32(res), D, lam, bc, dt,
                                  None if s is None else s.ctypes.data_as(C.c_void_p), x, tol, maxiter, substeps)
    return x, int(it)


def writevtk(path, x, y, z, ctype, force, rank=0):
    """src/io.c:372-516 for one rank; returns the file's bytes"""
    f = np.asarray(force, np.float64)
    a = lambda v: np.ascontiguousarray(v, np.float64)
    rc = lib().or_writevtk(os.fsencode(path), len(x), a(x), a(y), a(z), np.ascontiguousarray(ctype, np.int32),
                           a(f[:, 0]), a(f[:, 1]), a(f[:, 2]), int(rank))
    if rc != 0:
        raise OSError(f"cannot write {path}")
    return open(path, "rb").read()


def wtime():
    return lib().or_wtime()


def set_threads(n):
    lib().or_set_threads(int(n))


def max_threads():
    return lib().or_max_threads()
