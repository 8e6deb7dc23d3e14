"""timothy_amd -- MI355X-native cell-mechanics path of the Timothy tissue
simulator (octree neighbour search, SPH density/potential/force, cloud-in-cell
transfers), driven through the C ABI of libtimothy_hip.so.

The library is built in-tree by timothy_amd/build.py.  Importing this package
without the built library raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtimothy_hip.so")
COMPAT_PATH = os.path.join(_HERE, "libtimothy_compat.so")

TM_OK = 0
STATUS = {0: "TM_OK", 1: "TM_ERR_INVALID", 2: "TM_ERR_HIP", 3: "TM_ERR_NOMEM", 4: "TM_ERR_DEGENERATE",
          5: "TM_ERR_STATE", 6: "TM_ERR_UNSUPPORTED"}
FP64, MIXED = 0, 1
NOALLOC_MAX_NODES = 2097152  # TM_NOALLOC_MAX_NODES
STAGE_BUILD, STAGE_NEIGHBORS, STAGE_SPH, STAGE_CIC, STAGE_STEP, STAGE_TILES = 0, 1, 2, 3, 4, 5


class TMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class CellType(C.Structure):
    _fields_ = [("size", C.c_double), ("h", C.c_double), ("h2", C.c_double), ("h3", C.c_double),
                ("h4", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("lower", C.c_double * 3), ("resolution", C.c_float), ("loweridx", C.c_int64 * 3),
                ("upperidx", C.c_int64 * 3), ("localsize", C.c_int64 * 3)]

    @classmethod
    def from_dict(cls, g):
        o = cls()
        o.lower[:] = [float(v) for v in g["lower"]]
        o.resolution = float(np.float32(g["resolution"]))
        o.loweridx[:] = [int(v) for v in g["loweridx"]]
        o.upperidx[:] = [int(v) for v in g["upperidx"]]
        o.localsize[:] = [int(v) for v in g["localsize"]]
        return o


class Move(C.Structure):
    """tm_move: randommove per cell type, box edges, Philox key and step"""
    _fields_ = [("randommove", C.c_float * 16), ("box", C.c_double * 3), ("seed", C.c_uint64),
                ("step", C.c_uint64)]

    @classmethod
    def make(cls, randommove, box=(2048.0, 2048.0, 2048.0), seed=0, step=0):
        o = cls()
        rm = list(randommove) + [0.0] * (16 - len(randommove))
        o.randommove[:] = [float(np.float32(v)) for v in rm]
        o.box[:] = [float(v) for v in box]
        o.seed = int(seed)
        o.step = int(step)
        return o


class Diffusion(C.Structure):
    """tm_diffusion: the environment solve's parameters"""
    _fields_ = [("diffusion", C.c_double), ("lambda_", C.c_double), ("boundary", C.c_double), ("dt", C.c_double),
                ("tol", C.c_double), ("maxiter", C.c_int32), ("substeps", C.c_int32)]

    @classmethod
    def make(cls, diffusion, dt, boundary=0.0, lam=0.0, tol=1e-12, maxiter=50, substeps=1):
        return cls(float(diffusion), float(lam), float(boundary), float(dt), float(tol), int(maxiter),
                   int(substeps))


_lib = None
_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m timothy_amd.build` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "tm_create": (_i32, [C.POINTER(_vp), C.c_int]),
        "tm_destroy": (None, [_vp]),
        "tm_status_name": (C.c_char_p, [C.c_int]),
        "tm_last_error": (C.c_char_p, [_vp]),
        "tm_stream": (_vp, [_vp]),
        "tm_sync": (_i32, [_vp]),
        "tm_set_celltypes": (_i32, [_vp, C.POINTER(CellType), _i32, _i32]),
        "tm_set_param": (_i32, [_vp, C.c_char_p, _i64]),
        "tm_get_param": (_i32, [_vp, C.c_char_p, C.POINTER(_i64)]),
        "tm_upload": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "tm_bind_device": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "tm_build": (_i32, [_vp]),
        "tm_neighbors": (_i32, [_vp]),
        "tm_sph": (_i32, [_vp, C.c_int]),
        "tm_step": (_i32, [_vp, C.c_int]),
        "tm_get_results": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "tm_device_results": (_i32, [_vp] + [C.POINTER(_vp)] * 7),
        "tm_get_affine": (_i32, [_vp, _i32, _vp, C.POINTER(C.c_double)]),
        "tm_get_codes": (_i32, [_vp, _vp]),
        "tm_get_order": (_i32, [_vp, _vp, _vp]),
        "tm_get_neighbors": (_i32, [_vp, _vp, _vp, C.POINTER(_i64)]),
        "tm_get_qstar": (_i32, [_vp, _vp]),
        "tm_pair_count": (_i32, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
        "tm_get_octree": (_i32, [_vp, _i32, _vp, C.POINTER(_i64)]),
        "tm_cic_deposit": (_i32, [_vp, C.POINTER(Grid), _i32, _vp]),
        "tm_field_gradient": (_i32, [_vp, C.POINTER(Grid), _vp, _vp]),
        "tm_field_gradient_f32": (_i32, [_vp, C.POINTER(Grid), _vp, _vp]),
        "tm_cic_interp": (_i32, [_vp, C.POINTER(Grid), _vp, _vp, _vp]),
        "tm_cic_interp_f32": (_i32, [_vp, C.POINTER(Grid), _vp, _vp, _vp]),
        "tm_malloc": (_i32, [_vp, C.POINTER(_vp), C.c_size_t]),
        "tm_free": (_i32, [_vp, _vp]),
        "tm_memcpy_htod": (_i32, [_vp, _vp, _vp, C.c_size_t]),
        "tm_memcpy_dtoh": (_i32, [_vp, _vp, _vp, C.c_size_t]),
        "tm_memset": (_i32, [_vp, _vp, C.c_int, C.c_size_t]),
        "tm_host_alloc": (_i32, [_vp, C.POINTER(_vp), C.c_size_t]),
        "tm_memcpy_htod_async": (_i32, [_vp, _vp, _vp, C.c_size_t]),
        "tm_memcpy_dtoh_async": (_i32, [_vp, _vp, _vp, C.c_size_t]),
        "tm_host_free": (_i32, [_vp, _vp]),
        "tm_memcpy_dtod": (_i32, [_vp, _vp, _vp, C.c_size_t]),
        "tm_time_stage": (_i32, [_vp, C.c_int, C.c_int, _i32, _i32, C.POINTER(C.c_double)]),
        "tm_last_kernel_ms": (_i32, [_vp, C.POINTER(C.c_double)]),
        "tm_graph_capture": (_i32, [_vp, C.c_int]),
        "tm_graph_launch": (_i32, [_vp]),
        "tm_graph_destroy": (_i32, [_vp]),
        "tm_autotune": (_i32, [_vp, C.c_int, C.c_char_p, C.c_size_t]),
        "tm_update_positions": (_i32, [_vp, C.POINTER(Move), _vp, _vp, C.POINTER(_i64)]),
        "tm_move_cells": (_i32, [_vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(Move), _vp, _vp,
                                 C.POINTER(_i64)]),
        "tm_get_positions": (_i32, [_vp, _vp, _vp, _vp]),
        "tm_graph_capture_loop": (_i32, [_vp, C.c_int, C.POINTER(Move), _vp]),
        "tm_graph_removed": (_i32, [_vp, C.POINTER(_i64)]),
        "tm_diffusion_step": (_i32, [_vp, C.POINTER(Grid), C.POINTER(Diffusion), _vp, _vp, C.POINTER(_i32),
                                     C.POINTER(C.c_double)]),
        "tm_diffusion_step_noalloc": (_i32, [_vp, C.POINTER(Grid), C.POINTER(Diffusion), _vp, _vp,
                                             C.POINTER(_i32), C.POINTER(C.c_double)]),
        "tm_write_vtk": (_i32, [_vp, C.c_char_p, _i32]),
        "tm_write_vtk_arrays": (_i32, [_vp, C.c_char_p, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def lib():
    return _load()


def _ptr(a):
    return None if a is None else a.ctypes.data


class DeviceArray:
    """A device allocation owned through the context (tm_malloc / tm_free)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        ctx._check(lib().tm_malloc(ctx.h, C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(ctx, arr.shape, arr.dtype)
        ctx._check(lib().tm_memcpy_htod(ctx.h, d.ptr, arr.ctypes.data, arr.nbytes))
        return d

    def get(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(lib().tm_memcpy_dtoh(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def set(self, arr):
        arr = np.ascontiguousarray(arr, self.dtype)
        assert arr.nbytes == self.nbytes
        self.ctx._check(lib().tm_memcpy_htod(self.ctx.h, self.ptr, arr.ctypes.data, self.nbytes))

    def zero(self):
        self.ctx._check(lib().tm_memset(self.ctx.h, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().tm_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class Context:
    """One tm_ctx: a HIP stream on one device with its colonies and buffers."""

    def __init__(self, device: int = 0):
        L = lib()
        h = _vp()
        s = L.tm_create(C.byref(h), device)
        if s != TM_OK:
            raise TMError(s, "tm_create failed (no HIP device?)")
        self.h = h.value
        self.n = 0
        self.ncol = 0

    def close(self):
        if getattr(self, "h", None):
            lib().tm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, s):
        if s != TM_OK:
            msg = lib().tm_last_error(self.h)
            raise TMError(s, msg.decode() if msg else "")

    # ---- configuration
    def set_celltypes(self, types, dimension=3):
        arr = (CellType * len(types))()
        for i, t in enumerate(types):
            if isinstance(t, dict):
                t = (t["size"], t["h"], t["h2"], t["h3"], t["h4"])
            arr[i] = CellType(*[float(v) for v in t])
        self._check(lib().tm_set_celltypes(self.h, arr, len(types), dimension))

    def set_param(self, key, value):
        self._check(lib().tm_set_param(self.h, key.encode(), int(value)))

    def get_param(self, key):
        v = _i64()
        self._check(lib().tm_get_param(self.h, key.encode(), C.byref(v)))
        return v.value

    # ---- cells
    def upload(self, x, y, z, size, young=None, ctype=None, phase=None, colony_offsets=None):
        n = len(x)
        self._keep = [np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64),
                      np.ascontiguousarray(z, np.float64), np.ascontiguousarray(size, np.float64),
                      None if young is None else np.ascontiguousarray(young, np.float32),
                      np.ascontiguousarray(np.zeros(n) if ctype is None else ctype, np.int32),
                      None if phase is None else np.ascontiguousarray(phase, np.int32)]
        ncol = 1 if colony_offsets is None else len(colony_offsets) - 1
        off = None if colony_offsets is None else np.ascontiguousarray(colony_offsets, np.int64)
        self._check(lib().tm_upload(self.h, n, ncol, _ptr(off), *[_ptr(a) for a in self._keep]))
        self._keep = None
        self.n = n
        self.ncol = ncol

    def upload_colony(self, col):
        self.set_celltypes(col.get("types"), col.get("dimension", 3))
        self.upload(col["x"], col["y"], col["z"], col["size"], col.get("young"), col["ctype"], col.get("phase"))

    def upload_colonies(self, cols):
        self.set_celltypes(cols[0]["types"], 3)
        off = np.cumsum([0] + [c["n"] for c in cols]).astype(np.int64)
        cat = lambda k, dt: np.concatenate([np.asarray(c[k], dt) for c in cols])
        self.upload(cat("x", np.float64), cat("y", np.float64), cat("z", np.float64), cat("size", np.float64),
                    cat("young", np.float32), cat("ctype", np.int32), cat("phase", np.int32), colony_offsets=off)
        return off

    def bind_device(self, n, x, y, z, size, young, ctype, phase=None, colony_offsets=None):
        ncol = 1 if colony_offsets is None else len(colony_offsets) - 1
        off = None if colony_offsets is None else np.ascontiguousarray(colony_offsets, np.int64)
        ptrs = [a if isinstance(a, (int, type(None))) else a.ptr for a in (x, y, z, size, young, ctype, phase)]
        self._check(lib().tm_bind_device(self.h, n, ncol, _ptr(off), *ptrs))
        self.n = n
        self.ncol = ncol

    # ---- stages
    def build(self):
        self._check(lib().tm_build(self.h))

    def neighbors(self):
        self._check(lib().tm_neighbors(self.h))

    def sph(self, precision=FP64):
        self._check(lib().tm_sph(self.h, precision))

    def step(self, precision=FP64):
        self._check(lib().tm_step(self.h, precision))

    def sync(self):
        self._check(lib().tm_sync(self.h))

    # ---- results
    def results(self):
        n = self.n
        out = [np.empty(n) for _ in range(6)]
        self._check(lib().tm_get_results(self.h, *[o.ctypes.data for o in out]))
        return dict(density=out[0], v=out[1], mindist=out[2], force=np.stack(out[3:], axis=1))

    def affine(self, colony=0):
        sh = np.zeros(3)
        sc = C.c_double()
        self._check(lib().tm_get_affine(self.h, colony, sh.ctypes.data, C.byref(sc)))
        return sh, sc.value

    def codes(self):
        loc = np.empty((self.n, 3), np.uint32)
        self._check(lib().tm_get_codes(self.h, loc.ctypes.data))
        return loc

    def order(self):
        perm = np.empty(self.n, np.int32)
        span = np.empty(self.n, np.int32)
        self._check(lib().tm_get_order(self.h, perm.ctypes.data, span.ctypes.data))
        return perm, span

    def neighbor_lists(self):
        off = np.empty(self.n + 1, np.int64)
        tot = _i64()
        self._check(lib().tm_get_neighbors(self.h, off.ctypes.data, None, C.byref(tot)))
        nbr = np.empty(max(tot.value, 1), np.int32)
        self._check(lib().tm_get_neighbors(self.h, off.ctypes.data, nbr.ctypes.data, C.byref(tot)))
        return off, nbr[:tot.value]

    def qstar(self):
        q = np.empty(self.n, np.int32)
        self._check(lib().tm_get_qstar(self.h, q.ctypes.data))
        return q

    def pair_count(self):
        p = _i64()
        m = _i64()
        self._check(lib().tm_pair_count(self.h, C.byref(p), C.byref(m)))
        return p.value, m.value

    def octree(self, colony=0):
        cnt = _i64()
        self._check(lib().tm_get_octree(self.h, colony, None, C.byref(cnt)))
        dt = np.dtype([("code", np.uint32, 3), ("limit", np.uint32, 3), ("level", np.uint32), ("pad", np.uint32),
                       ("father", np.int64), ("child", np.int64, 8), ("data", np.int32), ("pad2", np.int32)])
        assert dt.itemsize == 112
        out = np.zeros(cnt.value, dt)
        if cnt.value:
            self._check(lib().tm_get_octree(self.h, colony, out.ctypes.data, C.byref(cnt)))
        return out

    # ---- grid transfers (device arrays)
    def cic_deposit(self, grid, production: DeviceArray, weight_mode=0):
        g = Grid.from_dict(grid)
        self._check(lib().tm_cic_deposit(self.h, C.byref(g), weight_mode, production.ptr))

    def field_gradient(self, grid, field: DeviceArray, gradient: DeviceArray):
        g = Grid.from_dict(grid)
        f = lib().tm_field_gradient_f32 if field.dtype == np.float32 else lib().tm_field_gradient
        self._check(f(self.h, C.byref(g), field.ptr, gradient.ptr))

    def cic_interp(self, grid, field: DeviceArray, gradient: DeviceArray, out: DeviceArray):
        g = Grid.from_dict(grid)
        f = lib().tm_cic_interp_f32 if field.dtype == np.float32 else lib().tm_cic_interp
        self._check(f(self.h, C.byref(g), field.ptr, gradient.ptr, out.ptr))

    def diffusion_step(self, grid, params: Diffusion, field: DeviceArray, source: DeviceArray = None,
                       noalloc: bool = False):
        """tm_diffusion_step (or, with noalloc, tm_diffusion_step_noalloc: work
        arrays in the library's static device buffer, grids of at most
        NOALLOC_MAX_NODES nodes): the environment solve on the device field;
        returns (CG iterations, final relative residual)"""
        g = Grid.from_dict(grid)
        it = _i32()
        res = C.c_double()
        f = lib().tm_diffusion_step_noalloc if noalloc else lib().tm_diffusion_step
        self._check(f(self.h, C.byref(g), C.byref(params), None if source is None else source.ptr, field.ptr,
                      C.byref(it), C.byref(res)))
        return it.value, res.value

    def write_vtk(self, path, rank=0):
        self._check(lib().tm_write_vtk(self.h, os.fsencode(path), rank))

    def write_vtk_arrays(self, path, n, x, y, z, ctype, fx, fy, fz, rank=0):
        self._check(lib().tm_write_vtk_arrays(self.h, os.fsencode(path), n, x.ptr, y.ptr, z.ptr, ctype.ptr, fx.ptr,
                                              fy.ptr, fz.ptr, rank))

    # ---- measurement
    def time_stage(self, stage, precision=FP64, warmup=2, iters=10):
        ms = C.c_double()
        self._check(lib().tm_time_stage(self.h, stage, precision, warmup, iters, C.byref(ms)))
        return ms.value

    def graph_capture(self, precision=FP64):
        self._check(lib().tm_graph_capture(self.h, precision))

    def graph_capture_loop(self, move: Move, precision=MIXED, removecell: "DeviceArray" = None):
        self._check(lib().tm_graph_capture_loop(self.h, precision, C.byref(move),
                                                None if removecell is None else removecell.ptr))

    def graph_removed(self):
        k = _i64()
        self._check(lib().tm_graph_removed(self.h, C.byref(k)))
        return k.value

    # ---- moving the cells
    def update_positions(self, move: Move, draws: "DeviceArray" = None, removecell: "DeviceArray" = None):
        """tm_update_positions; returns the number of cells newly marked"""
        k = _i64(0)
        self._check(lib().tm_update_positions(self.h, C.byref(move), None if draws is None else draws.ptr,
                                              None if removecell is None else removecell.ptr, C.byref(k)))
        return k.value

    def move_cells(self, n, dimension, x, y, z, size, ctype, fx, fy, fz, move: Move, draws=None, removecell=None):
        """tm_move_cells on device arrays; returns the number of cells newly marked"""
        k = _i64(0)
        p = lambda a: None if a is None else a.ptr
        self._check(lib().tm_move_cells(self.h, n, dimension, p(x), p(y), p(z), p(size), p(ctype), p(fx), p(fy),
                                        p(fz), C.byref(move), p(draws), p(removecell), C.byref(k)))
        return k.value

    def positions(self):
        out = [np.empty(self.n) for _ in range(3)]
        self._check(lib().tm_get_positions(self.h, *[o.ctypes.data for o in out]))
        return out

    def graph_launch(self):
        self._check(lib().tm_graph_launch(self.h))

    def graph_destroy(self):
        self._check(lib().tm_graph_destroy(self.h))

    def autotune(self, precision=MIXED):
        """runs tm_autotune; returns {"trials": [...], "chosen": {...}, "text": report}
        (values parsed as int or float where they are numbers)"""
        buf = C.create_string_buffer(16384)
        self._check(lib().tm_autotune(self.h, precision, buf, 16384))
        return parse_autotune(buf.value.decode())


def parse_autotune(text):
    """the tm_autotune report as {"trials": [{"stage": .., key: value, ..}],
    "chosen": {key: value}, "text": text}"""
    def fields(words):
        d = {}
        for w in words:
            k, _, v = w.partition("=")
            try:
                d[k] = int(v)
            except ValueError:
                try:
                    d[k] = float(v)
                except ValueError:
                    d[k] = v
        return d

    out = {"trials": [], "chosen": None, "text": text}
    for line in text.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "trial":
            out["trials"].append(dict(stage=w[1], **fields(w[2:])))
        elif w[0] == "chosen":
            out["chosen"] = fields(w[1:])
        else:
            raise ValueError(f"unparsed autotune line: {line!r}")
    return out


def header_functions(path=None):
    """names of the functions declared in a public header (for ABI checks)"""
    import re

    path = path or os.path.join(os.path.dirname(_HERE), "include", "timothy_hip.h")
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tm_[a-z0-9_]+)\s*\(", txt)))
