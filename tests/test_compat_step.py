"""The reference C API served by libtimothy_compat.so, driven by a C program
(tests/compat/step_harness.c) in the order of the reference's step_compute
(src/step.c:41-100), against the CPU oracle.

Bars: affShift/affScale and the octree size equal the oracle's; the
reference's own DFS over the emitted cellsinfo->octree (octcomputebox,
octlocateregion, octheap*) visits the oracle's candidates in the oracle's
order; FP64 density, v, mindist and forces are bit-exact; the CIC deposit
(production of field 0) is within 1e-12 relative (atomic summation order),
the other fields' production is zero; the interpolated cellenvdata and the
field gradients are bit-exact; with TM_PRECISION=mixed the forces are within
1e-5 of FP64 (normalised by the largest force).
"""
import os
import subprocess

import numpy as np
import pytest

from timothy_amd import build as B
from timothy_amd import generators as gen

pytestmark = pytest.mark.gpu

NWALK = 400


def _write_input(path, col, fields, grads):
    grid = col["grid"]
    G = int(grid["G"])
    n = col["n"]
    with open(path, "wb") as f:
        np.array([n, len(col["types"]), G, len(fields), min(NWALK, n)], np.int64).tofile(f)
        for t in col["types"]:
            np.array([t["size"], t["h"], t["h2"], t["h3"], t["h4"]], np.float64).tofile(f)
        np.array(list(grid["lower"]) + [float(grid["resolution"])], np.float64).tofile(f)
        np.stack([col["x"], col["y"], col["z"], col["size"]], axis=1).astype(np.float64).tofile(f)
        col["young"].astype(np.float32).tofile(f)
        col["ctype"].astype(np.int32).tofile(f)
        col["phase"].astype(np.int32).tofile(f)
        for fld, grd in zip(fields, grads):
            fld.astype(np.float64).tofile(f)
            grd.astype(np.float64).tofile(f)


def _read_output(path, n, G, nf, nwalk):
    b = open(path, "rb").read()
    o = 0

    def take(dtype, count):
        nonlocal o
        a = np.frombuffer(b, dtype, count, o)
        o += a.nbytes
        return a

    head = take(np.float64, 5)
    cells = take(np.float64, 6 * n).reshape(n, 6)
    prod = take(np.float64, nf * G ** 3).reshape(nf, -1)
    ced = take(np.float64, nf * n * 4).reshape(nf, n, 4)
    grad = take(np.float64, nf * 3 * G ** 3).reshape(nf, -1)
    woff = take(np.int64, nwalk + 1)
    wc = take(np.int32, int(woff[-1]))
    assert o == len(b)
    return dict(shift=head[:3], scale=head[3], octsize=int(head[4]), cells=cells, prod=prod, ced=ced, grad=grad,
                woff=woff, wc=wc)


def _run(tmp_path, col, precision, move=None, prebuild=""):
    if not os.path.exists(B.HARNESS):
        pytest.fail(f"{B.HARNESS} is not built (run the package build first)")
    grid = col["grid"]
    G = int(grid["G"])
    fields = [gen.smooth_field(grid, 31 + k, np.float64) for k in range(2)]
    grads = [gen.smooth_field(grid, 41 + k, np.float64) for k in range(2)]
    grads = [np.concatenate([g, 0.5 * g, -g]) for g in grads]
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(inp, col, fields, grads)
    env = dict(os.environ, TM_PRECISION=precision, STEP_HARNESS_PREBUILD=prebuild)
    print(f"step_harness n={col['n']} G={G} precision={precision} prebuild={prebuild!r}", flush=True)
    args = [B.HARNESS, inp, out]
    if move is not None:
        args += [str(move[0]), repr(float(move[1])), str(tmp_path / "move.bin")]
    r = subprocess.run(args, env=env, capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    got = _read_output(out, col["n"], G, len(fields), min(NWALK, col["n"]))
    if move is not None:
        b = open(tmp_path / "move.bin", "rb").read()
        n = col["n"]
        xyz = np.frombuffer(b, np.float64, 3 * n).reshape(3, n)
        rc = np.frombuffer(b, np.uint8, n, 24 * n)
        tail = np.frombuffer(b, np.int64, 2, 25 * n)
        got.update(x=xyz[0], y=xyz[1], z=xyz[2], removecell=rc, removecount=int(tail[0]), next_rand=int(tail[1]))
    return got, fields, grads


CASES = [
    ("ball20k", lambda: gen.config(0)),
    ("gauss2types", lambda: gen.gaussian_colony(30000, 77, 146.6 * (30000 / 2e6) ** (1 / 3), frac_type1=0.05)),
]


# prebuild: the octree of a moved copy of the cells is built first and
# released ("free": the step reuses the layer's kept octree buffer) or held
# ("held": the step's octree gets a plain allocation)
@pytest.mark.parametrize("name,make,prebuild", [c + ("",) for c in CASES] + [CASES[0] + ("free",),
                                                                           CASES[0] + ("held",)])
def test_step_compute_order_fp64(oracle_mod, tmp_path, name, make, prebuild):
    O = oracle_mod
    col = make()
    if "grid" not in col:
        col["grid"] = gen.bbox_grid(col, 64)
    got, fields, grads = _run(tmp_path, col, "fp64", prebuild=prebuild)
    oc = O.Colony.from_dict(col)
    tree = O.Octree(oc, col["types"])
    # octbuild: affine map and node count
    assert np.array_equal(got["shift"], tree.shift) and got["scale"] == tree.scale
    assert got["octsize"] == tree.t.nnodes
    # the reference's traversal over the emitted octree
    coff, cand, _ = tree.candidates()
    nw = len(got["woff"]) - 1
    assert np.array_equal(got["woff"], coff[:nw + 1] - coff[0])
    assert np.array_equal(got["wc"], cand[:coff[nw]])
    # computepotential / computegradient
    dens, v, md = tree.computepotential()
    F = tree.computegradient()
    c = got["cells"]
    assert np.array_equal(c[:, 0], dens) and np.array_equal(c[:, 1], v) and np.array_equal(c[:, 2], md)
    assert np.array_equal(c[:, 3:6], F)
    # patches: deposit into field 0, zero elsewhere
    grid = col["grid"]
    ref = O.cic_deposit(oc, grid, 0)
    p0 = got["prod"][0]
    assert np.array_equal(p0 != 0, ref != 0)
    assert np.all(np.abs(p0 - ref) <= 1e-12 * np.abs(ref))
    assert np.all(got["prod"][1:] == 0)
    # interpolation of the fields and of the gradients handed in
    for k in range(len(fields)):
        assert np.array_equal(got["ced"][k], O.cic_interp(oc, grid, fields[k], grads[k]))
    # fieldgradient
    for k in range(len(fields)):
        assert np.array_equal(got["grad"][k], O.fieldgradient(grid["localsize"], grid["resolution"], fields[k]))
    print(f"  {name}: octree {got['octsize']} nodes, walk {len(got['wc'])} candidates, all equal", flush=True)


def test_step_compute_order_mixed(oracle_mod, tmp_path):
    O = oracle_mod
    col = gen.config(0)
    got, _, _ = _run(tmp_path, col, "mixed")
    tree = O.Octree(O.Colony.from_dict(col), col["types"])
    F = tree.computegradient()
    dens, v, md = tree.computepotential()
    c = got["cells"]
    assert np.array_equal(c[:, 2], md)
    assert np.abs(c[:, 0] - dens).max() <= 1e-5 * np.abs(dens).max()
    assert np.abs(c[:, 3:6] - F).max() <= 1e-5 * np.abs(F).max()


def test_updatepositions_after_step(oracle_mod, tmp_path):
    """cellsupdate's updatepositions (src/cells.c:342-365) through the
    reference API after the step: the random steps come from the program's
    rand() in the reference's order (the stream advances by 3 draws per
    cell), positions are bit-exact with the oracle fed the same rand()
    stream, and cells reaching the box faces are marked and counted"""
    import ctypes

    O = oracle_mod
    col = gen.config(0)
    extra = np.array([[1023.2, 0.0, 0.0], [0.0, -1023.4, 0.0], [0.0, 0.0, 1022.5]])
    for k, a in enumerate("xyz"):
        col[a] = np.concatenate([col[a], extra[:, k]])
    m = len(extra)
    col["size"] = np.concatenate([col["size"], np.ones(m)])
    col["young"] = np.concatenate([col["young"], np.full(m, 2100.0, np.float32)])
    col["ctype"] = np.concatenate([col["ctype"], np.zeros(m, np.int32)])
    col["phase"] = np.concatenate([col["phase"], np.ones(m, np.int32)])
    col["n"] = len(col["x"])
    col["grid"] = gen.bbox_grid(col, 64)
    got, _, _ = _run(tmp_path, col, "fp64", move=(2024, 0.3))
    F = got["cells"][:, 3:6]
    tree = O.Octree(O.Colony.from_dict(col), col["types"])
    assert np.array_equal(F, tree.computegradient())
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    libc.srand(2024)
    ex, ey, ez, erc, ek = O.updatepositions(col["x"], col["y"], col["z"], col["size"], col["ctype"], [0.3],
                                            F[:, 0], F[:, 1], F[:, 2], None)
    assert np.array_equal(got["x"], ex) and np.array_equal(got["y"], ey) and np.array_equal(got["z"], ez)
    assert np.array_equal(got["removecell"], erc) and got["removecount"] == ek == 2
    assert got["next_rand"] == libc.rand()
    # writevtk after the update (settings.step 1): moved positions, the step's forces
    vtk = open(tmp_path / "vtk" / "step00000001.vtk", "rb").read()
    assert vtk == O.writevtk(str(tmp_path / "ref.vtk"), ex, ey, ez, col["ctype"], F)


def test_step_compute_with_no_cells(oracle_mod, tmp_path):
    """a rank whose cells have all left (localcount.n == 0): octbuild returns
    without a tree (src/octree.c:152-153), the CIC deposit adds nothing, so
    every field's production is zero (src/patches.c:172 loops over no
    cells, src/patches.c:327-328 zeroes production first), the interpolation
    has no cells to fill, and the field gradients are still computed"""
    O = oracle_mod
    G = 16
    grid = dict(lower=np.array([-10.0, -12.0, -8.0]), resolution=np.float32(1.5), loweridx=np.zeros(3, np.int64),
                upperidx=np.full(3, G - 1, np.int64), localsize=np.full(3, G, np.int64), G=G)
    col = dict(n=0, x=np.zeros(0), y=np.zeros(0), z=np.zeros(0), size=np.zeros(0), young=np.zeros(0, np.float32),
               ctype=np.zeros(0, np.int32), phase=np.zeros(0, np.int32), types=gen.CELLTYPES, grid=grid)
    got, fields, grads = _run(tmp_path, col, "fp64")
    assert got["octsize"] == 0 and got["cells"].shape == (0, 6) and got["ced"].shape == (2, 0, 4)
    assert np.all(got["prod"] == 0)
    for k in range(len(fields)):
        assert np.array_equal(got["grad"][k], O.fieldgradient(grid["localsize"], grid["resolution"], fields[k]))


def test_step_grid_copies_span_several_chunks(oracle_mod, tmp_path):
    """A 168^3 grid (38 MB per field, 114 MB per gradient): the layer's copies
    from and to the caller's arrays run through its 32 MB page-locked chunks
    several times each; production, the interpolated values and the field
    gradient still equal the oracle's."""
    O = oracle_mod
    col = gen.config(0, scale=0.25)
    col["grid"] = gen.bbox_grid(col, 168)
    got, fields, grads = _run(tmp_path, col, "fp64")
    grid = col["grid"]
    oc = O.Colony.from_dict(col)
    ref = O.cic_deposit(oc, grid, 0)
    p0 = got["prod"][0]
    assert np.array_equal(p0 != 0, ref != 0)
    assert np.all(np.abs(p0 - ref) <= 1e-12 * np.abs(ref))
    assert np.all(got["prod"][1:] == 0)
    for k in range(len(fields)):
        assert np.array_equal(got["ced"][k], O.cic_interp(oc, grid, fields[k], grads[k]))
        assert np.array_equal(got["grad"][k], O.fieldgradient(grid["localsize"], grid["resolution"], fields[k]))
