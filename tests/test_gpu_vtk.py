"""GPU VTK writer (writevtk, src/io.c:372-516): the reference's own step-0
file byte for byte, and the resident cells against the oracle's writer,
before and after a move."""
import json
import os

import numpy as np
import pytest

from timothy_amd import generators as gen

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx(tm):
    c = tm.Context(0)
    yield c
    c.close()


def test_vtk_reference_file(tm, ctx, tmp_path):
    with open(os.path.join(GOLD, "vtk_step0.json")) as f:
        fx = json.load(f)
    n = fx["n"]
    pos = np.array(fx["positions_after_move_f32"], np.float64)
    frc = np.array(fx["force_f32"], np.float64)
    d = {k: tm.DeviceArray.from_host(ctx, np.ascontiguousarray(v)) for k, v in
         dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], t=np.zeros(n, np.int32), fx=frc[:, 0], fy=frc[:, 1],
              fz=frc[:, 2]).items()}
    path = str(tmp_path / "step.vtk")
    ctx.write_vtk_arrays(path, n, d["x"], d["y"], d["z"], d["t"], d["fx"], d["fy"], d["fz"], 0)
    assert open(path, "rb").read() == bytes.fromhex(fx["file_hex"])
    for v in d.values():
        v.free()


def test_vtk_resident_cells(tm, oracle_mod, ctx, tmp_path):
    O = oracle_mod
    col = gen.gaussian_colony(30000, 77, 146.6 * (30000 / 2e6) ** (1 / 3), frac_type1=0.05)
    ctx.upload_colony(col)
    ctx.step(tm.FP64)
    F = ctx.results()["force"]
    path = str(tmp_path / "a.vtk")
    ctx.write_vtk(path, rank=0)
    assert open(path, "rb").read() == O.writevtk(str(tmp_path / "r.vtk"), col["x"], col["y"], col["z"],
                                                 col["ctype"], F)
    # after a move, before the next build: moved positions with the step's forces
    ctx.update_positions(tm.Move.make([0.1, 0.2], seed=3, step=9))
    x, y, z = ctx.positions()
    ctx.write_vtk(path, rank=0)
    assert open(path, "rb").read() == O.writevtk(str(tmp_path / "r.vtk"), x, y, z, col["ctype"], F)
    # after the next build the forces no longer match the order: refused until the sums are redone
    ctx.build()
    with pytest.raises(tm.TMError):
        ctx.write_vtk(path)
