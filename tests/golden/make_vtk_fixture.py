"""Decode the reference-produced VTK sample into a JSON fixture.

Source (data files the reference ships, read as data, never executed):
  new_run/vtk/step00000000.vtk  big-endian binary VTK written by writevtk
                                (src/io.c:372-516) after step 0: POINTS are
                                positions after the move x += F
                                (src/cells.c:350-352), VECTORS force = F
  new_run/cells.inp             cell type "healthy": SIZE 5.0, H 10.0
  new_run/healthy.dat           three seed cells + 10 random (Box-Muller) cells

Cell-type quirk recorded with the fixture: H from cells.inp only updates h;
h2,h3,h4 keep the values initialcelltype computed from the default h=1.5
(src/initialisation.c:80-82 then src/io.c:222), so the force normalisation
48/(pi*h4) uses h4 = 1.5**4 while u = d/h uses h = 10.

Run in a container that has the reference at /root/reference; writes
tests/golden/vtk_step0.json (decoded positions and forces, and the file's
797 bytes as hex).
"""
import json
import os
import struct

REF = "/root/reference/new_run"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vtk_step0.json")


def main():
    data = open(os.path.join(REF, "vtk", "step00000000.vtk"), "rb").read()
    tag = b"POINTS 13 float\n"
    i = data.index(tag) + len(tag)
    n = 13
    pts = struct.unpack(">%df" % (3 * n), data[i:i + 12 * n])
    tag = b"VECTORS force float\n"
    j = data.index(tag) + len(tag)
    frc = struct.unpack(">%df" % (3 * n), data[j:j + 12 * n])
    h0 = 1.5
    fixture = {
        "source": "new_run/vtk/step00000000.vtk (reference writevtk output, step 0)",
        "n": n,
        "positions_after_move_f32": [list(pts[3 * k:3 * k + 3]) for k in range(n)],
        "force_f32": [list(frc[3 * k:3 * k + 3]) for k in range(n)],
        "celltype": {"size": 5.0, "h": 10.0, "h2": h0 * h0, "h3": h0 * h0 * h0, "h4": h0 * h0 * h0 * h0},
        "seed_cells": [[1.0, 1.0, 1.0], [5.0, 5.0, 50.0], [20.0, 20.0, 20.0]],
        "dimension": 3,
        # the whole file, for a byte-for-byte check of a VTK writer
        "file_hex": data.hex(),
    }
    with open(OUT, "w") as f:
        json.dump(fixture, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
