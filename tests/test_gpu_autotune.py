"""The autotuner and the tile width.

- tm_autotune on a 30k-cell colony: a parseable report that covers the walk
  shapes, the FP32 screening and every tile width; the shapes it keeps leave
  the FP64 sums bit-identical to the oracle (src/potential.c:87-259) and the
  mixed sums within their bound.  It replaces the reference's only knob, the
  OpenMP schedule of the walks (src/potential.c:90-93).
- Every tile width (128, 256, 512 cells) gives the same lists, bit-exact FP64
  sums and bounded mixed sums; a captured graph is retired when the width
  changes.
"""
import numpy as np
import pytest

from timothy_amd import generators as gen

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx(tm):
    c = tm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def case(oracle_mod):
    O = oracle_mod
    col = gen.gaussian_colony(30000, 77, 146.6 * (30000 / 2e6) ** (1 / 3), frac_type1=0.05)
    tree = O.Octree(O.Colony.from_dict(col), col["types"])
    dens, v, md = tree.computepotential()
    F = tree.computegradient()
    return col, dict(density=dens, v=v, mindist=md, force=F)


def check_fp64(r, ref):
    for k in ("density", "v", "mindist", "force"):
        assert np.array_equal(r[k], ref[k]), k


def check_mixed(r, ref):
    assert np.array_equal(r["mindist"], ref["mindist"])
    assert np.abs(r["force"] - ref["force"]).max() <= 1e-5 * np.abs(ref["force"]).max()
    assert np.abs(r["density"] - ref["density"]).max() <= 1e-5 * np.abs(ref["density"]).max()


def test_tile_widths(tm, ctx, case):
    col, ref = case
    ctx.upload_colony(col)
    ctx.set_param("sph_tiles", 1)
    lists = None
    for tw in (128, 512, 256):
        ctx.set_param("tile_width", tw)
        assert ctx.get_param("tile_width") == tw
        ctx.build()
        ctx.neighbors()
        got = ctx.neighbor_lists()
        if lists is not None:
            assert np.array_equal(got[0], lists[0]) and np.array_equal(got[1], lists[1])
        lists = got
        say(f"tile_width {tw}: staged records per tile {ctx.get_param('tile_lds')}, "
            f"over-full tiles {ctx.get_param('tiles_over')}")
        for u in (2, 4, 8):
            ctx.set_param("sph_unroll_fp64", u)
            ctx.sph(tm.FP64)
            check_fp64(ctx.results(), ref)
            ctx.set_param("sph_unroll_mixed", u)
            ctx.sph(tm.MIXED)
            check_mixed(ctx.results(), ref)
    with pytest.raises(tm.TMError):
        ctx.set_param("tile_width", 64)
    # a graph captured at one width is retired when the width changes
    ctx.graph_capture(tm.FP64)
    ctx.graph_launch()
    check_fp64(ctx.results(), ref)
    ctx.set_param("tile_width", 128)
    with pytest.raises(tm.TMError) as e:
        ctx.graph_launch()
    assert e.value.status == 5  # TM_ERR_STATE: no captured graph
    ctx.graph_capture(tm.FP64)
    ctx.graph_launch()
    check_fp64(ctx.results(), ref)
    ctx.graph_destroy()
    ctx.set_param("tile_width", 256)
    ctx.set_param("sph_unroll_fp64", 4)
    ctx.set_param("sph_unroll_mixed", 4)


def test_autotune_report_and_results(tm, ctx, case):
    col, ref = case
    ctx.upload_colony(col)
    for prec in (tm.FP64, tm.MIXED):
        ctx.step(prec)
        rep = ctx.autotune(prec)
        say(rep["text"].splitlines()[-1])
        stages = {t["stage"] for t in rep["trials"]}
        assert {"neighbors", "tiles_sph"} <= stages
        assert {t["tile_width"] for t in rep["trials"] if t["stage"] == "tiles_sph"} == {128, 256, 512}
        assert all(t["ms"] > 0 for t in rep["trials"])
        ch = rep["chosen"]
        assert ch["precision"] == ("fp64" if prec == tm.FP64 else "mixed")
        assert {t.get("walk_coop") for t in rep["trials"] if "walk_coop" in t} == {0, 1}
        for k in ("nbr_block", "walk_window", "walk_coop", "walk_f32", "tile_width"):
            assert ctx.get_param(k) == ch[k], k
        # the kept tile shape is the fastest, or the starting one within 2% of it
        ts = [t for t in rep["trials"] if t["stage"] == "tiles_sph"]
        kept = [t for t in ts if (t["tile_width"], t["unroll"]) == (ch["tile_width"], ch["unroll"])]
        assert len(kept) == 1 and kept[0]["ms"] <= min(t["ms"] for t in ts) * 1.02 + 1e-9
        # the kept shapes change no result
        ctx.sph(tm.FP64)
        check_fp64(ctx.results(), ref)
        ctx.sph(tm.MIXED)
        check_mixed(ctx.results(), ref)
        ctx.step(tm.FP64)
        check_fp64(ctx.results(), ref)
