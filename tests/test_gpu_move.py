"""GPU parity of the cell move (updatepositions, src/cells.c:342-365) against
the oracle: the resident form over the sorted colony, the plain-array form,
the reference-API updatepositions fed from rand(), and the captured mechanics
loop against the same steps run eagerly."""
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


def colony_with_strays():
    """20k-cell ball plus cells whose spheres reach the box faces"""
    col = gen.config(0)
    extra = np.array([[1023.2, 0.0, 0.0], [0.0, -1023.4, 0.0], [0.0, 0.0, 1022.0], [500.0, 500.0, 500.0]])
    for k, a in enumerate("xyz"):
        col[a] = np.concatenate([col[a], extra[:, k]])
    m = len(extra)
    col["size"] = np.concatenate([col["size"], np.ones(m)])
    col["young"] = np.concatenate([col["young"], np.full(m, 2100.0, np.float32)])
    col["ctype"] = np.concatenate([col["ctype"], np.zeros(m, np.int32)])
    col["phase"] = np.concatenate([col["phase"], np.ones(m, np.int32)])
    col["n"] = len(col["x"])
    return col


def oracle_move(O, col, F, rm, draws, removecell, dimension=3, box=(2048.0, 2048.0, 2048.0)):
    return O.updatepositions(col["x"], col["y"], col["z"], col["size"], col["ctype"], rm, F[:, 0], F[:, 1], F[:, 2],
                             draws, box=box, removecell=removecell, dimension=dimension)


@pytest.mark.parametrize("source", ["philox", "draws"])
def test_update_positions_resident(tm, oracle_mod, ctx, source):
    O = oracle_mod
    col = colony_with_strays()
    n = col["n"]
    say(f"resident move n={n} source={source}")
    ctx.upload_colony(col)
    ctx.step(tm.FP64)
    F = ctx.results()["force"]
    rm = [0.2]
    seed, step = 0x1234_5678_9ABC_DEF0, 41
    move = tm.Move.make(rm, seed=seed, step=step)
    pre = np.zeros(n, np.uint8)
    pre[n - 3] = 1  # already marked: not counted again
    rc = tm.DeviceArray.from_host(ctx, pre)
    if source == "philox":
        draws = O.philox_draws(n, 3, seed, step)
        k = ctx.update_positions(move, None, rc)
    else:
        draws = np.random.default_rng(5).random(3 * n, dtype=np.float32)
        d = tm.DeviceArray.from_host(ctx, draws)
        k = ctx.update_positions(move, d, rc)
        d.free()
    x, y, z = ctx.positions()
    ex, ey, ez, erc, ek = oracle_move(O, col, F, rm, draws, pre)
    assert np.array_equal(x, ex) and np.array_equal(y, ey) and np.array_equal(z, ez)
    # the stray at x = 1023.2 is newly marked, the one at y = -1023.4 was marked before
    assert np.array_equal(rc.get(), erc) and k == ek == 1 and erc[n - 4] == 1 and erc[n - 2] == 0
    rc.free()
    # the tree is stale until the next build; the next step sees the moved cells
    with pytest.raises(tm.TMError):
        ctx.results()
    ctx.step(tm.FP64)
    col2 = dict(col, x=ex, y=ey, z=ez)
    t2 = O.Octree(O.Colony.from_dict(col2), col2["types"])
    assert np.array_equal(ctx.results()["force"], t2.computegradient())


@pytest.mark.parametrize("dimension", [3, 2])
def test_move_cells_arrays(tm, oracle_mod, ctx, dimension):
    O = oracle_mod
    rng = np.random.default_rng(dimension)
    n = 100_003
    say(f"array move n={n} dim={dimension}")
    x, y, z = (rng.uniform(-1030, 1030, n) for _ in range(3))
    size = rng.choice([1.0, 1.5], n)
    ctype = (size > 1.2).astype(np.int32)
    F = rng.normal(0, 0.3, (n, 3))
    rm = [0.1, 0.35]
    draws = rng.random(dimension * n, dtype=np.float32)
    pre = (rng.random(n) < 0.01).astype(np.uint8)
    dev = {k: tm.DeviceArray.from_host(ctx, v) for k, v in
           dict(x=x, y=y, z=z, size=size, ctype=ctype, fx=np.ascontiguousarray(F[:, 0]),
                fy=np.ascontiguousarray(F[:, 1]), fz=np.ascontiguousarray(F[:, 2]), draws=draws, rc=pre).items()}
    k = ctx.move_cells(n, dimension, dev["x"], dev["y"], dev["z"], dev["size"], dev["ctype"], dev["fx"], dev["fy"],
                       dev["fz"], tm.Move.make(rm), dev["draws"], dev["rc"])
    col = dict(x=x, y=y, z=z, size=size, ctype=ctype)
    ex, ey, ez, erc, ek = oracle_move(O, col, F, rm, draws, pre, dimension)
    assert np.array_equal(dev["x"].get(), ex) and np.array_equal(dev["y"].get(), ey)
    assert np.array_equal(dev["z"].get(), ez)
    assert np.array_equal(dev["rc"].get(), erc) and k == ek and ek > 100
    # Philox draws without a draws array
    dev["x"].set(x)
    dev["y"].set(y)
    dev["z"].set(z)
    k = ctx.move_cells(n, dimension, dev["x"], dev["y"], dev["z"], dev["size"], dev["ctype"], dev["fx"], dev["fy"],
                       dev["fz"], tm.Move.make(rm, seed=99, step=3), None, None)
    ex, ey, ez, erc, ek = oracle_move(O, col, F, rm, O.philox_draws(n, dimension, 99, 3), None, dimension)
    assert np.array_equal(dev["x"].get(), ex) and np.array_equal(dev["z"].get(), ez) and k == ek
    # a cell type past the table is an error, not an out-of-range read
    bad = ctype.copy()
    bad[7] = 16
    dev["ctype"].set(bad)
    with pytest.raises(tm.TMError):
        ctx.move_cells(n, dimension, dev["x"], dev["y"], dev["z"], dev["size"], dev["ctype"], dev["fx"], dev["fy"],
                       dev["fz"], tm.Move.make(rm), None, None)
    for d in dev.values():
        d.free()


def test_loop_graph_matches_eager(tm, ctx):
    """each replay of the captured mechanics loop equals one eager step + move
    (bit-exact, FP64), with eager work and synchronisation between replays"""
    col = gen.config(1, scale=0.02)
    n = col["n"]
    move = tm.Move.make([0.05], seed=7, step=100)
    with tm.Context(0) as eager:
        eager.upload_colony(col)
        ctx.upload_colony(col)
        rc_g = tm.DeviceArray.from_host(ctx, np.zeros(n, np.uint8))
        ctx.graph_capture_loop(move, tm.FP64, rc_g)
        for rep in range(4):
            say(f"loop replay {rep}")
            ctx.graph_launch()
            eager.step(tm.FP64)
            F_e = eager.results()["force"]
            eager.update_positions(tm.Move.make([0.05], seed=7, step=100 + rep))
            assert np.array_equal(ctx.results()["force"], F_e), rep
            for a, b in zip(ctx.positions(), eager.positions()):
                assert np.array_equal(a, b), rep
        assert ctx.graph_removed() == 0
        rc_g.free()
        ctx.graph_destroy()


def test_loop_graph_list_growth(tm, ctx):
    """an attracting cell type (negative h4 scales the force by a negative
    constant) packs the colony, so neighbour counts outgrow the captured
    lists; the overflowing replay moves nothing, the step is redone eagerly
    and recaptured, and every step still equals the eager one"""
    col = gen.ball_colony(4000, 3)
    col["types"] = [dict(size=1.0, h=2.0, h2=4.0, h3=8.0, h4=-16.0)]
    move = tm.Move.make([0.01], seed=11, step=0)
    with tm.Context(0) as eager:
        eager.upload_colony(col)
        ctx.upload_colony(col)
        ctx.set_param("nbr_cap", 8)
        eager.set_param("nbr_cap", 8)
        ctx.graph_capture_loop(move, tm.FP64, None)
        alloc0 = ctx.get_param("nbr_alloc")
        grew = False
        for rep in range(8):
            say(f"attracting loop replay {rep}, lists {ctx.get_param('nbr_alloc')}")
            ctx.graph_launch()
            eager.step(tm.FP64)
            F_e = eager.results()["force"]
            eager.update_positions(tm.Move.make([0.01], seed=11, step=rep))
            assert np.array_equal(ctx.results()["force"], F_e), rep
            for a, b in zip(ctx.positions(), eager.positions()):
                assert np.array_equal(a, b), rep
            grew = grew or ctx.get_param("nbr_alloc") > alloc0
        ctx.graph_destroy()
    assert grew, "the lists never grew: the overflow path was not exercised"
