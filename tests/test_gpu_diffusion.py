"""GPU environment solve (src/environment.c:131-511) against the oracle and
the closed forms: random fields, sources and boundaries over small and large
z, a Dirichlet sine eigenmode, a constant field at full size, and the CIC
deposit's production as the source (cells -> field)."""
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


def grid_of(ls, res, lower=(0.0, 0.0, 0.0)):
    return dict(lower=list(lower), resolution=np.float32(res), loweridx=[0, 0, 0],
                upperidx=[int(v) - 1 for v in ls], localsize=[int(v) for v in ls])


def z_of(D, dt, res):
    return D * dt / float(np.float32(res) * np.float32(res))


@pytest.mark.parametrize("z_target,lam,substeps", [(1.8e-5, 0.0, 1), (0.5, 0.01, 1), (20.0, 0.0, 2)])
def test_diffusion_matches_oracle(tm, oracle_mod, ctx, z_target, lam, substeps):
    O = oracle_mod
    ls = (24, 20, 28)
    res, dt = 2.0, 16000.0
    D = z_target * 4.0 / dt
    rng = np.random.default_rng(11)
    u0 = 0.1575 + 0.01 * rng.random(ls)
    src = np.zeros(ls)
    src[5, 5, 5] = 100.0  # the reference's point source
    src += 1e-3 * rng.random(ls)
    p = dict(diffusion=D, boundary=0.1575, dt=dt, **{"lambda": lam})
    say(f"diffusion z={z_of(D, dt, res):.3g} lambda={lam} substeps={substeps}")
    f = tm.DeviceArray.from_host(ctx, u0.ravel())
    s = tm.DeviceArray.from_host(ctx, src.ravel())
    it, rel = ctx.diffusion_step(grid_of(ls, res), tm.Diffusion.make(D, dt, 0.1575, lam, 1e-12, 500, substeps), f, s)
    got = f.get()
    ref, it_ref = O.diffusion_solve(ls, res, p, u0, src, tol=1e-12, maxiter=500, substeps=substeps)
    assert rel <= 1e-12 and abs(it - it_ref) <= max(2 * substeps, it_ref // 20)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    # the solution satisfies the system assembled by the oracle
    b = O.diffusion_rhs(ls, res, p, u0.ravel(), src) if substeps == 1 else None
    if b is not None:
        r = b - O.diffusion_apply(ls, res, p, got)
        assert np.linalg.norm(r) <= 2e-12 * np.linalg.norm(b)
    f.free()
    s.free()


def test_diffusion_eigenmode(tm, ctx):
    ls = (64, 48, 80)
    res = 1.0
    D, dt = 0.8, 1.0
    modes = (3, 2, 5)
    axes = [np.sin(np.pi * m * (np.arange(L) + 1) / (L + 1)) for m, L in zip(modes, ls)]
    u0 = axes[0][:, None, None] * axes[1][None, :, None] * axes[2][None, None, :]
    z = z_of(D, dt, res)
    mu = 1 + z * sum(2 * (1 - np.cos(np.pi * m / (L + 1))) for m, L in zip(modes, ls))
    f = tm.DeviceArray.from_host(ctx, u0.ravel())
    it, rel = ctx.diffusion_step(grid_of(ls, res), tm.Diffusion.make(D, dt, 0.0, 0.0, 1e-13, 200), f)
    assert np.abs(f.get().reshape(ls) - u0 / mu).max() <= 1e-12 and it <= 3
    f.free()


def test_diffusion_constant_full_size(tm, ctx):
    """256^3: a constant field with the boundary at its value is the solution
    (no iteration needed); with a lower boundary it decays only near the faces"""
    ls = (256, 256, 256)
    f = tm.DeviceArray.from_host(ctx, np.full(256 ** 3, 0.1575))
    it, rel = ctx.diffusion_step(grid_of(ls, 8.0), tm.Diffusion.make(1.82e-5, 16000.0, 0.1575), f)
    v = f.get()
    assert it == 0 and np.abs(v - 0.1575).max() <= 1e-15
    it, rel = ctx.diffusion_step(grid_of(ls, 8.0), tm.Diffusion.make(1.82e-5 * 1e3, 16000.0, 0.0, maxiter=200), f)
    v = f.get().reshape(ls)
    say(f"256^3 z={z_of(1.82e-2, 16000.0, 8.0):.3g}: {it} iterations, residual {rel:.2e}")
    assert rel <= 1e-12 and 0 < it < 200
    assert np.abs(v[64:192, 64:192, 64:192] - 0.1575).max() <= 1e-12  # far from the faces: unchanged
    assert v[0, 128, 128] < 0.1575 and v.min() >= 0.0
    f.free()


def test_diffusion_source_from_deposit(tm, oracle_mod, ctx):
    """the CIC deposit's production feeds the solve as its source"""
    O = oracle_mod
    col = gen.config(0)
    grid = gen.bbox_grid(col, 48)
    ls = grid["localsize"]
    ctx.upload_colony(col)
    ctx.build()
    prod = tm.DeviceArray(ctx, int(np.prod(ls)), np.float64)
    ctx.cic_deposit(grid, prod, 0)
    u0 = np.full(int(np.prod(ls)), 0.1575)
    f = tm.DeviceArray.from_host(ctx, u0)
    D, dt = 1.82e-5 * 200, 16000.0
    it, rel = ctx.diffusion_step(grid, tm.Diffusion.make(D, dt, 0.1575, 0.0, 1e-12, 200), f, prod)
    ref, it_ref = O.diffusion_solve(ls, grid["resolution"], dict(diffusion=D, boundary=0.1575, dt=dt, **{"lambda": 0.0}),
                                    u0, prod.get(), tol=1e-12, maxiter=200)
    assert np.abs(f.get() - ref).max() <= 1e-10 * np.abs(ref).max() and it > 0
    prod.free()
    f.free()
