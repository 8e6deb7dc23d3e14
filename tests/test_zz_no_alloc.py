"""The allocation-free environment solve (tm_diffusion_step_noalloc) against
the allocating one (tm_diffusion_step) through the C ABI: identical kernels,
so the stepped fields, iteration counts and residuals must be identical, on
two calls in a row, the second on a smaller grid; the static work buffer
starts each call from whatever the last call left.  Grids above
TM_NOALLOC_MAX_NODES are refused, and the calls leave the device's free
memory unchanged."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def grid_of(lx, ly, lz, res=2.0):
    return dict(lower=np.zeros(3), resolution=np.float32(res), loweridx=np.zeros(3, np.int64),
                upperidx=np.array([lx - 1, ly - 1, lz - 1], np.int64), localsize=np.array([lx, ly, lz], np.int64),
                G=lx)


def device_free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free = C.c_size_t()
    total = C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_noalloc_solve_matches_allocating_solve(tm):
    rng = np.random.default_rng(12)
    with tm.Context(0) as ctx:
        # z = D dt / res^2 = 3: tens of CG iterations per step
        par = tm.Diffusion.make(3.0 * 4.0 / 100.0, 100.0, boundary=0.25, lam=1e-4, tol=1e-12, maxiter=200,
                                substeps=2)
        for shape in ((96, 112, 128), (40, 56, 72)):  # the second call smaller than the first
            grid = grid_of(*shape)
            nodes = int(np.prod(shape))
            u0 = rng.uniform(0.0, 1.0, nodes)
            src = rng.uniform(-0.01, 0.02, nodes)
            s_d = tm.DeviceArray.from_host(ctx, src)
            a = tm.DeviceArray.from_host(ctx, u0)
            b = tm.DeviceArray.from_host(ctx, u0)
            say(f"grid {shape}: allocating solve")
            ra = ctx.diffusion_step(grid, par, a, s_d)
            say(f"grid {shape}: allocation-free solve")
            rb = ctx.diffusion_step(grid, par, b, s_d, noalloc=True)
            say(f"  iterations {ra[0]} / {rb[0]}, residual {ra[1]:.3e} / {rb[1]:.3e}")
            assert ra[0] > 10 and ra == rb
            assert np.array_equal(a.get(), b.get())
            for d in (a, b, s_d):
                d.free()


def test_noalloc_refuses_large_grids_and_allocates_nothing(tm):
    with tm.Context(0) as ctx:
        par = tm.Diffusion.make(0.04, 100.0, boundary=0.1, tol=1e-10, maxiter=50)
        small = grid_of(32, 32, 32)
        f = tm.DeviceArray.from_host(ctx, np.full(32 ** 3, 0.5))
        ctx.diffusion_step(small, par, f, noalloc=True)  # first launch of each kernel
        ctx.sync()
        before = device_free_bytes()
        for shape in ((128, 128, 128), (64, 64, 64)):
            g = tm.DeviceArray.from_host(ctx, np.full(int(np.prod(shape)), 0.5))
            after_input = device_free_bytes()
            ctx.diffusion_step(grid_of(*shape), par, g, noalloc=True)
            ctx.sync()
            assert device_free_bytes() == after_input, shape
            g.free()
        assert tm.NOALLOC_MAX_NODES == 128 ** 3
        big = grid_of(129, 128, 128)
        g = tm.DeviceArray(ctx, 129 * 128 * 128, np.float64)
        with pytest.raises(tm.TMError) as e:
            ctx.diffusion_step(big, par, g, noalloc=True)
        assert e.value.status == 1 and "TM_NOALLOC_MAX_NODES" in str(e.value)
        g.free()
        f.free()
        say(f"free device memory {before} bytes before, unchanged by the allocation-free calls")
