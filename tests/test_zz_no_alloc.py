"""fvm_step_noalloc (the step with no allocation: its table, time slot and
norm partials in the device's static scratch array) against fvm_step,
through the C ABI: bit for bit on two calls in a row, the second on a
smaller grid (the scratch keeps the first call's contents), with connected
blocks, CFL steps, an end time and fixed steps; sizes beyond the scratch and
unsupported faces are refused."""
import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


def uneven_blocks(ctx, seed=3):
    """27 connected blocks of uneven sizes (8-14 vertices) over a smooth state."""
    rng = np.random.default_rng(seed)
    sizes = [rng.integers(8, 15, 3) for _ in range(3)]
    ranges = []
    for d in range(3):
        lo, r = 0, []
        for s in sizes[d]:
            r.append((lo, lo + int(s) - 1))
            lo += int(s) - 1
        ranges.append(r)
    n = [r[-1][1] + 1 for r in ranges]
    p = cases.smooth_3d((n[2], n[1], n[0]), seed=12)
    blocks = {}
    for a, (x0, x1) in enumerate(ranges[0]):
        for b, (y0, y1) in enumerate(ranges[1]):
            for c, (z0, z1) in enumerate(ranges[2]):
                blk = ctx.block((x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1), (0.05, 0.06, 0.07))
                blk.upload(cases.to_conserved(p[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1]))
                blocks[(a, b, c)] = blk
    for (a, b, c), blk in blocks.items():
        idx = [a, b, c]
        for d in range(3):
            if idx[d] + 1 < 3:
                nb = list(idx)
                nb[d] += 1
                blk.connect(2 * d + 1, blocks[tuple(nb)], shared=True)
    return list(blocks.values())


def run_3d(noalloc, t_end=None):
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
        blks = uneven_blocks(ctx)
        if t_end is not None:
            ctx.set_time(0.0, t_end)
        step = ctx.step_noalloc if noalloc else ctx.step
        dts = [step(3, cfl=0.8), step(4, cfl=0.8)]
        norms, flag = ctx.norms()
        return [b.download() for b in blks], dts, norms, flag, ctx.time()


def run_2d(noalloc):
    p, dx = cases.riemann2d(41, seed=0)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        b = ctx.block((41, 41, 1), (dx, dx, 1.0))
        b.upload(cases.to_conserved(p))
        step = ctx.step_noalloc if noalloc else ctx.step
        dt = step(10, dt=2e-4)
        norms, flag = ctx.norms()
        return [b.download()], [dt], norms, flag, ctx.time()


def same(a, b):
    qa, dta, na, fa, ta = a
    qb, dtb, nb, fb, tb = b
    assert all(np.array_equal(x, y) for x, y in zip(qa, qb))
    assert dta == dtb and ta == tb and fa == fb
    assert np.array_equal(na, nb)


def test_noalloc_equals_step_two_calls():
    """27 connected blocks with CFL steps, then a smaller 2-D grid with fixed
    steps: each equals fvm_step bit for bit (state, dt, time, norms)."""
    print("case no-alloc 27-block CFL then riemann2d 41^2 fixed", flush=True)
    big = run_3d(True)
    small = run_2d(True)         # right after: the scratch still holds the first call's
    same(big, run_3d(False))
    same(small, run_2d(False))


def test_noalloc_end_time_halts_on_device():
    """With an end time the steps past it are skipped on the device; the
    norms are those of the last step that ran."""
    dt = run_3d(False)[1][0]
    t_end = 4.5 * dt             # reached in the second call, at its second step
    ref = run_3d(False, t_end=t_end)
    got = run_3d(True, t_end=t_end)
    same(got, ref)
    assert ref[4] == (t_end, 5) or (ref[4][1] == 5 and abs(ref[4][0] - t_end) < 1e-15)


def test_noalloc_refuses_beyond_its_scratch(monkeypatch):
    p, dx = cases.riemann2d(9, seed=0)
    q = cases.to_conserved(p)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        for _ in range(65):
            ctx.block((9, 9, 1), (dx, dx, 1.0)).upload(q)
        with pytest.raises(fvmgpu.FvmError, match="at most 64 blocks"):
            ctx.step_noalloc(1, dt=1e-4)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        b = ctx.block((9, 9, 1), (dx, dx, 1.0))
        b.upload(q)
        b.set_bc(0, fvmgpu.BC_WALL)
        with pytest.raises(fvmgpu.FvmError, match="patches"):
            ctx.step_noalloc(1, dt=1e-4)
        b.set_bc(0, fvmgpu.BC_EXTRAP)
    # a block with more stage tiles than the scratch holds (one-row chunks:
    # 49 x 3000 tiles of the 2-D row kernel)
    monkeypatch.setenv("FVM_CHUNK", "1")
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        b = ctx.block((3000, 3000, 1), (1e-3, 1e-3, 1.0))
        b.init(fvmgpu.IC_REGIONS, fvmgpu.region_params([((0.0, 1e9), (0.0, 1e9),
                                                          (1.0, 0.0, 0.0, 0.0, 1.0))]))
        with pytest.raises(fvmgpu.FvmError, match="65536 stage tiles"):
            ctx.step_noalloc(1, dt=1e-5)
        ctx.step(1, dt=1e-5)                 # the allocating step takes it
