"""Initial states computed on the device (fvm_block_init) against the host
generators the other tests upload (fvmgpu.cases).

* regions (the reference's init_value, src/base/defs_block_data.f90:166-201,
  and its four-quadrant case, src/solve/class_flow_field.f90:207-243): only
  comparisons, products and sums, so bit for bit, before and after steps;
* vortex and Taylor-Green: the device's exp / pow / sin / cos against
  numpy's, so to a few ulp (relative 1e-14 of each component's scale).
"""
import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [0, 3])
def test_regions_riemann_bitwise(seed):
    n = 41
    p, dx = cases.riemann2d(n, seed=seed)
    if seed == 0:
        quads = cases.REF_QUADRANTS
    else:   # the same draws as cases.riemann2d
        rng = np.random.default_rng(seed)
        quads = []
        for xr, yr in (((0.8, 1.0), (0.8, 1.0)), ((0.0, 0.8), (0.8, 1.0)),
                       ((0.0, 0.8), (0.0, 0.8)), ((0.8, 1.0), (0.0, 0.8))):
            rho = rng.uniform(0.1, 2.0)
            pr = rng.uniform(0.02, 1.5)
            u, v = rng.uniform(-1.3, 1.3, 2)
            quads.append((xr, yr, (rho, u, v, 0.0, pr)))
    print(f"case regions riemann2d seed={seed}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        ctx.numerics(0)
        a = ctx.block((n, n, 1), (dx, dx, 1.0))
        a.init(fvmgpu.IC_REGIONS, fvmgpu.region_params(quads))
        b = ctx.block((n, n, 1), (dx, dx, 1.0))
        b.upload(cases.to_conserved(p))
        assert np.array_equal(a.download(), cases.to_conserved(p))
        ctx.step(5, cfl=0.5)
        assert np.array_equal(a.download(), b.download())


def test_regions_3d_boxes_and_origin():
    n = (20, 14, 10)
    h = (0.1, 0.2, 0.3)
    origin = (-1.0, 0.5, 2.0)
    boxes = [((-9, 9), (-9, 9), (-9, 9), (1.0, 0.1, 0.2, 0.3, 1.0)),
             ((-1.0, -0.3), (0.5, 1.5), (2.0, 2.9), (0.5, -0.2, 0.0, 0.1, 0.4)),
             ((0.0, 0.5), (1.0, 2.0), (2.3, 3.0), (2.0, 0.0, 0.3, 0.0, 1.8))]
    params = []
    for (xa, xb), (ya, yb), (za, zb), st in boxes:
        params += [xa, xb, ya, yb, za, zb, *st]
    x = origin[0] + h[0] * np.arange(n[0])
    y = origin[1] + h[1] * np.arange(n[1])
    z = origin[2] + h[2] * np.arange(n[2])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    p = np.zeros(Z.shape + (5,))
    for (xa, xb), (ya, yb), (za, zb), st in boxes:
        m = (X >= xa) & (X <= xb) & (Y >= ya) & (Y <= yb) & (Z >= za) & (Z <= zb)
        p[m] = st
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block(n, h)
        blk.init(fvmgpu.IC_REGIONS, params, origin)
        assert np.array_equal(blk.download(), cases.to_conserved(p))


def rel_err(a, b):
    return max(np.abs(a[c] - b[c]).max() / max(np.abs(b[c]).max(), 1.0) for c in range(a.shape[0]))


@pytest.mark.parametrize("seed", [0, 5])
def test_vortex(seed):
    n = 48
    p, h, centre = cases.isentropic_vortex(n, seed=seed)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block((n, n, n), (h, h, h))
        blk.init(fvmgpu.IC_VORTEX, [10.0, 5.0, centre[0], centre[1], 1.0, 1.0, 0.0])
        q = blk.download()
    err = rel_err(q, cases.to_conserved(p))
    print("vortex rel err", err)
    assert err < 1e-14


def test_vortex_2d():
    n = 64
    p3, h, centre = cases.isentropic_vortex(n, seed=2)
    p = p3[:1].copy()
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block((n, n, 1), (h, h, 1.0))
        blk.init(fvmgpu.IC_VORTEX, [10.0, 5.0, centre[0], centre[1], 1.0, 1.0, 0.0])
        q = blk.download()
    assert rel_err(q, cases.to_conserved(p)) < 1e-14


def test_tgv():
    n = 40
    p, h = cases.tgv(n)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block((n, n, n), (h, h, h))
        blk.init(fvmgpu.IC_TGV, [0.1])
        q = blk.download()
    err = rel_err(q, cases.to_conserved(p))
    print("tgv rel err", err)
    assert err < 1e-14


def test_init_errors():
    with fvmgpu.Context() as ctx:
        blk = ctx.block((8, 8, 1), (0.1, 0.1, 1.0))
        with pytest.raises(fvmgpu.FvmError, match="scheme"):
            blk.init(fvmgpu.IC_TGV, [0.1])
        ctx.scheme()
        with pytest.raises(fvmgpu.FvmError):
            blk.init(fvmgpu.IC_TGV, [0.1])            # 2-D block
        with pytest.raises(fvmgpu.FvmError):
            blk.init(fvmgpu.IC_REGIONS, [0.0] * 10)
        with pytest.raises(fvmgpu.FvmError):
            blk.init(fvmgpu.IC_VORTEX, [10.0, 5.0, 5.0, 5.0, 1.0, 1.0, 0.5])   # w in 2-D
        with pytest.raises(fvmgpu.FvmError):
            blk.init(99, [])


def test_cut_pieces_initialised_at_their_offsets():
    """A 101^2 Riemann grid cut 3 x 2 (the reference's uniform cut), every
    piece initialised on the device at its offset in the grid and connected
    to its neighbours: initial state and ten CFL steps equal the whole
    grid's (the oracle's) bit for bit."""
    from _parity import run_oracle
    n = 101
    p, dx = cases.riemann2d(n, seed=0)
    quads = fvmgpu.region_params(cases.REF_QUADRANTS)
    rx, ry = fvmgpu.cut_uniform(n, 3), fvmgpu.cut_uniform(n, 2)
    print("case cut 3x2 pieces initialised at their offsets", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        ctx.numerics(0)
        blocks = {}
        for a, (x0, x1) in enumerate(rx):
            for b, (y0, y1) in enumerate(ry):
                blk = ctx.block((x1 - x0 + 1, y1 - y0 + 1, 1), (dx, dx, 1.0))
                blk.init_at(fvmgpu.IC_REGIONS, quads, (x0, y0, 0))
                blocks[(a, b)] = (blk, x0, y0)
        for (a, b), (blk, _, _) in blocks.items():
            if a + 1 < len(rx):
                blk.connect(1, blocks[(a + 1, b)][0], shared=True)
            if b + 1 < len(ry):
                blk.connect(3, blocks[(a, b + 1)][0], shared=True)

        def gather():
            out = np.zeros((5, 1, n, n))
            for blk, x0, y0 in blocks.values():
                q = blk.download()
                out[:, :, y0:y0 + q.shape[2], x0:x0 + q.shape[3]] = q
            return out

        assert np.array_equal(gather(), cases.to_conserved(p))
        ctx.step(10, cfl=0.5)
        out = gather()
    qo, _ = run_oracle(p, (dx, dx, 1.0), 10, cfl=0.5)
    assert np.array_equal(out, qo)
