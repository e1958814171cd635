"""Wall faces with viscous terms on the GPU against the oracle (whose wall
rules are pinned in test_oracle_walls.py): adiabatic no-slip walls (BC 33)
and slip walls / symmetry planes (30, 31, 50), single and connected blocks,
corners where walls meet, fixed dt and CFL steps; bit for bit in the exact
family, within 1e-12 in the fast family."""
import numpy as np
import pytest

import _oracle as O
from _parity import rel_linf
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

MUSCL_ROE = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK_SSP3)
WENO_AUSMP = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK4, ref_compat=0)
VISC = (0.03, 1.0, 0.0, 0.72, 1.0)


def field(shape, seed):
    """Smooth positive state with velocities of both signs (nk, nj, ni, 5)."""
    p = cases.smooth_3d(shape, seed=seed)
    if shape[0] == 1:
        p[..., 3] = 0.0
    return p


def gpu(p, sp, bc, nsteps, visc=VISC, numerics=0, dt=1e-3, cfl=0.0, **scheme):
    nk, nj, ni, _ = p.shape
    scheme = {k: v for k, v in scheme.items()}
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(numerics)
        if visc:
            ctx.viscous(*visc)
        b = ctx.block((ni, nj, nk), sp)
        for f, kind in enumerate(bc):
            if kind and (nk > 1 or f < 4):
                b.set_bc(f, kind)
        b.upload(cases.to_conserved(p))
        ctx.step(nsteps, dt=dt, cfl=cfl)
        return b.download(), ctx.norms()[0]


def oracle(p, sp, bc, nsteps, visc=VISC, dt=1e-3, cfl=0.0, recon=O.MUSCL, limiter=O.LIM_MINMOD,
           flux=O.ROE, rk=O.RK_SSP3, ref_compat=1):
    nk = p.shape[0]
    dim = 3 if nk > 1 else 2
    r = O.oracle_solve(p, dim, sp, recon=recon, limiter=limiter, flux=flux, rk=rk, dt=dt,
                       cfl=cfl, nsteps=nsteps, p_from_q=1, bc=tuple(bc),
                       weno_textbook=0 if ref_compat else 1, visc=visc)
    gz = 3 if dim == 3 else 0
    q = r["q"][:, gz:gz + nk, 3:-3, 3:-3].copy()
    if dim == 2:
        q[3] = 0.0
    return q, r


P, N, W, S, E = O.BC_PERIODIC, O.BC_NOSLIP, O.BC_WALL, O.BC_SYMMETRY, O.BC_EXTRAP


@pytest.mark.parametrize("bc", [(P, P, N, N, 0, 0), (N, N, N, N, 0, 0), (N, W, S, N, 0, 0)])
@pytest.mark.parametrize("scheme", [MUSCL_ROE, WENO_AUSMP])
def test_walls_2d(bc, scheme):
    p = field((1, 17, 20), seed=11)
    sp = (1 / 19, 1 / 16, 1.0)
    print(f"case walls 2-D 20x17 bc {bc} {scheme['recon']}/{scheme['flux']}", flush=True)
    q, norms = gpu(p, sp, bc, 6, **scheme)
    qo, r = oracle(p, sp, bc, 6, **scheme)
    assert np.array_equal(q, qo)
    assert np.allclose(norms, r["res"][-1], rtol=1e-12, atol=0)
    for f, kind in enumerate(bc[:4]):                    # the no-slip wall vertices' momentum
        if kind == N:
            side = 0 if f % 2 == 0 else -1
            wall = q[1:3, 0, :, side] if f < 2 else q[1:3, 0, side, :]
            assert np.all(wall == 0.0)


@pytest.mark.parametrize("bc,scheme", [((N,) * 6, MUSCL_ROE), ((N, N, P, P, W, N), MUSCL_ROE),
                                       ((S, E, N, N, N, W), MUSCL_ROE), ((N, W, N, N, S, N), WENO_AUSMP)])
def test_walls_3d(bc, scheme):
    """Walls on several faces of a 3-D block, corners where no-slip walls
    meet (each crossing flips the velocity again), viscous terms; WENO's
    five-point stencil reads three mirror layers."""
    p = field((12, 13, 14), seed=12)
    sp = (1 / 13, 1 / 12, 1 / 11)
    print(f"case walls 3-D 14x13x12 bc {bc} {scheme['recon']}/{scheme['flux']}", flush=True)
    q, _ = gpu(p, sp, bc, 4, **scheme)
    qo, _ = oracle(p, sp, bc, 4, **scheme)
    assert np.array_equal(q, qo)
    qf, _ = gpu(p, sp, bc, 4, numerics=1, **scheme)
    assert rel_linf(qf, qo) <= 1e-12


def test_noslip_inviscid_and_cfl():
    """No-slip walls without viscous terms, and CFL steps (the wave speed of
    the wall vertices after their momentum is removed sets the next dt)."""
    p = field((10, 12, 13), seed=13)
    sp = (1 / 12, 1 / 11, 1 / 9)
    bc = (N, N, P, P, N, E)
    q, _ = gpu(p, sp, bc, 5, visc=None, **MUSCL_ROE)
    qo, _ = oracle(p, sp, bc, 5, visc=None, **MUSCL_ROE)
    assert np.array_equal(q, qo)
    q, _ = gpu(p, sp, bc, 5, dt=0.0, cfl=0.5, **MUSCL_ROE)
    qo, r = oracle(p, sp, bc, 5, dt=0.0, cfl=0.5, **MUSCL_ROE)
    assert np.array_equal(q, qo)


@pytest.mark.parametrize("ncut", [(2, 1, 1), (2, 2, 2)])
def test_walls_on_connected_blocks(ncut):
    """A walled box cut into connected blocks: each block's outer faces carry
    the walls; the viscous stencil crosses an interface and then a wall (an
    edge of the box), and the wall vertices on the shared planes are fixed in
    both blocks; bit for bit the oracle on the whole box."""
    p = field((12, 14, 15), seed=14)
    sp = (1 / 14, 1 / 13, 1 / 11)
    bc = (N, W, N, N, S, N)
    print(f"case walled box 15x14x12 cut {ncut}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**MUSCL_ROE)
        ctx.viscous(*VISC)
        blocks = ctx.cut_blocks(cases.to_conserved(p), sp, ncut)
        for (a, b, c), (blk, _) in blocks.items():
            idx = (a, b, c)
            for d in range(3):
                if idx[d] == 0:
                    blk.set_bc(2 * d, bc[2 * d])
                if idx[d] == ncut[d] - 1:
                    blk.set_bc(2 * d + 1, bc[2 * d + 1])
        ctx.step(4, dt=1e-3)
        out = np.zeros((5,) + p.shape[:3])
        for blk, (x0, y0, z0) in blocks.values():
            q = blk.download()
            out[:, z0:z0 + q.shape[1], y0:y0 + q.shape[2], x0:x0 + q.shape[3]] = q
    qo, _ = oracle(p, sp, bc, 4, **MUSCL_ROE)
    assert np.array_equal(out, qo)
