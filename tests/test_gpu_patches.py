"""1-to-1 connection patches (fvm_conn_add): ghost layers filled on the
device from donor ranges under the reference's direction transform
(src/base/defs_face_geom.f90:43-63, 384-403; CGNS Conn-1to1 point range,
donor range and transform, src/base/interface_cgns.f90:105-112).

Invariants:
* identity transforms, any tiling of a face by patches from several donors:
  the blocks reproduce the unsplit grid (the oracle on the union) bit for bit;
* a donor stored with its axes permuted and reversed (momentum along its own
  axes): bit for bit the unsplit grid for the schemes whose oracle is mirror
  symmetric.  MUSCL-minmod + AUSM+ and textbook WENO5 + Van Leer are (the
  oracle run on the transposed, reversed field equals the original exactly);
  the reference's Roe flux and its WENO variant are not (3e-5 and 3e-3 after
  three steps), so they are left out here.
"""
import numpy as np
import pytest

import _oracle as O
from _parity import run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


class Part:
    """A block of the union grid: global vertex offset and extents (x, y, z)."""

    def __init__(self, blk, off, ext):
        self.blk, self.off, self.ext = blk, off, ext


def ident_patch(recv, donor, gmin, gmax):
    """Identity-transform patch over the global vertex range gmin..gmax
    (0-based, inclusive, one plane in the face direction)."""
    rmin = [g - o + 1 for g, o in zip(gmin, recv.off)]
    rmax = [g - o + 1 for g, o in zip(gmax, recv.off)]
    dmin = [g - o + 1 for g, o in zip(gmin, donor.off)]
    dmax = [g - o + 1 for g, o in zip(gmax, donor.off)]
    recv.blk.conn_add(rmin, rmax, donor.blk, dmin, dmax, (1, 2, 3))


def make_part(ctx, q, sp, off, ext):
    (x0, y0, z0), (nx, ny, nz) = off, ext
    blk = ctx.block((nx, ny, nz), sp)
    blk.upload(q[:, z0:z0 + nz, y0:y0 + ny, x0:x0 + nx])
    return Part(blk, off, ext)


def gather(parts, shape):
    out = np.zeros((5,) + shape)
    for pt in parts:
        (x0, y0, z0), (nx, ny, nz) = pt.off, pt.ext
        out[:, z0:z0 + nz, y0:y0 + ny, x0:x0 + nx] = pt.blk.download()
    return out


def t_layout(nx, ny, nz, xa, yb):
    """Block A: x 0..xa, all y; B1: x xa..nx-1, y 0..yb; B2: x xa..nx-1,
    y yb..ny-1 (interface planes shared).  A's i+ face takes two patches."""
    return {"A": ((0, 0, 0), (xa + 1, ny, nz)),
            "B1": ((xa, 0, 0), (nx - xa, yb + 1, nz)),
            "B2": ((xa, yb, 0), (nx - xa, ny - yb, nz))}


def link_t(P, nx, ny, nz, xa, yb):
    A, B1, B2 = P["A"], P["B1"], P["B2"]
    zt = nz - 1
    # A's i+ face: y 0..yb from B1, y yb+1..ny-1 from B2
    ident_patch(A, B1, (xa, 0, 0), (xa, yb, zt))
    ident_patch(A, B2, (xa, yb + 1, 0), (xa, ny - 1, zt))
    # B1, B2 i- faces from A
    ident_patch(B1, A, (xa, 0, 0), (xa, yb, zt))
    ident_patch(B2, A, (xa, yb, 0), (xa, ny - 1, zt))
    # B1 j+ <- B2, B2 j- <- B1
    ident_patch(B1, B2, (xa, yb, 0), (nx - 1, yb, zt))
    ident_patch(B2, B1, (xa, yb, 0), (nx - 1, yb, zt))


@pytest.mark.parametrize("mode,graph", [(fvmgpu.LAUNCH_BATCHED, False),
                                        (fvmgpu.LAUNCH_PER_BLOCK, False),
                                        (fvmgpu.LAUNCH_BATCHED, True)])
def test_t_junction_equals_single_block(mode, graph):
    nx, ny, nz, xa, yb = 34, 28, 18, 15, 11
    p = cases.smooth_3d((nz, ny, nx), seed=21)
    q = cases.to_conserved(p)
    sp = (1 / 33, 1 / 27, 1 / 17)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    print(f"case T-junction 34x28x18 mode={mode} graph={graph}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(0)
        ctx.launch_mode(mode)
        if graph:
            ctx.graph(True)
        P = {k: make_part(ctx, q, sp, *v) for k, v in t_layout(nx, ny, nz, xa, yb).items()}
        link_t(P, nx, ny, nz, xa, yb)
        ctx.step(3, dt=1e-3)
        ctx.step(2, dt=1e-3)
        out = gather(P.values(), (nz, ny, nx))
    qo, _ = run_oracle(p, sp, 5, dt=1e-3, **scheme)
    assert np.array_equal(out, qo)


def test_t_junction_2d_cfl():
    """2-D Riemann problem over the T layout, CFL time steps to a fixed end
    time (the CFL step is a max over all blocks)."""
    n, xa, yb = 60, 27, 35
    p, h = cases.riemann2d(n, seed=4)
    q = cases.to_conserved(p)
    sp = (h, h, 1.0)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.AUSMPLUS, rk=O.RK_SSP3)
    print("case T-junction riemann2d 60 cfl", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(0)
        P = {k: make_part(ctx, q, sp, *v) for k, v in t_layout(n, n, 1, xa, yb).items()}
        link_t(P, n, n, 1, xa, yb)
        ctx.step(6, cfl=0.5)
        out = gather(P.values(), (1, n, n))
    qo, _ = run_oracle(p, sp, 6, cfl=0.5, **scheme)
    assert np.array_equal(out, qo)


def rotate_in(sub):
    """Union-frame conserved block (5, nz, ny, nx) -> a donor frame whose i
    axis is -y and j axis is +x: array (5, nz, nx, ny)."""
    s = sub[:, :, ::-1, :].transpose(0, 1, 3, 2)
    out = np.array(s)
    out[1] = -s[2]
    out[2] = s[1]
    return np.ascontiguousarray(out)


def rotate_out(qb):
    s = qb.transpose(0, 1, 3, 2)[:, :, ::-1, :]
    out = np.array(s)
    out[1] = s[2]
    out[2] = -s[1]
    return out


@pytest.mark.parametrize("scheme", [
    dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.AUSMPLUS, rk=O.RK_SSP3),
    dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.VANLEER, rk=O.RK_SSP3, ref_compat=0),
], ids=["muscl_ausmplus", "weno_vanleer"])
@pytest.mark.parametrize("dim", [3, 2])
def test_rotated_donor_equals_single_block(scheme, dim):
    nx, ny, xa = 30, 22, 13
    nz = 12 if dim == 3 else 1
    if dim == 3:
        p = cases.smooth_3d((nz, ny, nx), seed=8)
        sp = (1 / 29, 1 / 21, 1 / 11)
    else:
        p = cases.smooth_3d((2, ny, nx), seed=8)[:1].copy()
        p[..., 3] = 0.0
        sp = (1 / 29, 1 / 21, 1.0)
    q = cases.to_conserved(p)
    nxb = nx - xa
    print(f"case rotated donor dim={dim} {scheme}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**{k: v for k, v in scheme.items()})
        ctx.numerics(0)
        A = ctx.block((xa + 1, ny, nz), sp)
        A.upload(q[:, :, :, :xa + 1])
        B = ctx.block((ny, nxb, nz), (sp[1], sp[0], sp[2]))
        B.upload(rotate_in(q[:, :, :, xa:]))
        # A's i+ face: A direction i runs along B's j (forward), A's j along
        # B's i backward, k along k (the reference's donor_type)
        A.conn_add((xa + 1, 1, 1), (xa + 1, ny, nz), B, (1, ny, 1), (1, 1, nz), (2, -1, 3))
        # B's j- face: the inverse transform (defs_face_geom.f90:361-373)
        B.conn_add((1, 1, 1), (ny, 1, nz), A, (ny, xa + 1, 1), (1, xa + 1, nz), (-2, 1, 3))
        ctx.step(4, dt=1e-3)
        out = np.zeros_like(q)
        out[:, :, :, :xa + 1] = A.download()
        out[:, :, :, xa:] = rotate_out(B.download())
    qo, _ = run_oracle(p, sp, 4, dt=1e-3, **scheme)
    assert np.array_equal(out, qo)


def test_patch_validation():
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        A = ctx.block((10, 8, 6), (0.1, 0.1, 0.1))
        B = ctx.block((10, 8, 6), (0.1, 0.1, 0.1))
        ok = dict(range_min=(10, 1, 1), range_max=(10, 8, 6), donor=B,
                  donor_range_min=(1, 1, 1), donor_range_max=(1, 8, 6))
        bad = [
            dict(ok, donor_type=(1, 1, 3)),                       # not a permutation
            dict(ok, range_min=(9, 1, 1), range_max=(9, 8, 6)),   # not on a face
            dict(ok, donor_range_min=(2, 1, 1), donor_range_max=(2, 8, 6)),  # donor not on a face
            dict(ok, donor_range_max=(1, 7, 6)),                  # extents differ
            dict(ok, donor_range_min=(10, 1, 1), donor_range_max=(10, 8, 6)),  # ghosts map outward
            dict(ok, range_max=(10, 9, 6)),                       # out of the block
        ]
        for kw in bad:
            with pytest.raises(fvmgpu.FvmError):
                A.conn_add(**kw)
        # a patch covering half the face: the step reports the gap
        A.conn_add((10, 1, 1), (10, 4, 6), B, (1, 1, 1), (1, 4, 6))
        u = np.zeros((5, 6, 8, 10))
        u[0] = 1.0
        u[4] = 2.5
        A.upload(u)
        B.upload(u)
        with pytest.raises(fvmgpu.FvmError, match="cover"):
            ctx.step(1, dt=1e-3)
        # overlapping patches
        with pytest.raises(fvmgpu.FvmError):
            A.conn_add((10, 4, 1), (10, 8, 6), B, (1, 4, 1), (1, 8, 6))
            ctx.step(1, dt=1e-3)
        # a patched face is not a periodic or block-connected face
        with pytest.raises(fvmgpu.FvmError):
            A.set_bc(1, fvmgpu.BC_PERIODIC)
        ctx.viscous(1e-3)                    # viscous terms read across patches
        with pytest.raises(fvmgpu.FvmError, match="viscous"):
            B.set_bc_state(1, fvmgpu.BC_FARFIELD, (1.0, 0.0, 0.0, 0.0, 1.0))


VISC = (0.03, 1.0, 0.0, 0.72, 1.0)


def test_t_junction_viscous_equals_single_block():
    """Viscous terms over the T layout: stencil points beyond A's i+ face are
    read from whichever donor covers them (B1 or B2), edge points through
    the donors' own j faces; bit for bit the oracle on the union."""
    nx, ny, nz, xa, yb = 30, 26, 14, 13, 10
    p = cases.smooth_3d((nz, ny, nx), seed=23)
    q = cases.to_conserved(p)
    sp = (1 / 29, 1 / 25, 1 / 13)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    print("case T-junction viscous 30x26x14", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.viscous(*VISC)
        P = {k: make_part(ctx, q, sp, *v) for k, v in t_layout(nx, ny, nz, xa, yb).items()}
        link_t(P, nx, ny, nz, xa, yb)
        ctx.step(4, dt=1e-3)
        out = gather(P.values(), (nz, ny, nx))
    r = O.oracle_solve(p, 3, sp, dt=1e-3, cfl=0.0, nsteps=4, p_from_q=1, bc=(0,) * 6,
                       weno_textbook=0, visc=VISC, **scheme)
    assert np.array_equal(out, r["q"][:, 3:-3, 3:-3, 3:-3])


@pytest.mark.parametrize("dim", [3, 2])
def test_rotated_donor_viscous_equals_single_block(dim):
    """Viscous terms across a donor stored with its axes permuted and
    reversed: the stencil's velocities come back rotated into the receiver's
    axes; bit for bit the oracle on the union (MUSCL-minmod + AUSM+, whose
    oracle is mirror symmetric, as for the inviscid case)."""
    nx, ny, xa = 28, 20, 12
    nz = 10 if dim == 3 else 1
    if dim == 3:
        p = cases.smooth_3d((nz, ny, nx), seed=9)
        sp = (1 / 27, 1 / 19, 1 / 9)
    else:
        p = cases.smooth_3d((2, ny, nx), seed=9)[:1].copy()
        p[..., 3] = 0.0
        sp = (1 / 27, 1 / 19, 1.0)
    q = cases.to_conserved(p)
    nxb = nx - xa
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.AUSMPLUS, rk=O.RK_SSP3)
    print(f"case rotated donor viscous dim={dim}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.viscous(*VISC)
        A = ctx.block((xa + 1, ny, nz), sp)
        A.upload(q[:, :, :, :xa + 1])
        B = ctx.block((ny, nxb, nz), (sp[1], sp[0], sp[2]))
        B.upload(rotate_in(q[:, :, :, xa:]))
        A.conn_add((xa + 1, 1, 1), (xa + 1, ny, nz), B, (1, ny, 1), (1, 1, nz), (2, -1, 3))
        B.conn_add((1, 1, 1), (ny, 1, nz), A, (ny, xa + 1, 1), (1, xa + 1, nz), (-2, 1, 3))
        ctx.step(4, dt=1e-3)
        out = np.zeros_like(q)
        out[:, :, :, :xa + 1] = A.download()
        out[:, :, :, xa:] = rotate_out(B.download())
    dimo = 3 if dim == 3 else 2
    r = O.oracle_solve(p, dimo, sp, dt=1e-3, cfl=0.0, nsteps=4, p_from_q=1, bc=(0,) * 6,
                       weno_textbook=0, visc=VISC, **scheme)
    gz = 3 if dim == 3 else 0
    qo = r["q"][:, gz:gz + nz, 3:-3, 3:-3].copy()
    if dim == 2:
        qo[3] = 0.0
    assert np.array_equal(out, qo)
