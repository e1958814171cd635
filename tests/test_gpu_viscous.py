"""Central viscous fluxes on the GPU against the oracle (whose viscous term
is pinned by closed forms in test_oracle_viscous.py): bit for bit in the
exact family, within 1e-12 after one step in the fast family; conservation
on the periodic Taylor-Green box."""
import numpy as np
import pytest

import _oracle as O
from _parity import rel_linf
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


def gpu_run(p, spacing, nsteps, visc, numerics=0, dt=0.0, cfl=0.0, periodic=True, **scheme):
    nk, nj, ni, _ = p.shape
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(numerics)
        ctx.viscous(*visc)
        b = ctx.block((ni, nj, nk), spacing)
        if periodic:
            b.set_periodic()
        b.upload(cases.to_conserved(p))
        ctx.step(nsteps, dt=dt, cfl=cfl)
        return b.download(), ctx.norms()


def oracle_run(p, spacing, nsteps, visc, dt=0.0, cfl=0.0, periodic=True, recon=O.MUSCL,
               limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3, ref_compat=1):
    nk = p.shape[0]
    dim = 3 if nk > 1 else 2
    r = O.oracle_solve(p, dim, spacing, recon=recon, limiter=limiter, flux=flux, rk=rk, dt=dt,
                       cfl=cfl, nsteps=nsteps, p_from_q=1, bc=(1 if periodic else 0,) * 6,
                       weno_textbook=0 if ref_compat else 1, visc=visc)
    gz = 3 if dim == 3 else 0
    return r["q"][:, gz:gz + nk, 3:-3, 3:-3], r


TGV_SCHEME = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.VANLEER, rk=O.RK4, ref_compat=0)


@pytest.mark.parametrize("scheme,tile", [(TGV_SCHEME, ""),
                                         (dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA,
                                               flux=O.ROE, rk=O.RK_SSP3), ""),
                                         (dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA,
                                               flux=O.ROE, rk=O.RK_SSP3), "4")])
def test_tgv_exact_bitwise(scheme, tile, monkeypatch):
    # Re = 100 so the viscous term is not negligible at this resolution;
    # shape 4: the register-reconstruction kernel adding the viscous residual
    if tile:
        monkeypatch.setenv("FVM_TILE", tile)
    p, h = cases.tgv(20)
    visc = cases.tgv_params(re=100.0)
    print(f"case tgv20 exact {scheme} tile={tile}", flush=True)
    q, (norms, nan) = gpu_run(p, (h, h, h), 3, visc, dt=2e-3, **scheme)
    qo, r = oracle_run(p, (h, h, h), 3, visc, dt=2e-3, **scheme)
    assert np.array_equal(q, qo)
    assert np.allclose(norms, r["res"][-1], rtol=1e-13, atol=0)


def test_viscous_2d_extrapolation_bitwise():
    """2-D, transmissive boundaries (stencil clamped at the block edge),
    non-uniform temperature so Sutherland's law varies."""
    n = 33
    x = np.linspace(0, 1, n)
    Y, X = np.meshgrid(x, x, indexing="ij")
    p = np.zeros((1, n, n, 5))
    p[0, ..., 0] = 1.0 + 0.2 * np.sin(3 * X) * np.cos(2 * Y)
    p[0, ..., 1] = 0.3 * np.sin(2 * np.pi * Y)
    p[0, ..., 2] = 0.2 * np.cos(2 * np.pi * X)
    p[0, ..., 4] = 1.0 + 0.1 * X * Y
    visc = (0.02, 1.0, 0.4, 0.71, 1.0)
    sp = (1 / (n - 1), 1 / (n - 1), 1.0)
    print("case viscous 2-D extrapolation", flush=True)
    q, _ = gpu_run(p, sp, 4, visc, dt=5e-4, periodic=False)
    qo, _ = oracle_run(p, sp, 4, visc, dt=5e-4, periodic=False)
    assert np.array_equal(q, qo)


def test_tgv_fast_one_step_and_conservation():
    p, h = cases.tgv(24)
    visc = cases.tgv_params(re=400.0)
    print("case tgv24 fast vs exact, CFL", flush=True)
    qe, _ = gpu_run(p, (h, h, h), 1, visc, numerics=0, cfl=0.5, **TGV_SCHEME)
    qf, _ = gpu_run(p, (h, h, h), 1, visc, numerics=1, cfl=0.5, **TGV_SCHEME)
    assert rel_linf(qf, qe) <= 1e-12
    q20, _ = gpu_run(p, (h, h, h), 20, visc, numerics=1, cfl=0.5, **TGV_SCHEME)
    q0 = cases.to_conserved(p)
    for c in range(5):
        assert abs(q20[c].sum() - q0[c].sum()) <= 1e-11 * np.abs(q0[c]).sum() + 1e-12
    # viscosity dissipates kinetic energy
    assert cases.kinetic_energy(q20) < cases.kinetic_energy(q0)


def test_tgv_initial_dissipation_rate():
    """Early kinetic-energy decay of the TGV equals the closed form
    -dE/dt = nu <|omega|^2> = 0.75 nu of the initial field (incompressible
    limit; Ma = 0.1 here), within 5 %."""
    p, h = cases.tgv(32)
    re = 100.0
    print("case tgv32 dissipation", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**TGV_SCHEME)
        ctx.numerics(1)
        ctx.viscous(*cases.tgv_params(re=re))
        b = ctx.block((32, 32, 32), (h, h, h))
        b.set_periodic()
        b.upload(cases.to_conserved(p))
        e0 = cases.kinetic_energy(b.download())
        ctx.step(10, cfl=0.5)
        t, _ = ctx.time()
        e1 = cases.kinetic_energy(b.download())
    rate = (e0 - e1) / t
    print("dissipation rate", rate, "closed form", 0.75 / re)
    assert abs(rate / (0.75 / re) - 1.0) < 0.05


def test_tgv192_full_size_bitwise():
    """The Taylor-Green workload at its full size (192^3, Re 1600, WENO5 +
    Van Leer + viscous, RK4), exact family: one CFL step equals the oracle's
    bit for bit (the oracle on the host's threads: a few seconds)."""
    import os
    p, h = cases.tgv(192)
    visc = cases.tgv_params()
    print("case tgv192 full size, 1 CFL step", flush=True)
    q, _ = gpu_run(p, (h, h, h), 1, visc, cfl=0.5, **TGV_SCHEME)
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1)
    qo, r = oracle_run(p, (h, h, h), 1, visc, cfl=0.5, **TGV_SCHEME)
    assert np.array_equal(q, qo)


def _cut_run(p, spacing, ncut, periodic, nsteps, visc, numerics=0, launch=None, dt=2e-3, **scheme):
    """The field p cut into ncut blocks (connected across their shared
    planes, periodic directions wrapping across blocks), viscous terms on;
    the blocks' results gathered into the whole grid (shared planes checked
    equal in both blocks)."""
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(numerics)
        ctx.viscous(*visc)
        if launch is not None:
            ctx.launch_mode(launch)
        blocks = ctx.cut_blocks(cases.to_conserved(p), spacing, ncut, periodic=periodic)
        ctx.step(nsteps, dt=dt)
        out = np.full((5,) + p.shape[:3], np.nan)
        for blk, (x0, y0, z0) in blocks.values():
            q = blk.download()
            sl = (slice(None), slice(z0, z0 + q.shape[1]), slice(y0, y0 + q.shape[2]),
                  slice(x0, x0 + q.shape[3]))
            seen = ~np.isnan(out[sl])
            assert np.array_equal(out[sl][seen], q[seen])   # shared planes agree
            out[sl] = q
        return out, ctx.norms()


@pytest.mark.parametrize("ncut,launch", [((3, 2, 2), None), ((2, 3, 1), fvmgpu.LAUNCH_PER_BLOCK),
                                         ((1, 1, 3), fvmgpu.LAUNCH_PERSISTENT)])
def test_tgv_cut_grid_equals_oracle(ncut, launch):
    """Viscous terms on connected blocks: the periodic Taylor-Green box cut
    into blocks (interior interfaces share a vertex plane, the periodic wraps
    abut), three RK4 steps, bit for bit the oracle on the whole box; the
    stencil's edge and corner points cross two or three interfaces."""
    p, h = cases.tgv(18)
    visc = cases.tgv_params(re=100.0)
    print(f"case tgv18 cut {ncut} launch {launch}", flush=True)
    q, _ = _cut_run(p, (h, h, h), ncut, (True, True, True), 3, visc, launch=launch, **TGV_SCHEME)
    qo, _ = oracle_run(p, (h, h, h), 3, visc, dt=2e-3, **TGV_SCHEME)
    assert np.array_equal(q, qo)


def test_viscous_cut_grid_open_boundaries():
    """A smooth 3-D field with extrapolated outer faces, cut 2 x 2 x 2, MUSCL
    + Roe, SSP-RK3, viscous: bit for bit the oracle; the fast family within
    1e-12."""
    p = cases.smooth_3d((14, 16, 18), seed=4)
    sp = (1 / 17, 1 / 15, 1 / 13)
    visc = (0.05, 1.0, 0.0, 0.72, 1.0)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK_SSP3)
    print("case smooth3d 14x16x18 cut 2x2x2 open, viscous", flush=True)
    q, _ = _cut_run(p, sp, (2, 2, 2), (False, False, False), 3, visc, dt=1e-3, **scheme)
    qo, _ = oracle_run(p, sp, 3, visc, dt=1e-3, periodic=False, **scheme)
    assert np.array_equal(q, qo)
    qf, _ = _cut_run(p, sp, (2, 2, 2), (False, False, False), 3, visc, numerics=1, dt=1e-3,
                     **scheme)
    assert rel_linf(qf, qo) <= 1e-12


def test_viscous_cut_grid_2d():
    """2-D: a Riemann-like smooth field cut 3 x 2 with extrapolated faces, and
    periodic in x across the blocks; bit for bit the oracle."""
    p = cases.smooth_3d((1, 20, 23), seed=6)
    p[..., 3] = 0.0
    sp = (1 / 22, 1 / 19, 1.0)
    visc = (0.02, 1.0, 0.0, 0.72, 1.0)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
    for periodic in ((False, False, False), (True, False, False)):
        print(f"case smooth2d 23x20 cut 3x2 periodic {periodic}", flush=True)
        with fvmgpu.Context() as ctx:
            ctx.scheme(**scheme)
            ctx.viscous(*visc)
            blocks = ctx.cut_blocks(cases.to_conserved(p), sp, (3, 2), periodic=periodic)
            ctx.step(3, dt=1e-3)
            out = np.zeros((5, 1, 20, 23))
            for blk, (x0, y0, _) in blocks.values():
                qb = blk.download()
                out[:, :, y0:y0 + qb.shape[2], x0:x0 + qb.shape[3]] = qb
        bc = (1, 1, 0, 0, 0, 0) if periodic[0] else (0,) * 6
        r = O.oracle_solve(p, 2, sp, recon=scheme["recon"], limiter=scheme["limiter"],
                           flux=scheme["flux"], rk=scheme["rk"], dt=1e-3, cfl=0.0, nsteps=3,
                           p_from_q=1, bc=bc, weno_textbook=0, visc=visc)
        qo = r["q"][:, 0:1, 3:-3, 3:-3].copy()
        qo[3] = 0.0                            # 2-D blocks hold no rho w
        assert np.array_equal(out, qo)
