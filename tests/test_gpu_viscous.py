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
