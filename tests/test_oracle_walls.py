"""The oracle's wall faces (CPU).  The reference lists wall codes
(src/defs/type_defs_bc.f90:11-30) but never applies them; this library defines
them: ghost -l is the mirror vertex l with the normal velocity negated (slip
wall 30/31, symmetry 50) or every velocity component negated (adiabatic
no-slip wall 33); a no-slip wall's vertices lose their momentum (and its
kinetic energy) after every stage; the viscous stencil crosses a wall to the
mirror vertex with the same sign rule.  Pinned here by:
- the viscous residual of a shear mode sin(pi y / H) between no-slip walls,
  which the discrete operator maps to a multiple of itself (0 at the walls);
- a slip wall ≡ the mirror-symmetric grid restricted to its half;
- the decay of that shear mode in time against the discrete viscous rate,
  with the wall velocity exactly zero after every step."""
import numpy as np
import pytest

import _oracle as O

GAMMA = 1.4


def shear_mode(ni, nj, U, p0, mu):
    """Parallel flow u(y) = U sin(pi y / H) between walls at j = 0 and j = nj-1
    (H = 1), rho = 1, v = w = 0, uniform pressure p0 (so T and mu uniform)."""
    dy = 1.0 / (nj - 1)
    y = np.arange(nj) * dy
    u = U * np.sin(np.pi * y)
    u[0] = u[-1] = 0.0
    p = np.zeros((1, nj, ni, 5))
    p[..., 0] = 1.0
    p[..., 1] = u[None, :, None]
    p[..., 4] = p0
    return p, dy, u


def test_noslip_viscous_residual_of_the_shear_mode():
    """x-momentum residual of u = U sin(pi y) with no-slip walls at j = 0 and
    j = n-1, periodic in x: the discrete operator maps the mode to
    mu (2 cos(pi dy) - 2) / dy^2 u_j at every vertex -- 0 at the walls, where
    the stencil reads the mirror vertex with u negated (a mirror without the
    negation gives 2 mu u_1 / dy^2 there, a clamp mu u_1 / dy^2)."""
    ni, nj, mu = 6, 33, 0.05
    p, dy, u = shear_mode(ni, nj, 0.01, 1.0, mu)
    sp = (0.1, dy, 1.0)
    visc = (mu, 1.0, 0.0, 0.72, 1.0)
    for bc_y in (O.BC_NOSLIP,):
        rv = O.viscous_residual(p, 2, sp, visc, bc=(O.BC_PERIODIC, O.BC_PERIODIC, bc_y, bc_y, 0, 0))
        lam = mu * (2.0 * np.cos(np.pi * dy) - 2.0) / dy ** 2
        got = rv[1, 0, :, 2]
        assert abs(got[0]) < 1e-12 and abs(got[-1]) < 1e-12
        assert np.allclose(got[1:-1], lam * u[1:-1], rtol=1e-9, atol=0)
        assert np.allclose(rv[1, 0], rv[1, 0, :, :1], rtol=0, atol=1e-15)   # uniform along x
    # the plausible mistakes named above, planted in the boundary values
    bad_mirror = mu * (u[1] - 2 * u[0] + u[1]) / dy ** 2
    bad_clamp = mu * (u[1] - 2 * u[0] + u[0]) / dy ** 2
    assert abs(bad_mirror) > 1e-3 and abs(bad_clamp) > 1e-3


@pytest.mark.parametrize("kind", [O.BC_WALL, O.BC_SYMMETRY])
def test_slip_wall_equals_mirrored_grid(kind):
    """A smooth 2-D field on x in [0, 1] with a slip wall (or symmetry plane)
    at x = 0, viscous terms on, 4 SSP-RK3 steps ≡ the grid mirrored about
    x = 0 (u negated on the mirror half) restricted to x >= 0, bit for bit.
    Van Leer splitting keeps the mirrored grid exactly symmetric; the
    reference's Roe flux does not (its Harten width eps |V + c| changes with
    the sign of V), so it is not used here."""
    rng = np.random.default_rng(5)
    ni, nj = 14, 11
    x = np.linspace(0.0, 1.0, ni)
    y = np.linspace(0.0, 1.0, nj)
    X, Y = np.meshgrid(x, y)
    p = np.zeros((1, nj, ni, 5))
    p[0, ..., 0] = 1.0 + 0.1 * np.cos(np.pi * X) * (1 + 0.2 * np.sin(3 * Y + rng.uniform()))
    p[0, ..., 1] = 0.2 * np.sin(np.pi * X) * np.cos(2 * Y)          # odd in x: 0 at x = 0
    p[0, ..., 2] = 0.1 * np.cos(np.pi * X) * np.sin(2 * Y + 0.3)
    p[0, ..., 4] = 1.0 + 0.05 * np.cos(np.pi * X) * np.cos(Y)
    full = np.concatenate([p[:, :, :0:-1].copy(), p], axis=2)       # x from -1 to 1
    full[:, :, :ni - 1, 1] *= -1.0
    dx, dy = 1.0 / (ni - 1), 1.0 / (nj - 1)
    visc = (0.02, 1.0, 0.0, 0.72, 1.0)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.VANLEER, rk=O.RK_SSP3, dt=2e-3,
                  cfl=0.0, nsteps=4, p_from_q=1, weno_textbook=0, visc=visc)
    half = O.oracle_solve(p, 2, (dx, dy, 1.0), bc=(kind, 0, 0, 0, 0, 0), **scheme)["q"]
    whole = O.oracle_solve(full, 2, (dx, dy, 1.0), bc=(0,) * 6, **scheme)["q"]
    h = half[:, 0, 3:-3, 3:-3]
    w = whole[:, 0, 3:-3, 3 + ni - 1:-3]
    assert np.array_equal(h, w)
    # extrapolation at x = 0 instead of the mirror changes the result
    clamp = O.oracle_solve(p, 2, (dx, dy, 1.0), bc=(0, 0, 0, 0, 0, 0), **scheme)["q"][:, 0, 3:-3, 3:-3]
    assert np.abs(clamp - w).max() > 1e-6


def test_noslip_shear_mode_decays_at_the_viscous_rate():
    """The shear mode between no-slip walls, low Mach, integrated in time
    (MUSCL + Roe, RK4): its amplitude decays as exp(lambda t) with the
    discrete rate lambda = nu (2 cos(pi dy) - 2) / dy^2 (within 1e-4 of the
    decay at t = 0.4; measured 1.1e-5), and the wall vertices' momentum is
    exactly zero."""
    ni, nj, mu, U = 5, 33, 0.05, 0.01
    p, dy, u = shear_mode(ni, nj, U, 1.0 / GAMMA, mu)
    sp = (0.1, dy, 1.0)
    visc = (mu, 1.0 / GAMMA, 0.0, 0.72, 1.0)       # T = p / rho = t_ref: mu = mu_ref
    dt, nsteps = 2e-3, 200
    r = O.oracle_solve(p, 2, sp, recon=O.MUSCL, limiter=O.LIM_NONE, flux=O.ROE, rk=O.RK4, dt=dt,
                       cfl=0.0, nsteps=nsteps, p_from_q=1,
                       bc=(O.BC_PERIODIC, O.BC_PERIODIC, O.BC_NOSLIP, O.BC_NOSLIP, 0, 0),
                       weno_textbook=0, visc=visc)
    q = r["q"][:, 0, 3:-3, 3:-3]
    assert np.all(q[1, 0, :] == 0.0) and np.all(q[1, -1, :] == 0.0)
    assert np.all(q[2, 0, :] == 0.0) and np.all(q[2, -1, :] == 0.0)
    uu = q[1, :, 2] / q[0, :, 2]
    amp = np.dot(uu, u) / np.dot(u, u)
    lam = mu * (2.0 * np.cos(np.pi * dy) - 2.0) / dy ** 2
    expect = np.exp(lam * dt * nsteps)
    print("amplitude", amp, "expected", expect)
    assert abs(amp / expect - 1.0) < 1e-4
    # a slip wall there instead lets the wall vertices slide: a different decay
    r2 = O.oracle_solve(p, 2, sp, recon=O.MUSCL, limiter=O.LIM_NONE, flux=O.ROE, rk=O.RK4, dt=dt,
                        cfl=0.0, nsteps=nsteps, p_from_q=1,
                        bc=(O.BC_PERIODIC, O.BC_PERIODIC, O.BC_WALL, O.BC_WALL, 0, 0),
                        weno_textbook=0, visc=visc)
    q2 = r2["q"][:, 0, 3:-3, 3:-3]
    amp2 = np.dot(q2[1, :, 2] / q2[0, :, 2], u) / np.dot(u, u)
    assert abs(amp2 / expect - 1.0) > 5e-2
