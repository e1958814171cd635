"""The oracle's central viscous fluxes (no reference counterpart: the
reference's viscous_flux_sub is an empty stub, src/solve/class_space_discret.f90:
671-675).  Pinned by closed forms: Sutherland's law at a textbook point,
exactness on linear and quadratic fields, telescoping conservation on a
periodic box and second-order convergence to the analytic Navier-Stokes
viscous term."""
import numpy as np
import pytest

import _oracle as O

MU, TREF, S, PR, R = 0.01, 1.0, 0.5, 0.7, 1.0


def grid(n, length):
    return [length[d] / (n[d] - 1) * np.arange(n[d]) for d in range(3)]


def test_sutherland_air():
    # air: mu = 1.716e-5 Pa s at 273.15 K, S = 110.4 K; mu(300 K) = 1.846e-5 (tables)
    mu = O.sutherland(300.0, 1.716e-5, 273.15, 110.4)
    assert abs(mu - 1.846e-5) < 1e-3 * 1.846e-5
    assert O.sutherland(273.15, 1.716e-5, 273.15, 110.4) == 1.716e-5
    # S = 0: (T/T_ref)^(3/2) T_ref / T, the square-root law of hard spheres
    assert abs(O.sutherland(4.0, 2.0, 1.0, 0.0) - 4.0) < 1e-14


def linear_state(n, A, length=(1.0, 1.5, 2.0)):
    x, y, z = grid(n, length)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    xs = (X, Y, Z)
    p = np.zeros(X.shape + (5,))
    p[..., 0] = 1.0
    for i in range(3):
        p[..., 1 + i] = sum(A[i][j] * xs[j] for j in range(3))
    p[..., 4] = R * TREF          # T = t_ref: mu = mu_ref
    return p, [length[d] / (n[d] - 1) for d in range(3)]


def test_linear_velocity_exact():
    """u = A x, T const: the stress is constant, so the momentum residual
    vanishes and the energy residual is the dissipation tau : grad u."""
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, (3, 3))
    n = (9, 8, 7)
    p, dr = linear_state(n, A)
    rv = O.viscous_residual(p, 3, dr, (MU, TREF, S, PR, R))
    inner = (slice(1, -1),) * 3
    assert np.abs(rv[0]).max() == 0.0
    for c in range(1, 4):
        assert np.abs(rv[c][inner]).max() < 1e-13
    div = np.trace(A)
    tau = MU * (A + A.T) - 2.0 / 3.0 * MU * div * np.eye(3)
    diss = np.sum(tau * A)        # tau_ij d_j u_i
    assert np.allclose(rv[4][inner], diss, rtol=1e-12, atol=1e-15)


def test_quadratic_shear_exact():
    """u = y^2, T const: x-momentum residual d/dy(mu du/dy) = 2 mu exactly."""
    n = (6, 9, 6)
    x, y, z = grid(n, (1.0, 1.0, 1.0))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    p = np.zeros(X.shape + (5,))
    p[..., 0] = 1.0
    p[..., 1] = Y * Y
    p[..., 4] = R * TREF
    rv = O.viscous_residual(p, 3, [1 / 5, 1 / 8, 1 / 5], (MU, TREF, S, PR, R))
    assert np.allclose(rv[1][1:-1, 1:-1, 1:-1], 2.0 * MU, rtol=1e-11)
    assert np.abs(rv[2][1:-1, 1:-1, 1:-1]).max() < 1e-12


def periodic_field(n, dim):
    h = 2 * np.pi / n
    x = h * np.arange(n)
    if dim == 3:
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    else:
        Y, X = np.meshgrid(x, x, indexing="ij")
        Z = np.zeros_like(X)
        Y, X, Z = Y[None], X[None], Z[None]
    p = np.zeros(X.shape + (5,))
    p[..., 0] = 1.0
    p[..., 1] = np.sin(X + 2 * Y)
    p[..., 2] = np.cos(Y - Z)
    p[..., 3] = np.sin(Z + X) if dim == 3 else 0.0
    p[..., 4] = R * TREF
    return p, h, (X, Y, Z)


def analytic(X, Y, Z, dim):
    """mu (lap u_i + 1/3 d_i div u) of the field above."""
    if dim == 3:
        return (MU * (-5 * np.sin(X + 2 * Y) + (-np.sin(X + 2 * Y) - np.sin(Z + X)) / 3),
                MU * (-2 * np.cos(Y - Z) + (-2 * np.sin(X + 2 * Y) - np.cos(Y - Z)) / 3),
                MU * (-2 * np.sin(Z + X) + (np.cos(Y - Z) - np.sin(Z + X)) / 3))
    # 2-D: v = cos(y), w = 0
    return (MU * (-5 * np.sin(X + 2 * Y) + (-np.sin(X + 2 * Y)) / 3),
            MU * (-np.cos(Y) + (-2 * np.sin(X + 2 * Y) - np.cos(Y)) / 3),
            0.0 * X)


@pytest.mark.parametrize("dim", [2, 3])
def test_second_order_convergence(dim):
    errs = []
    for n in (16, 32):
        p, h, (X, Y, Z) = periodic_field(n, dim)
        rv = O.viscous_residual(p, dim, (h, h, h), (MU, TREF, S, PR, R), bc=(1,) * 6)
        ex = analytic(X, Y, Z, dim)
        errs.append(max(np.abs(rv[1 + c] - ex[c]).max() for c in range(dim)))
    rate = np.log2(errs[0] / errs[1])
    assert 1.8 < rate < 2.3, (errs, rate)


def test_periodic_conservation():
    """Telescoping: on a periodic box the residual of every component sums to
    zero (to round-off), whatever the field."""
    rng = np.random.default_rng(4)
    n = 12
    h = 2 * np.pi / n
    x = h * np.arange(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    p = np.zeros(X.shape + (5,))
    ph = rng.uniform(0, 2 * np.pi, 6)
    p[..., 0] = 1.0 + 0.2 * np.sin(X + ph[0])
    p[..., 1] = np.sin(Y + ph[1]) * np.cos(Z)
    p[..., 2] = np.cos(X + ph[2])
    p[..., 3] = np.sin(X + Y + ph[3])
    p[..., 4] = 1.0 + 0.1 * np.cos(Z + ph[4])
    rv = O.viscous_residual(p, 3, (h, h, h), (MU, TREF, S, PR, R), bc=(1,) * 6)
    for c in range(1, 5):
        assert abs(rv[c].sum()) < 1e-12 * np.abs(rv[c]).sum() + 1e-15


def test_solve_with_viscosity_differs_and_conserves():
    """Through orc_solve: the viscous run conserves mass, momentum and energy
    on a periodic box and differs from the inviscid one."""
    p, h, _ = periodic_field(12, 3)
    p[..., 4] = 1.0 + 0.05 * np.sin(np.arange(12))[None, None, :]
    kw = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4, dt=1e-3,
              nsteps=3, bc=(1,) * 6)
    a = O.oracle_solve(p, 3, (h, h, h), visc=(0.05, TREF, S, PR, R), **kw)
    b = O.oracle_solve(p, 3, (h, h, h), **kw)
    qa = a["q"][:, 3:-3, 3:-3, 3:-3]
    qb = b["q"][:, 3:-3, 3:-3, 3:-3]
    assert np.abs(qa - qb).max() > 1e-6
    q0 = np.zeros_like(qa)
    from fvmgpu import cases
    q0 = cases.to_conserved(p)
    for c in range(5):
        assert abs(qa[c].sum() - q0[c].sum()) < 1e-12 * np.abs(q0[c]).sum() + 1e-13
