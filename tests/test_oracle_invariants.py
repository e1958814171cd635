"""Closed forms and invariants that pin the oracle where the reference has no
counterpart to compare with (RK4, periodic ghosts, CFL step, textbook
WENO5-JS, AUSM+), plus physics checks of the reference schemes themselves.
Each check is chosen so that a plausible slip (dropped term, wrong sign,
wrong index, transposed operand) fails it."""
import math

import numpy as np
import pytest

import _oracle as O
import fvmgpu.cases as cases

G = 1.4


def phys_flux(q, d, gamma=G):
    """Euler flux f(q) . e_d for AoS primitive states q (n, 5)."""
    rho, u, v, w, p = q.T
    vel = (u, v, w)
    V = vel[d]
    H = gamma / (gamma - 1) * p / rho + 0.5 * (u * u + v * v + w * w)
    f = np.empty_like(q)
    f[:, 0] = rho * V
    for m in range(3):
        f[:, 1 + m] = rho * vel[m] * V + (p if m == d else 0.0)
    f[:, 4] = rho * H * V
    return f


def states(rng, n, mach_lo=0.0, mach_hi=0.9, d=0, sign=1.0):
    s = np.empty((n, 5))
    s[:, 0] = rng.uniform(0.2, 2.0, n)
    s[:, 4] = rng.uniform(0.1, 2.0, n)
    c = np.sqrt(G * s[:, 4] / s[:, 0])
    s[:, 1:4] = rng.uniform(-0.3, 0.3, (n, 3)) * c[:, None]
    s[:, 1 + d] = sign * rng.uniform(mach_lo, mach_hi, n) * c
    return s


# ----------------------------------------------------------------- fluxes
@pytest.mark.parametrize("kind", [O.ROE, O.AUSM, O.AUSMPLUS])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("mach", [(0.0, 0.9), (1.1, 3.0)])
def test_flux_consistency(kind, d, mach):
    """F(q, q) = f(q) for Roe, the reference AUSM and AUSM+ at any Mach."""
    rng = np.random.default_rng(7 + d)
    q = states(rng, 200, *mach, d=d, sign=rng.choice([-1.0, 1.0]))
    got = O.flux(O.oracle(), kind, q, q, d)
    np.testing.assert_allclose(got, phys_flux(q, d), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_vanleer_consistency_and_supersonic_quirk(d):
    """The reference Van Leer split is consistent for |M| < 1; in the
    supersonic branch it keeps the subsonic momentum/energy factors
    (src/scheme/scheme_FVS_VanLeer.f90:82-95), so F(q, q) != f(q) there."""
    rng = np.random.default_rng(11)
    sub = states(rng, 100, 0.0, 0.95, d=d)
    np.testing.assert_allclose(O.flux(O.oracle(), O.VANLEER, sub, sub, d), phys_flux(sub, d),
                               rtol=1e-13, atol=1e-13)
    sup = states(rng, 100, 1.2, 3.0, d=d)
    got = O.flux(O.oracle(), O.VANLEER, sup, sup, d)
    rho, p = sup[:, 0], sup[:, 4]
    c = np.sqrt(G * p / rho)
    V = sup[:, 1 + d]
    mass = rho * c * (V / c)
    assert np.allclose(got[:, 0], mass, rtol=1e-13)            # mass flux is rho V
    mom = mass * ((-V + 2 * c) / G + V)                        # subsonic factor kept
    np.testing.assert_allclose(got[:, 1 + d], mom, rtol=1e-12)
    assert np.all(np.abs(got[:, 1 + d] - phys_flux(sup, d)[:, 1 + d]) > 1e-6)


@pytest.mark.parametrize("kind", [O.ROE, O.AUSM, O.AUSMPLUS])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_supersonic_upwinding(kind, d):
    """Both states supersonic in +e_d (every Roe eigenvalue above the Harten
    threshold): the face flux is the left physical flux."""
    rng = np.random.default_rng(13 + d)
    qL = states(rng, 100, 2.5, 3.5, d=d)
    qR = qL.copy()
    qR[:, 0] *= rng.uniform(0.95, 1.05, 100)
    qR[:, 4] *= rng.uniform(0.95, 1.05, 100)
    got = O.flux(O.oracle(), kind, qL, qR, d, entfix=0.2)
    np.testing.assert_allclose(got, phys_flux(qL, d), rtol=1e-11, atol=1e-12)


def test_roe_entropy_fix_threshold():
    """Harten fix: lambda <= delta -> (lambda^2 + delta^2) / (2 delta), which
    makes the dissipation continuous; with entfix = 0 the Roe flux of a
    stagnation face reduces to the central average minus |lambda| terms."""
    qL = np.array([[1.0, 0.0, 0.0, 0.0, 1.0]])
    qR = np.array([[0.5, 0.0, 0.0, 0.0, 0.4]])
    f_fix = O.flux(O.oracle(), O.ROE, qL, qR, 0, entfix=0.2)
    f_eps = O.flux(O.oracle(), O.ROE, qL, qR, 0, entfix=0.2 + 1e-9)
    np.testing.assert_allclose(f_fix, f_eps, rtol=1e-7)          # continuous in entfix
    # pressure-only jump at rest: mass flux comes from the acoustic waves only
    assert abs(f_fix[0, 1] - 0.5 * (1.0 + 0.4)) < 0.2


# ------------------------------------------------------- reconstruction
@pytest.mark.parametrize("lim", [O.LIM_NONE, O.LIM_MINMOD])
def test_muscl_exact_on_linear(lim):
    x = np.arange(30, dtype=float)
    pg = np.stack([1.0 + 0.1 * x, -0.3 + 0.05 * x, 0.2 * x, 0.01 * x, 2.0 + 0.02 * x], axis=1)
    slope = np.array([0.1, 0.05, 0.2, 0.01, 0.02])
    nface = 25
    L = O.recon(O.oracle(), O.MUSCL, pg, nface, 3, -1, kappa=0.33333, limiter=lim)
    R = O.recon(O.oracle(), O.MUSCL, pg, nface, 3, 1, kappa=0.33333, limiter=lim)
    f = np.arange(1, nface + 1)
    # face f (1-based) lies between ghost-mesh vertices f+2 and f+3 (0-based f+1, f+2)
    exact = pg[f + 1] + 0.5 * slope
    np.testing.assert_allclose(L, exact, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(R, exact, rtol=1e-14, atol=1e-14)


def test_muscl_limiters_clip_extrema():
    """At a local extremum every limiter returns the vertex value (TVD)."""
    pg = np.ones((12, 5))
    pg[6] = 2.0
    for lim in (O.LIM_MINMOD, O.LIM_VANLEER):
        L = O.recon(O.oracle(), O.MUSCL, pg, 7, 3, -1, limiter=lim)
        # face 5 (1-based) has its left state centred on 0-based vertex 6
        assert np.all(L[4] == 2.0)


def _cell_avg_sin(n):
    h = 2 * np.pi / n
    xf = h * np.arange(-3, n + 4)
    avg = (np.cos(xf[:-1]) - np.cos(xf[1:])) / h
    return h, xf, avg


@pytest.mark.parametrize("textbook,order", [(1, 4.5), (0, 0.9)])
def test_weno_order(textbook, order):
    """Textbook WENO5-JS is 5th order on smooth cell averages; the
    reference's IS_1 term (u1 + u3)^2 skews the weights and costs order, but
    the candidate stencils keep it consistent."""
    errs = []
    for n in (40, 80):
        h, xf, avg = _cell_avg_sin(n)
        qg = np.repeat(avg[:, None], 5, axis=1) + 2.0
        npg = qg.shape[0]
        nface = npg - 5
        L = O.recon(O.oracle(), O.WENO, qg, nface, 3, -1, textbook=textbook)[:, 0] - 2.0
        f = np.arange(1, nface + 1)
        exact = np.sin(xf[f + 2])
        errs.append(np.abs(L - exact)[5:-5].max())
    rate = math.log2(errs[0] / errs[1])
    assert rate > order, (errs, rate)


def test_weno_textbook_right_state_is_mirror():
    rng = np.random.default_rng(5)
    qg = rng.uniform(1.0, 2.0, (20, 5))
    R = O.recon(O.oracle(), O.WENO, qg, 15, 3, 1, textbook=1)
    Lm = O.recon(O.oracle(), O.WENO, qg[::-1].copy(), 15, 3, -1, textbook=1)
    # the mirrored left state at the mirrored face is the right state
    np.testing.assert_array_equal(R, Lm[::-1])


def _weno_js_published(u):
    """Jiang & Shu, JCP 126 (1996), eqs. (2.10), (3.2)-(3.4), in its published
    quotient form: alpha_k = d_k / (eps + IS_k)^2, w_k = alpha_k / sum alpha."""
    u0, u1, u2, u3, u4 = u
    q = (u0 / 3 - 7 * u1 / 6 + 11 * u2 / 6,
         -u1 / 6 + 5 * u2 / 6 + u3 / 3,
         u2 / 3 + 5 * u3 / 6 - u4 / 6)
    IS = (13 / 12 * (u0 - 2 * u1 + u2) ** 2 + (u0 - 4 * u1 + 3 * u2) ** 2 / 4,
          13 / 12 * (u1 - 2 * u2 + u3) ** 2 + (u1 - u3) ** 2 / 4,
          13 / 12 * (u2 - 2 * u3 + u4) ** 2 + (3 * u2 - 4 * u3 + u4) ** 2 / 4)
    a = [d / (1e-6 + s) ** 2 for d, s in zip((0.1, 0.6, 0.3), IS)]
    return sum(ai * qi for ai, qi in zip(a, q)) / sum(a)


def test_weno_textbook_matches_published_weights():
    """The textbook variant's product-form weights (one division per face)
    are Jiang-Shu's quotient form: random stencils with jumps of every size,
    so each weight matters somewhere; a weight paired with the wrong
    indicator product, the reference's (u1 + u3) term or an unsquared
    indicator each miss by far more than rounding."""
    rng = np.random.default_rng(11)
    n = 400
    base = rng.uniform(0.5, 2.0, n)
    jumps = rng.choice([0.0, 1e-3, 0.1, 1.0, 10.0], n) * (rng.uniform(size=n) < 0.3)
    col = base + np.cumsum(jumps)
    qg = np.repeat(col[:, None], 5, axis=1)
    nface = n - 5
    L = O.recon(O.oracle(), O.WENO, qg, nface, 3, -1, textbook=1)[:, 0]
    R = O.recon(O.oracle(), O.WENO, qg, nface, 3, 1, textbook=1)[:, 0]
    f = np.arange(1, nface + 1)
    ig = f + 1                       # vertex left of face f (0-based; 3 ghosts)
    wantL = np.array([_weno_js_published(col[i - 2:i + 3]) for i in ig])
    wantR = np.array([_weno_js_published(col[i - 1:i + 4][::-1]) for i in ig])
    tol = 1e-13 * np.abs(col).max()
    np.testing.assert_allclose(L, wantL, rtol=0, atol=tol)
    np.testing.assert_allclose(R, wantR, rtol=0, atol=tol)
    # not vacuous: the nonlinear weights move the values well away from the
    # linear (smooth-data) combination at the jumps
    lin = np.array([0.1 * (col[i - 2] / 3 - 7 * col[i - 1] / 6 + 11 * col[i] / 6)
                    + 0.6 * (-col[i - 1] / 6 + 5 * col[i] / 6 + col[i + 1] / 3)
                    + 0.3 * (col[i] / 3 + 5 * col[i + 1] / 6 - col[i + 2] / 6) for i in ig])
    assert np.abs(L - lin).max() > 0.1


def test_weno_reference_right_equals_next_left():
    """scheme_WENO.f90:54-56: the right state of face f is the left state of f+1."""
    rng = np.random.default_rng(6)
    qg = rng.uniform(1.0, 2.0, (20, 5))
    L = O.recon(O.oracle(), O.WENO, qg, 15, 3, -1, textbook=0)
    R = O.recon(O.oracle(), O.WENO, qg, 15, 3, 1, textbook=0)
    np.testing.assert_array_equal(R[:-1], L[1:])


# ------------------------------------------------------------ time stepping
@pytest.mark.parametrize("z", [-2.5, -1.0, -0.3, 0.2, 0.5])
def test_ssp_rk3_stability_polynomial(z):
    u0 = np.array([1.0])
    um = u0.copy()
    for m in (1, 2, 3):
        _, um = O.runge_kutta(O.oracle(), m, 1.0, u0, um, z * um)
    assert abs(um[0] - (1 + z + z * z / 2 + z ** 3 / 6)) < 1e-14


@pytest.mark.parametrize("z", [-2.7, -1.0, -0.3, 0.2, 0.5])
def test_rk4_stability_polynomial(z):
    u0 = np.array([1.0])
    um, acc = u0.copy(), np.zeros(1)
    for s in (1, 2, 3, 4):
        um, acc = O.rk4_stage(s, 1.0, u0, um, acc, z * um)
    assert abs(um[0] - (1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24)) < 1e-14


def test_cfl_dt_closed_form():
    """Uniform state with c = 1: dt = cfl / ((|u|+1)/dx + (|v|+1)/dy + (|w|+1)/dz)."""
    n = (5, 4, 3)
    p = np.zeros((n[2], n[1], n[0], 5))
    p[...] = [1.0, 1.0, -0.5, 0.25, 1.0 / G]
    q = cases.to_conserved(p)
    dr = np.array([0.5, 0.25, 0.125])
    dt = O.cfl_dt(q, 3, dr, cfl=0.8)
    assert abs(dt - 0.8 / (2.0 / 0.5 + 1.5 / 0.25 + 1.25 / 0.125)) < 1e-15
    dt2 = O.cfl_dt(q[:, :1], 2, dr, cfl=0.8)
    assert abs(dt2 - 0.8 / (2.0 / 0.5 + 1.5 / 0.25)) < 1e-15


# --------------------------------------------------------- whole-block
COMBOS = [(O.SIMPLE, O.LIM_NONE, O.ROE), (O.MUSCL, O.LIM_MINMOD, O.ROE),
          (O.MUSCL, O.LIM_VANALBADA, O.VANLEER), (O.WENO, O.LIM_NONE, O.AUSM),
          (O.MUSCL, O.LIM_VANLEER, O.AUSMPLUS)]


@pytest.mark.parametrize("recon,lim,flux", COMBOS)
@pytest.mark.parametrize("bc", [O.BC_EXTRAP, O.BC_PERIODIC])
def test_free_stream_preserved_bitwise(recon, lim, flux, bc):
    p = np.zeros((6, 7, 8, 5))
    p[...] = [1.3, 0.4, -0.2, 0.1, 0.9]
    r = O.oracle_solve(p, 3, (0.1, 0.2, 0.3), recon=recon, limiter=lim, flux=flux, rk=O.RK4,
                       dt=1e-3, nsteps=3, bc=(bc,) * 6)
    q0 = cases.to_conserved(p)
    assert np.array_equal(r["q"][:, 3:-3, 3:-3, 3:-3], q0)
    assert np.all(r["res"] == 0.0)


@pytest.mark.parametrize("recon,lim,flux", COMBOS)
@pytest.mark.parametrize("rk", [O.RK_SSP3, O.RK4])
def test_periodic_conservation(recon, lim, flux, rk):
    """Nodal periodic box: face n of the last vertex is face 0 of the first,
    so the flux differences telescope and the totals are conserved."""
    p = cases.smooth_3d((8, 9, 10), seed=2)
    for m in (1, 2, 3):
        p[..., m] = p[..., m] - p[..., m].mean()
    r = O.oracle_solve(p, 3, (0.1, 0.1, 0.1), recon=recon, limiter=lim, flux=flux, rk=rk,
                       dt=0.0, cfl=0.5, nsteps=4, bc=(O.BC_PERIODIC,) * 6)
    q0 = cases.to_conserved(p)
    q1 = r["q"][:, 3:-3, 3:-3, 3:-3]
    for c in range(5):
        a, b = math.fsum(q0[c].ravel()), math.fsum(q1[c].ravel())
        scale = math.fsum(np.abs(q0[c]).ravel()) + 1.0
        assert abs(a - b) / scale < 2e-14, (c, a, b)
    # periodic ghost images hold the wrapped vertex
    q = r["q"]
    np.testing.assert_array_equal(q[:, 3:-3, 3:-3, 0], q[:, 3:-3, 3:-3, -6])
    np.testing.assert_array_equal(q[:, 3:-3, 3:-3, -1], q[:, 3:-3, 3:-3, 5])
    np.testing.assert_array_equal(q[:, 0, 3:-3, 3:-3], q[:, -6, 3:-3, 3:-3])


def test_extrapolation_ghosts_copy_boundary():
    p = cases.smooth_3d((6, 7, 8), seed=4)
    r = O.oracle_solve(p, 3, (0.1, 0.1, 0.1), dt=1e-3, nsteps=2)
    q = r["q"]
    for l in range(3):
        np.testing.assert_array_equal(q[:, 3:-3, 3:-3, l], q[:, 3:-3, 3:-3, 3])
        np.testing.assert_array_equal(q[:, 3:-3, l, 3:-3], q[:, 3:-3, 3, 3:-3])
        np.testing.assert_array_equal(q[:, -1 - l, 3:-3, 3:-3], q[:, -4, 3:-3, 3:-3])


def test_riemann2d_transpose_symmetry():
    """The reference's quadrant states are symmetric under x <-> y with
    u <-> v except on the x = y = 0.8 corner vertex, which the region order
    assigns to region 4; on a 40 x 40 grid 0.8 is not a vertex coordinate.
    The x-then-y sweep order breaks the symmetry only at round-off."""
    p, dx = cases.riemann2d(40, seed=0)
    r = O.oracle_solve(p, 2, (dx, dx, 1.0), dt=2e-4, nsteps=20)
    q = r["q"][:, 0, 3:-3, 3:-3]
    qt = q.transpose(0, 2, 1)[[0, 2, 1, 3, 4]]
    assert np.abs(q - qt).max() < 1e-12 * np.abs(q).max()


def test_vortex_second_order_convergence():
    """MUSCL-vanAlbada + Roe + RK4 + periodic + CFL on Yee's vortex (the
    3-D benchmark's scheme, run on its 2-D cross-section): the error
    against the exact translated vortex falls at second order."""
    errs = []
    for n in (24, 48):
        L = 10.0
        h = L / n
        x = h * np.arange(n)
        p0 = cases.vortex_2d(x, x, (5.0, 5.0), L, 5.0, G, (1.0, 1.0, 0.0))[None]
        t_end, t = 1.0, 0.0
        # fixed dt that divides t_end, CFL ~ 0.4
        nsteps = int(math.ceil(t_end / (0.4 * h / (1.0 + 1.0 + 2 * 1.2))))
        r = O.oracle_solve(p0, 2, (h, h, 1.0), recon=O.MUSCL, limiter=O.LIM_VANALBADA,
                           flux=O.ROE, rk=O.RK4, dt=t_end / nsteps, nsteps=nsteps,
                           bc=(O.BC_PERIODIC,) * 4 + (0, 0))
        q = r["q"][:, 0, 3:-3, 3:-3]
        exact = cases.to_conserved(cases.vortex_2d(x, x, (5.0, 5.0), L, 5.0, G, (1.0, 1.0, 0.0),
                                                   t=t_end)[None])[:, 0]
        errs.append(np.abs(q[0] - exact[0]).mean())
    rate = math.log2(errs[0] / errs[1])
    assert rate > 1.6, (errs, rate)


@pytest.mark.parametrize("rk,cfl", [(O.RK4, 0.8), (O.RK_SSP3, 0.5)])
def test_solve_from_conserved_state_continues_a_run(rk, cfl):
    """orc_solve_q from an earlier run's final conserved state continues that
    run bit for bit (chunked long runs), CFL dt included; the state is the
    run's whole memory, so 3 + 4 + 3 steps equal 10 steps."""
    p, h, _ = cases.isentropic_vortex(12, seed=2)
    kw = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=rk, cfl=cfl, dt=0.0,
              bc=(O.BC_PERIODIC,) * 6)
    whole = O.oracle_solve(p, 3, (h, h, h), nsteps=10, p_from_q=1, **kw)
    first = O.oracle_solve(p, 3, (h, h, h), nsteps=3, p_from_q=1, **kw)
    q, dts = first["q"][:, 3:-3, 3:-3, 3:-3], list(first["dt"])
    for n in (4, 3):
        r = O.oracle_solve_q(q, 3, (h, h, h), nsteps=n, **kw)
        q = r["q_real"]
        dts += list(r["dt"])
    assert np.array_equal(q, whole["q"][:, 3:-3, 3:-3, 3:-3])
    assert dts == list(whole["dt"])
    assert np.array_equal(r["res"][-1], whole["res"][-1])
