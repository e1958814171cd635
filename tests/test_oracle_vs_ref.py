"""The oracle against the reference's own scheme modules on fresh seeded
inputs (needs oracle/_ref/libfvmref.so, built where the reference sources
are present; skipped elsewhere)."""
import numpy as np
import pytest

import _oracle as O

pytestmark = pytest.mark.skipif(O.ref() is None, reason="reference harness not built here")


def states(rng, n, vmax=1.5):
    s = np.empty((n, 5))
    s[:, 0] = rng.uniform(0.05, 3.0, n)
    s[:, 1:4] = rng.uniform(-vmax, vmax, (n, 3))
    s[:, 4] = rng.uniform(0.01, 2.0, n)
    return s


@pytest.mark.parametrize("seed", range(5))
def test_per_function_random(seed):
    rng = np.random.default_rng(1000 + seed)
    pg = states(rng, 40)
    nface = 35
    orc, ref = O.oracle(), O.ref()
    for w in (-1, 1):
        for lim in (O.LIM_NONE, O.LIM_MINMOD, O.LIM_VANLEER, O.LIM_VANALBADA):
            kap = rng.uniform(-1, 1)
            a = O.recon(orc, O.MUSCL, pg, nface, 3, w, kappa=kap, limiter=lim)
            b = O.recon(ref, O.MUSCL, pg, nface, 3, w, kappa=kap, limiter=lim)
            assert np.array_equal(a, b)
        assert np.array_equal(O.recon(orc, O.WENO, np.abs(pg), nface, 3, w),
                              O.recon(ref, O.WENO, np.abs(pg), nface, 3, w))
    ql, qr = states(rng, 500, vmax=3.0), states(rng, 500, vmax=3.0)
    for kind in (O.ROE, O.VANLEER, O.AUSM):
        for d in range(3):
            g, e = rng.uniform(1.1, 1.7), rng.uniform(0.0, 0.5)
            assert np.array_equal(O.flux(orc, kind, ql, qr, d, gamma=g, entfix=e),
                                  O.flux(ref, kind, ql, qr, d, gamma=g, entfix=e))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_riemann_random_quadrants(seed):
    """Seeded quadrant states (rho in [0.1,2], p in [0.02,1.5], |u|,|v| <= 1.3)."""
    import fvmgpu.cases as cases
    p1, dx = cases.riemann2d(25, seed=seed)
    p = np.concatenate([p1, p1], axis=0)
    for rc, lim, fl in ((O.MUSCL, O.LIM_MINMOD, O.ROE), (O.WENO, O.LIM_NONE, O.AUSM),
                        (O.MUSCL, O.LIM_VANALBADA, O.VANLEER)):
        r = O.ref_solve(p, 2, (dx, dx, 1.0), recon=rc, limiter=lim, flux=fl, dt=1e-4, nsteps=5,
                        p_from_q=1)
        o = O.oracle_solve(p, 2, (dx, dx, 1.0), recon=rc, limiter=lim, flux=fl, dt=1e-4,
                           nsteps=5, p_from_q=1)
        assert np.array_equal(r["q"][:, 3, 3:-3, 3:-3], o["q"][:, 0, 3:-3, 3:-3])
        assert np.array_equal(r["res"], o["res"])
