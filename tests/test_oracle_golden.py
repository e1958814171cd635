"""The oracle against the golden fixtures written from the reference's own
compiled scheme modules (tests/golden/make_golden.py).  Bit-exact."""
import os

import numpy as np
import pytest

import _oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def sv():
    return np.load(os.path.join(GOLDEN, "scheme_vectors.npz"))


@pytest.fixture(scope="module")
def bs():
    return np.load(os.path.join(GOLDEN, "block_states.npz"))


@pytest.mark.parametrize("pname", ["rand", "edge"])
@pytest.mark.parametrize("wind", [-1, 1])
def test_reconstruction_vectors(sv, pname, wind):
    orc = O.oracle()
    pg = sv[f"pencil_{pname}"]
    nface = pg.shape[0] - 5
    assert np.array_equal(O.recon(orc, O.SIMPLE, pg, nface, 3, wind), sv[f"simple_{pname}_{wind}"])
    for lim in (O.LIM_NONE, O.LIM_MINMOD, O.LIM_VANLEER, O.LIM_VANALBADA):
        got = O.recon(orc, O.MUSCL, pg, nface, 3, wind, kappa=0.33333, limiter=lim)
        want = sv[f"muscl_{pname}_{lim}_{wind}"]
        # signed zeros included: compare the bit patterns
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), lim
    qg = sv[f"wenoin_{pname}"]
    assert np.array_equal(O.recon(orc, O.WENO, qg, nface, 3, wind), sv[f"weno_{pname}_{wind}"])


@pytest.mark.parametrize("kind,name", [(O.ROE, "roe"), (O.VANLEER, "vanleer"), (O.AUSM, "ausm")])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_flux_vectors(sv, kind, name, d):
    got = O.flux(O.oracle(), kind, sv["flux_ql"], sv["flux_qr"], d, gamma=1.4, entfix=0.2)
    assert np.array_equal(got, sv[f"flux_{name}_{d}"])


def test_rk_vectors(sv):
    for m in (1, 2, 3):
        got = O.runge_kutta(O.oracle(), m, 2e-4, sv["rk_u0"], sv["rk_um"], sv["rk_r"])[1]
        assert np.array_equal(got, sv[f"rk_{m}"])


COMBOS = [(O.SIMPLE, O.LIM_NONE), (O.MUSCL, O.LIM_MINMOD), (O.MUSCL, O.LIM_VANLEER),
          (O.MUSCL, O.LIM_VANALBADA), (O.WENO, O.LIM_NONE)]


@pytest.mark.parametrize("recon,lim", COMBOS)
@pytest.mark.parametrize("flux", [O.ROE, O.VANLEER, O.AUSM])
@pytest.mark.parametrize("steps", [1, 10])
def test_riemann21_states(bs, recon, lim, flux, steps):
    p = bs["riemann21_p"]
    dx = float(bs["riemann21_dx"])
    modes = (0, 1) if (recon, lim, flux) == (O.MUSCL, O.LIM_MINMOD, O.ROE) else (0,)
    for pfq in modes:
        r = O.oracle_solve(p, 2, (dx, dx, 1.0), recon=recon, limiter=lim, flux=flux,
                           kappa=0.33333, entfix=0.2, dt=2e-4, nsteps=steps, p_from_q=pfq)
        key = f"riemann21_{recon}_{lim}_{flux}_{steps}_{pfq}"
        assert np.array_equal(r["q"][:, 0, 3:-3, 3:-3], bs[key + "_q"])
        assert np.array_equal(r["res"], bs[key + "_res"])


def test_riemann64_state(bs):
    p = bs["riemann64_p"]
    dx = 1.0 / 63
    r = O.oracle_solve(p, 2, (dx, dx, 1.0), dt=2e-4, nsteps=10, p_from_q=0)
    assert np.array_equal(r["q"][:, 0, 3:-3, 3:-3], bs["riemann64_q"])
    assert np.array_equal(r["res"], bs["riemann64_res"])


@pytest.mark.parametrize("recon,lim", COMBOS)
@pytest.mark.parametrize("flux", [O.ROE, O.VANLEER, O.AUSM])
def test_smooth3d_states(bs, recon, lim, flux):
    p = bs["smooth3d_p"]
    dr = tuple(bs["smooth3d_dr"])
    r = O.oracle_solve(p, 3, dr, recon=recon, limiter=lim, flux=flux, dt=1e-3, nsteps=3,
                       p_from_q=1)
    key = f"smooth3d_{recon}_{lim}_{flux}"
    assert np.array_equal(r["q"][:, 3:-3, 3:-3, 3:-3], bs[key + "_q"])
    assert np.array_equal(r["res"], bs[key + "_res"])
