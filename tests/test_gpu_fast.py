"""The fast kernel family (FMA contraction, refined hardware reciprocals)
against the oracle and the exact family.  Tolerances are the stated ones:
relative L-inf per conserved component (normalised by max|q| with a floor of
1) at most 1e-12 after one step and at most 1e-9 after 1000 steps."""
import numpy as np
import pytest

import _oracle as O
from _parity import rel_linf, run_gpu, run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

COMBOS = [(O.SIMPLE, O.LIM_NONE, O.ROE), (O.MUSCL, O.LIM_MINMOD, O.ROE),
          (O.MUSCL, O.LIM_VANLEER, O.VANLEER), (O.MUSCL, O.LIM_VANALBADA, O.ROE),
          (O.WENO, O.LIM_NONE, O.AUSM), (O.MUSCL, O.LIM_VANALBADA, O.AUSMPLUS)]


@pytest.mark.parametrize("tile", ["", "4"])
@pytest.mark.parametrize("recon,lim,flux", COMBOS)
def test_fast_one_step_3d(recon, lim, flux, tile, monkeypatch):
    if tile:   # the register-reconstruction kernel (WENO falls back to the default)
        monkeypatch.setenv("FVM_TILE", tile)
    p = cases.smooth_3d((12, 14, 40), seed=5)
    sp = (1 / 39, 1 / 13, 1 / 11)
    print(f"case fast 1-step 3d recon={recon} lim={lim} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, _, _, _ = run_gpu(ctx, p, sp, 1, dt=2e-3, numerics=fvmgpu.NUMERICS_FAST,
                              recon=recon, limiter=lim, flux=flux, rk=O.RK4)
    qo, _ = run_oracle(p, sp, 1, dt=2e-3, recon=recon, limiter=lim, flux=flux, rk=O.RK4)
    err = rel_linf(qg, qo)
    print("rel linf", err)
    assert err <= 1e-12


@pytest.mark.parametrize("tile2", ["0", "1"])
@pytest.mark.parametrize("recon,lim,flux", COMBOS[:4])
def test_fast_one_step_riemann2d(recon, lim, flux, tile2, monkeypatch):
    monkeypatch.setenv("FVM_TILE2", tile2)
    p, dx = cases.riemann2d(101, seed=0)
    print(f"case fast 1-step riemann2d recon={recon} lim={lim} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, _, _, _ = run_gpu(ctx, p, (dx, dx, 1.0), 1, dt=1e-3, numerics=fvmgpu.NUMERICS_FAST,
                              recon=recon, limiter=lim, flux=flux)
    qo, _ = run_oracle(p, (dx, dx, 1.0), 1, dt=1e-3, recon=recon, limiter=lim, flux=flux)
    err = rel_linf(qg, qo)
    print("rel linf", err)
    assert err <= 1e-12


def test_fast_1000_steps_vortex():
    """The benchmark's scheme (MUSCL-vanAlbada + Roe, RK4, periodic, CFL 0.8)
    on the isentropic vortex, 1000 steps: fast vs exact.  The exact family is
    bit-identical to the oracle step by step (test_gpu_parity), so this bounds
    the fast family's distance from the oracle at 1000 steps."""
    p, h, _ = cases.isentropic_vortex(48, seed=0)
    print("case fast vs exact 1000 steps vortex48", flush=True)
    out = {}
    for mode in (fvmgpu.NUMERICS_EXACT, fvmgpu.NUMERICS_FAST):
        with fvmgpu.Context() as ctx:
            out[mode] = run_gpu(ctx, p, (h, h, h), 1000, cfl=0.8, bc=(1,) * 6, numerics=mode,
                                recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    err = rel_linf(out[1][0], out[0][0])
    print("rel linf after 1000 steps", err)
    assert err <= 1e-9
    # conservation on the periodic box, to round-off, in both families
    q0 = cases.to_conserved(p)
    for mode in out:
        for c in range(5):
            assert abs(out[mode][0][c].sum() - q0[c].sum()) <= 1e-11 * np.abs(q0[c]).sum() + 1e-12
