"""The fast kernel family (FMA contraction, refined hardware reciprocals)
against the oracle and the exact family.  Tolerances are the stated ones:
relative L-inf per conserved component (normalised by max|q| with a floor of
1) at most 1e-12 after one step and at most 1e-9 after 1000 steps."""
import os

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


def _vortex128_run(mode, q0, h, marks):
    out, done = [], 0
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
        ctx.numerics(mode)
        b = ctx.block((128, 128, 128), (h, h, h))
        b.set_periodic()
        b.upload(q0)
        for m in marks:
            ctx.step(m - done, cfl=0.8)
            done = m
            out.append(b.download())
    return out


def test_vortex128_amplifies_rounding_differences():
    """The benchmark workload (128^3 vortex, MUSCL-vanAlbada + Roe, RK4, CFL
    0.8) amplifies any rounding-level difference: the Yee vortex's vorticity
    changes sign at r = sqrt 2 (a shielded vortex, unstable to azimuthal
    perturbations), and at 128^3 the scheme's dissipation no longer damps it.
    Measured: the exact family (bit for bit the oracle) started from a state
    one ulp away in rho*E differs from itself by more than 1e-10 after 256
    steps, and the fast family (FMA contraction, refined reciprocals) differs
    from the exact family by no more than that same exact-family spread at
    1000 steps.  So at this size the 1000-step 1e-9 gate is met only by
    arithmetic bit-identical to the oracle: the exact family, which is what
    the benchmark reports; the fast family is as close as the problem's own
    conditioning under one-ulp input changes allows."""
    p, h, _ = cases.isentropic_vortex(128, seed=0)
    q0 = cases.to_conserved(p)
    q1 = q0.copy()
    q1[4] = np.nextafter(q1[4], np.inf)
    marks = [16, 256, 1000]
    print("case vortex128 conditioning: exact, exact+1ulp, fast; 1000 steps", flush=True)
    ex = _vortex128_run(fvmgpu.NUMERICS_EXACT, q0, h, marks)
    ex1 = _vortex128_run(fvmgpu.NUMERICS_EXACT, q1, h, marks)
    fa = _vortex128_run(fvmgpu.NUMERICS_FAST, q0, h, marks)
    spread = [rel_linf(a, b) for a, b in zip(ex1, ex)]
    fast = [rel_linf(a, b) for a, b in zip(fa, ex)]
    print("exact one-ulp spread", spread, "fast vs exact", fast)
    assert spread[0] < 1e-13 and spread[1] > 1e-10            # exponential growth
    assert fast[0] < 1e-13
    assert fast[2] <= max(spread[1:]) * 1.5 + 1e-12
    # conservation to round-off in both families
    for q in (ex[2], fa[2]):
        for c in range(5):
            assert abs(q[c].sum() - q0[c].sum()) <= 1e-11 * np.abs(q0[c]).sum() + 1e-12


def test_exact_1000_steps_equals_oracle():
    """The exact family against the oracle over the full 1000 steps (not only
    step by step), on a grid the host oracle runs in seconds: the 32^3 vortex
    with the benchmark's scheme and CFL steps; bit for bit, including every
    CFL dt."""
    p, h, _ = cases.isentropic_vortex(32, seed=7)
    print("case exact vs oracle 1000 steps vortex32", flush=True)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    with fvmgpu.Context() as ctx:
        qg, norms, _, dt_last = run_gpu(ctx, p, (h, h, h), 1000, cfl=0.8, bc=(1,) * 6, **scheme)
    O.set_threads(16)
    qo, r = run_oracle(p, (h, h, h), 1000, cfl=0.8, bc=(1,) * 6, **scheme)
    print("rel linf", rel_linf(qg, qo))
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(norms, r["res"][-1], rtol=1e-12, atol=1e-300)


@pytest.mark.timeout(1200)
@pytest.mark.skipif(not os.environ.get("FVM_LONG_TESTS"),
                    reason="runs the host oracle over 1000 steps of a 128^3 grid (about 6 minutes "
                           "on 16 threads); FVM_LONG_TESTS=1 runs it")
def test_exact_1000_steps_equals_oracle_vortex128():
    """The benchmark workload itself, the full run: the exact family on the
    128^3 vortex (MUSCL-vanAlbada + Roe, RK4, CFL 0.8) equals the oracle bit
    for bit after every 100 of its 1000 steps, CFL dt included.  The oracle
    runs in 25-step chunks (orc_solve_q continues a run bit for bit) so the
    run reports its progress."""
    p, h, _ = cases.isentropic_vortex(128, seed=0)
    q0 = cases.to_conserved(p)
    marks = [100 * m for m in range(1, 11)]
    print("case exact vs oracle 1000 steps vortex128", flush=True)
    gpu, gpu_dt = [], []
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
        ctx.numerics(fvmgpu.NUMERICS_EXACT)
        b = ctx.block((128, 128, 128), (h, h, h))
        b.set_periodic()
        b.upload(q0)
        done = 0
        for m in marks:
            gpu_dt.append(ctx.step(m - done, cfl=0.8))
            done = m
            gpu.append(b.download())
    O.set_threads(min(16, os.cpu_count() or 1))
    kw = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4, cfl=0.8, dt=0.0,
              bc=(O.BC_PERIODIC,) * 6)
    q, chunk = None, 25
    for mark, qg, dtg in zip(marks, gpu, gpu_dt):
        for s in range(mark - 100, mark, chunk):
            if q is None:
                r = O.oracle_solve(p, 3, (h, h, h), nsteps=chunk, p_from_q=1, **kw)
                q = r["q"][:, 3:-3, 3:-3, 3:-3].copy()
            else:
                r = O.oracle_solve_q(q, 3, (h, h, h), nsteps=chunk, **kw)
                q = r["q_real"]
            print(f"oracle step {s + chunk} of 1000", flush=True)
        assert np.array_equal(qg, q), f"step {mark}: rel linf {rel_linf(qg, q)}"
        assert dtg == r["dt"][-1]
        print(f"step {mark}: bit-identical, dt {dtg!r}", flush=True)
