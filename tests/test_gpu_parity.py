"""GPU path vs the oracle on the same seeded inputs, through the C ABI.

The kernels keep the reference's IEEE operation order and are built without
FMA contraction, so the expected result is bit-identical: the assertion is
exact equality of every conserved value, with the relative L-inf printed."""
import os

import numpy as np
import pytest

import _oracle as O
from _parity import rel_linf, run_gpu, run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

RECONS = [(O.SIMPLE, O.LIM_NONE), (O.MUSCL, O.LIM_NONE), (O.MUSCL, O.LIM_MINMOD),
          (O.MUSCL, O.LIM_VANLEER), (O.MUSCL, O.LIM_VANALBADA), (O.WENO, O.LIM_NONE)]
FLUXES = [O.ROE, O.VANLEER, O.AUSM]


# unlimited MUSCL is unstable at the Riemann problem's shocks (both paths
# produce NaN); it is covered by the smooth 3-D case
@pytest.mark.parametrize("recon,lim", [r for r in RECONS if r != (O.MUSCL, O.LIM_NONE)])
@pytest.mark.parametrize("flux", FLUXES)
def test_riemann2d_bitwise(recon, lim, flux):
    p, dx = cases.riemann2d(41, seed=0)
    print(f"case riemann2d n=41 recon={recon} lim={lim} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, _ = run_gpu(ctx, p, (dx, dx, 1.0), 10, dt=2e-4, recon=recon, limiter=lim,
                               flux=flux)
    qo, r = run_oracle(p, (dx, dx, 1.0), 10, dt=2e-4, recon=recon, limiter=lim, flux=flux)
    print("rel linf", rel_linf(qg, qo))
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(ng, r["res"][-1], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("recon,lim", RECONS)
@pytest.mark.parametrize("flux", FLUXES + [O.AUSMPLUS])
def test_smooth3d_bitwise(recon, lim, flux):
    p = cases.smooth_3d((11, 13, 70), seed=3)
    sp = (1 / 69, 1 / 12, 1 / 10)
    print(f"case smooth3d 70x13x11 recon={recon} lim={lim} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, _ = run_gpu(ctx, p, sp, 3, dt=1e-3, recon=recon, limiter=lim, flux=flux)
    qo, r = run_oracle(p, sp, 3, dt=1e-3, recon=recon, limiter=lim, flux=flux)
    print("rel linf", rel_linf(qg, qo))
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(ng, r["res"][-1], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("tile", ["0", "1", "2", "3", "4", "5"])
@pytest.mark.parametrize("chunk", ["3", "8", "32"])
def test_tile_shapes_bitwise(tile, chunk, monkeypatch):
    """Every compiled tile shape and k-chunk gives the same bits (partial
    tiles in i, j and k included); shapes 4 and 5 are the
    register-reconstruction kernel."""
    monkeypatch.setenv("FVM_TILE", tile)
    monkeypatch.setenv("FVM_CHUNK", chunk)
    p = cases.smooth_3d((19, 11, 67), seed=9)
    sp = (1 / 66, 1 / 10, 1 / 18)
    print(f"case tiles tile={tile} chunk={chunk} 67x11x19", flush=True)
    with fvmgpu.Context() as ctx:
        qg, _, _, _ = run_gpu(ctx, p, sp, 2, dt=1e-3, bc=(1, 1, 0, 0, 1, 1), recon=O.MUSCL,
                              limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    qo, _ = run_oracle(p, sp, 2, dt=1e-3, bc=(1, 1, 0, 0, 1, 1), recon=O.MUSCL,
                       limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    assert np.array_equal(qg, qo)


@pytest.mark.parametrize("tile", ["1", "3"])
@pytest.mark.parametrize("chunk", ["6", "10", "12"])
def test_weno_chunks_bitwise(tile, chunk, monkeypatch):
    """The autotuner's k-chunks that are not powers of two, with the WENO5
    stencil (five-plane k-queue) and a k extent that leaves a partial last
    chunk, give the oracle's bits."""
    monkeypatch.setenv("FVM_TILE", tile)
    monkeypatch.setenv("FVM_CHUNK", chunk)
    p = cases.smooth_3d((29, 13, 37), seed=4)
    sp = (1 / 36, 1 / 12, 1 / 28)
    print(f"case weno chunks tile={tile} chunk={chunk} 37x13x29", flush=True)
    sch = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3, ref_compat=0)
    with fvmgpu.Context() as ctx:
        qg, _, _, _ = run_gpu(ctx, p, sp, 2, dt=1e-3, bc=(1, 1, 0, 0, 1, 1), **sch)
    qo, _ = run_oracle(p, sp, 2, dt=1e-3, bc=(1, 1, 0, 0, 1, 1), **sch)
    assert np.array_equal(qg, qo)


@pytest.mark.parametrize("tile2", ["1", "2"])
@pytest.mark.parametrize("recon,lim,flux", [(O.SIMPLE, O.LIM_NONE, O.ROE),
                                            (O.MUSCL, O.LIM_MINMOD, O.ROE),
                                            (O.MUSCL, O.LIM_VANALBADA, O.VANLEER),
                                            (O.WENO, O.LIM_NONE, O.AUSM)])
def test_riemann2d_tiles_bitwise(tile2, recon, lim, flux, monkeypatch):
    """The tiled 2-D stage kernel (both shapes; 41 vertices leave partial
    tiles in i and j) gives the oracle's bits, as the row kernel does."""
    monkeypatch.setenv("FVM_TILE2", tile2)
    p, dx = cases.riemann2d(41, seed=0)
    print(f"case riemann2d tiled {tile2} recon={recon} lim={lim} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, _ = run_gpu(ctx, p, (dx, dx, 1.0), 10, dt=2e-4, recon=recon, limiter=lim,
                               flux=flux)
    qo, r = run_oracle(p, (dx, dx, 1.0), 10, dt=2e-4, recon=recon, limiter=lim, flux=flux)
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(ng, r["res"][-1], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("tile2", ["0", "1", "2"])
def test_riemann2d_periodic_cfl_bitwise(tile2, monkeypatch):
    """2-D CFL steps (dt from the device slot inside the first stage) with a
    periodic i direction, every 2-D launch shape, against the oracle."""
    monkeypatch.setenv("FVM_TILE2", tile2)
    p, dx = cases.riemann2d(37, seed=0)
    bc = (1, 1, 0, 0, 0, 0)
    print(f"case riemann2d periodic cfl tile {tile2}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, dt_last = run_gpu(ctx, p, (dx, dx, 1.0), 6, cfl=0.5, bc=bc)
    qo, r = run_oracle(p, (dx, dx, 1.0), 6, cfl=0.5, bc=bc)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(ng, r["res"][-1], rtol=1e-12)


@pytest.mark.parametrize("recon,lim,flux", [(O.MUSCL, O.LIM_VANALBADA, O.ROE),
                                            (O.WENO, O.LIM_NONE, O.AUSM)])
def test_rk4_periodic_cfl_bitwise(recon, lim, flux):
    """RK4 + periodic ghosts + the device CFL step, against the oracle."""
    p, h, _ = cases.isentropic_vortex(24, seed=1)
    print(f"case vortex24 rk4 cfl recon={recon} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, dt_last = run_gpu(ctx, p, (h, h, h), 5, cfl=0.8, bc=(1,) * 6, recon=recon,
                                     limiter=lim, flux=flux, rk=O.RK4)
    qo, r = run_oracle(p, (h, h, h), 5, cfl=0.8, bc=(1,) * 6, recon=recon, limiter=lim,
                       flux=flux, rk=O.RK4)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(ng, r["res"][-1], rtol=1e-12)


@pytest.mark.parametrize("tile", ["4", "5"])
@pytest.mark.parametrize("recon,lim,flux", [(O.SIMPLE, O.LIM_NONE, O.AUSM),
                                            (O.MUSCL, O.LIM_MINMOD, O.VANLEER),
                                            (O.MUSCL, O.LIM_NONE, O.AUSMPLUS),
                                            (O.MUSCL, O.LIM_VANLEER, O.ROE)])
def test_register_kernel_schemes_bitwise(tile, recon, lim, flux, monkeypatch):
    """The register-reconstruction kernel with every flux and limiter kind,
    CFL steps (its in-kernel wave-speed maximum and norms included)."""
    monkeypatch.setenv("FVM_TILE", tile)
    p = cases.smooth_3d((13, 21, 37), seed=17)
    sp = (1 / 36, 1 / 20, 1 / 12)
    print(f"case register kernel tile={tile} recon={recon} lim={lim} flux={flux}", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, dt_last = run_gpu(ctx, p, sp, 3, cfl=0.6, recon=recon, limiter=lim, flux=flux)
    qo, r = run_oracle(p, sp, 3, cfl=0.6, recon=recon, limiter=lim, flux=flux)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)
    np.testing.assert_allclose(ng, r["res"][-1], rtol=1e-12, atol=1e-300)


def test_register_tile_refused_for_weno(monkeypatch):
    """A WENO context asked for a register-reconstruction shape launches the
    default shape (same bits as the oracle)."""
    monkeypatch.setenv("FVM_TILE", "4")
    p = cases.smooth_3d((10, 12, 20), seed=2)
    sp = (1 / 19, 1 / 11, 1 / 9)
    with fvmgpu.Context() as ctx:
        qg, _, _, _ = run_gpu(ctx, p, sp, 2, dt=1e-3, recon=O.WENO, limiter=O.LIM_NONE, flux=O.ROE)
    qo, _ = run_oracle(p, sp, 2, dt=1e-3, recon=O.WENO, limiter=O.LIM_NONE, flux=O.ROE)
    assert np.array_equal(qg, qo)


def test_vortex128_full_size_bitwise():
    """The benchmark's workload at its full size (128^3 isentropic vortex,
    MUSCL-vanAlbada + Roe, RK4, periodic, CFL 0.8), exact family: two CFL
    steps equal the oracle's bit for bit (the oracle runs on the host's
    threads: a few seconds), and mass, momentum and energy are conserved on
    the periodic box to round-off."""
    p, h, _ = cases.isentropic_vortex(128, seed=0)
    print("case vortex128 full size, 2 CFL steps", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, dt_last = run_gpu(ctx, p, (h, h, h), 2, cfl=0.8, bc=(1,) * 6,
                                     recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1)
    qo, r = run_oracle(p, (h, h, h), 2, cfl=0.8, bc=(1,) * 6, recon=O.MUSCL,
                       limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)
    q0 = cases.to_conserved(p)
    for c in range(5):
        assert abs(qg[c].sum() - q0[c].sum()) <= 1e-12 * np.abs(q0[c]).sum() + 1e-12


def test_riemann400_full_size_bitwise():
    """The reference's own case at its full size (400^2 four-quadrant
    Riemann problem, MUSCL-minmod + Roe, SSP-RK3, CFL 0.5), exact family:
    ten CFL steps equal the oracle's bit for bit."""
    p, dx = cases.riemann2d(400, seed=0)
    print("case riemann400 full size, 10 CFL steps", flush=True)
    with fvmgpu.Context() as ctx:
        qg, ng, _, dt_last = run_gpu(ctx, p, (dx, dx, 1.0), 10, cfl=0.5)
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1)
    qo, r = run_oracle(p, (dx, dx, 1.0), 10, cfl=0.5)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)
