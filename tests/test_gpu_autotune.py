"""The launch-shape autotuner (fvm_autotune): its probe launches leave the
uploaded state as it was, and the steps after tuning, on whatever shape and
k-chunk it picked, equal the oracle bit for bit -- one 3-D block with CFL
steps, 2 x 2 x 2 connected blocks in one batched launch, a 2-D plane, and
viscous terms (whose launches the tuner times with the stage)."""
import numpy as np
import pytest

import _oracle as O
from _parity import run_oracle
import fvmgpu
from fvmgpu import cases
from test_gpu_multiblock import split_bounds
from test_gpu_viscous import TGV_SCHEME, oracle_run

pytestmark = pytest.mark.gpu


def test_autotune_vortex_keeps_state_then_bitwise():
    """Exact MUSCL-vanAlbada + Roe, RK4, periodic, CFL 0.8 (the benchmark's
    scheme on a 40 x 36 x 30 box): state unchanged by tuning, then five CFL
    steps equal the oracle, dt included."""
    p, h, _ = cases.isentropic_vortex(40, seed=2)
    p = np.ascontiguousarray(p[:30, :36, :40])
    sch = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    q0 = cases.to_conserved(p)
    print("case autotune vortex 40x36x30", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**sch)
        blk = ctx.block((40, 36, 30), (h, h, h))
        blk.set_periodic()
        blk.upload(q0)
        ctx.autotune(reps=1)
        shape = blk.launch_shape()
        assert np.array_equal(blk.download(), q0)
        dt_last = ctx.step(5, cfl=0.8)
        qg = blk.download()
    print("picked", shape, flush=True)
    qo, r = run_oracle(p, (h, h, h), 5, cfl=0.8, bc=(1,) * 6, **sch)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)


def test_autotune_connected_blocks_weno_bitwise():
    """WENO5 + AUSM+ (textbook weights), SSP-RK3, a 45 x 39 x 33 grid cut into
    2 x 2 x 2 connected blocks of uneven sizes in one batched launch: after
    tuning for all blocks together, three steps equal the oracle on the
    union grid."""
    p = cases.smooth_3d((33, 39, 45), seed=11)
    sp = (1 / 44, 1 / 38, 1 / 32)
    sch = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3, ref_compat=0)
    rx, ry, rz = split_bounds(45, 2), split_bounds(39, 2), split_bounds(33, 2)
    print("case autotune blocks 2x2x2 of 45x39x33", flush=True)
    out = np.zeros((5,) + p.shape[:3])
    with fvmgpu.Context() as ctx:
        ctx.scheme(**sch)
        blocks = {}
        for a, (x0, x1) in enumerate(rx):
            for b, (y0, y1) in enumerate(ry):
                for c, (z0, z1) in enumerate(rz):
                    blk = ctx.block((x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1), sp)
                    blk.upload(cases.to_conserved(p[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1]))
                    blocks[(a, b, c)] = blk
        for (a, b, c), blk in blocks.items():
            idx = [a, b, c]
            for d in range(3):
                if idx[d] == 0:
                    nb = list(idx)
                    nb[d] = 1
                    blk.connect(2 * d + 1, blocks[tuple(nb)], shared=True)
        ctx.autotune(reps=1)
        shapes = sorted(set(blk.launch_shape() for blk in blocks.values()))
        ctx.step(3, dt=1e-3)
        for (a, b, c), blk in blocks.items():
            (x0, x1), (y0, y1), (z0, z1) = rx[a], ry[b], rz[c]
            out[:, z0:z1 + 1, y0:y1 + 1, x0:x1 + 1] = blk.download()
    print("picked", shapes, flush=True)
    assert len(shapes) == 1   # one shape for the whole batched launch
    qo, _ = run_oracle(p, sp, 3, dt=1e-3, **sch)
    assert np.array_equal(out, qo)


def test_autotune_riemann2d_bitwise():
    """The 2-D four-quadrant Riemann case (101 x 101, MUSCL-minmod + Roe,
    SSP-RK3, CFL 0.5): the tuner picks among the row kernel and the plane
    tiles, then eight CFL steps equal the oracle."""
    p, dx = cases.riemann2d(101, seed=0)
    print("case autotune riemann2d 101", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        blk = ctx.block((101, 101, 1), (dx, dx, 1.0))
        q0 = cases.to_conserved(p)
        blk.upload(q0)
        ctx.autotune(reps=1)
        assert np.array_equal(blk.download(), q0)
        dt_last = ctx.step(8, cfl=0.5)
        qg = blk.download()
    qo, r = run_oracle(p, (dx, dx, 1.0), 8, cfl=0.5)
    assert dt_last == r["dt"][-1]
    assert np.array_equal(qg, qo)


def test_autotune_viscous_tgv_bitwise():
    """Navier-Stokes Taylor-Green box (28^3, WENO5 + Van Leer, RK4): the
    tuner's probes include the viscous launches; two CFL steps after tuning
    equal the oracle."""
    n = 28
    p, h = cases.tgv(n)
    visc = cases.tgv_params(re=1600.0, ma=0.1, pr=0.71)
    q0 = cases.to_conserved(p)
    print(f"case autotune tgv {n}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**TGV_SCHEME)
        ctx.viscous(*visc)
        blk = ctx.block((n, n, n), (h, h, h))
        blk.set_periodic()
        blk.upload(q0)
        ctx.autotune(reps=1)
        assert np.array_equal(blk.download(), q0)
        ctx.step(2, cfl=0.5)
        qg = blk.download()
    qo, _ = oracle_run(p, (h, h, h), 2, visc, cfl=0.5, **TGV_SCHEME)
    assert np.array_equal(qg, qo)
