"""hipGraph replay of time steps: every replay gives exactly the eager
launches' result (checked after each of several calls, with eager work and
host synchronisation between the replays), in CFL and fixed-dt mode, across
a re-capture after a configuration change."""
import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


def make(ctx, p, h, periodic=True, **scheme):
    ctx.scheme(**scheme)
    nk, nj, ni, _ = p.shape
    b = ctx.block((ni, nj, nk), (h, h, h))
    if periodic:
        b.set_periodic()
    b.upload(cases.to_conserved(p))
    return b


@pytest.mark.parametrize("scheme", [
    dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4),
    dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3, ref_compat=0)])
def test_graph_equals_eager(scheme):
    p, h, _ = cases.isentropic_vortex(24, seed=3)
    print(f"case graph vs eager {scheme}", flush=True)
    with fvmgpu.Context() as eg, fvmgpu.Context() as gr:
        be = make(eg, p, h, **scheme)
        bg = make(gr, p, h, **scheme)
        gr.graph(True)
        for call in range(4):
            n = 3 + call
            de = eg.step(n, cfl=0.6)
            dg = gr.step(n, cfl=0.6)
            assert de == dg
            qe, qg = be.download(), bg.download()
            assert np.array_equal(qe, qg), call
            ne, _ = eg.norms()
            ng, _ = gr.norms()
            assert np.array_equal(ne, ng)
        # fixed dt, then a numerics change forces a new capture
        eg.step(3, dt=1e-3)
        gr.step(3, dt=1e-3)
        assert np.array_equal(be.download(), bg.download())
        eg.numerics(fvmgpu.NUMERICS_FAST)
        gr.numerics(fvmgpu.NUMERICS_FAST)
        eg.step(4, cfl=0.6)
        gr.step(4, cfl=0.6)
        assert np.array_equal(be.download(), bg.download())


def test_graph_end_time_and_viscous():
    """End-time halting inside replayed steps, with the viscous kernels in
    the captured step."""
    p, h = cases.tgv(16)
    scheme = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.VANLEER, rk=O.RK4, ref_compat=0)
    print("case graph tgv16 end time", flush=True)
    out = []
    for use in (False, True):
        with fvmgpu.Context() as ctx:
            ctx.viscous(*cases.tgv_params(re=100.0))
            b = make(ctx, p, h, **scheme)
            ctx.graph(use)
            ctx.set_time(0.0, 0.05)
            for _ in range(20):
                ctx.step(4, cfl=0.5)
                t, n = ctx.time()
                if t >= 0.05 * (1 - 1e-14):
                    break
            out.append((b.download(), ctx.time()))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
