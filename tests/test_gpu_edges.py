"""Edge cases of the GPU path: the smallest blocks, extents one past and one
short of a tile, thin blocks, and the errors the C ABI reports for invalid
requests (forced from the host; no kernel sees bad input)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from _parity import run_gpu, run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

SCHEMES = [dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3),
           dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK4, ref_compat=0),
           dict(recon=O.SIMPLE, limiter=O.LIM_NONE, flux=O.VANLEER, rk=O.RK_SSP3)]


@pytest.mark.parametrize("shape", [(4, 4, 4), (5, 7, 9), (33, 9, 4), (65, 4, 5), (31, 17, 6)])
@pytest.mark.parametrize("si", range(len(SCHEMES)))
@pytest.mark.parametrize("periodic", [False, True])
def test_3d_edge_shapes_bitwise(shape, si, periodic):
    ni, nj, nk = shape
    p = cases.smooth_3d((nk, nj, ni), seed=ni + 3 * nj + 7 * nk)
    sp = (0.11, 0.13, 0.17)
    bc = (1,) * 6 if periodic else None
    print(f"case 3-D edge {shape} scheme {si} periodic {periodic}", flush=True)
    with fvmgpu.Context() as ctx:
        q, _, _, _ = run_gpu(ctx, p, sp, 3, dt=2e-3, bc=bc, **SCHEMES[si])
    qo, _ = run_oracle(p, sp, 3, dt=2e-3, bc=bc, **SCHEMES[si])
    assert np.array_equal(q, qo)


@pytest.mark.parametrize("shape", [(4, 4), (61, 5), (62, 9), (63, 4), (125, 7), (5, 130)])
@pytest.mark.parametrize("si", range(len(SCHEMES)))
def test_2d_edge_shapes_bitwise(shape, si):
    ni, nj = shape
    p = cases.smooth_3d((1, nj, ni), seed=ni + nj)
    p[..., 3] = 0.0
    sp = (0.02, 0.03, 1.0)
    print(f"case 2-D edge {shape} scheme {si}", flush=True)
    with fvmgpu.Context() as ctx:
        q, _, _, _ = run_gpu(ctx, p, sp, 3, dt=1e-3, **SCHEMES[si])
    qo, _ = run_oracle(p, sp, 3, dt=1e-3, **SCHEMES[si])
    assert np.array_equal(q, qo)


def test_invalid_requests_are_errors():
    print("case invalid requests", flush=True)
    with fvmgpu.Context() as ctx:
        with pytest.raises(fvmgpu.FvmError):
            ctx.step(1, dt=1e-3)                          # no scheme, no block
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        with pytest.raises(fvmgpu.FvmError):
            ctx.block((3, 8, 8), (0.1, 0.1, 0.1))         # below the minimum extent
        with pytest.raises(fvmgpu.FvmError):
            ctx.block((8, 8, 8), (0.1, -0.1, 0.1))        # negative spacing
        with pytest.raises(fvmgpu.FvmError, match="2\\^29"):
            ctx.block((1024, 1024, 512), (0.1, 0.1, 0.1))  # 32-bit component offsets
        with pytest.raises(fvmgpu.FvmError):
            ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=999, rk=O.RK_SSP3)
        a = ctx.block((8, 8, 8), (0.1, 0.1, 0.1))
        b = ctx.block((8, 9, 8), (0.1, 0.1, 0.1))
        with pytest.raises(fvmgpu.FvmError):
            a.connect(1, b)                               # faces do not conform
        with pytest.raises(fvmgpu.FvmError):
            ctx.step(1, dt=1e-3)                          # state not uploaded
        with pytest.raises(fvmgpu.FvmError):
            a.upload(np.zeros((5, 8, 8, 9)))              # wrong host shape
        with pytest.raises(fvmgpu.FvmError):
            ctx.step(1, dt=0.0, cfl=0.0)                  # neither dt nor CFL
        c = ctx.block((8, 8, 8), (0.1, 0.1, 0.1))
        a.connect(1, c)
        ctx.viscous(0.01)                                  # viscous on connected blocks: fine
        c.conn_add((8, 1, 1), (8, 8, 8), a, (1, 1, 1), (1, 8, 8), (1, 2, 3))       # and on patches
        with pytest.raises(fvmgpu.FvmError):
            b.set_bc_state(0, fvmgpu.BC_FARFIELD, (1.0, 0.5, 0.0, 0.0, 1.0))   # farfield + viscous: not
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        b2 = ctx.block((8, 8, 1), (0.1, 0.1, 1.0))
        q = np.zeros((5, 1, 8, 8))
        q[0] = 1.0
        q[4] = 2.5
        q[3] = 0.1                                         # rho w in a 2-D block
        with pytest.raises(fvmgpu.FvmError):
            b2.upload(q)
    with fvmgpu.Server(max_active=2) as srv:
        with pytest.raises(fvmgpu.FvmError):
            srv.submit(np.ones((5, 1, 8, 3)), (0.1, 0.1, 1.0), t_end=0.1, cfl=0.5)   # ni < 4
        with pytest.raises(fvmgpu.FvmError):
            srv.submit(np.ones((5, 1, 8, 8)), (0.1, 0.1, 1.0), t_end=0.0, cfl=0.5)   # t_end
        with pytest.raises(fvmgpu.FvmError):
            srv.result(12345)                                                        # unknown id
        r = fvmgpu.Request()
        r.n[0], r.n[1], r.n[2] = 8, 8, 8
        r.spacing[0] = r.spacing[1] = r.spacing[2] = 0.1
        r.recon, r.limiter, r.flux, r.rk = O.MUSCL, O.LIM_MINMOD, O.ROE, O.RK_SSP3
        r.gamma, r.t_end, r.cfl = 1.4, 0.1, 0.5
        r.bc[2] = fvmgpu.BC_WALL                                                     # not served
        r.ic_kind, r.ic_nparams = fvmgpu.IC_TGV, 1
        r.ic_params[0] = 0.1
        rid = ctypes.c_int64()
        assert fvmgpu.load_library().fvm_server_submit(srv._h, ctypes.byref(r), ctypes.byref(rid)) < 0


def test_nan_state_is_reported():
    """A diverging state raises the NaN flag of the norms (the reference's
    abort test, class_iteration.f90:183-191) and the CFL step reports it."""
    p = cases.smooth_3d((6, 6, 6), seed=5)
    print("case NaN state", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        b = ctx.block((6, 6, 6), (0.1, 0.1, 0.1))
        q = cases.to_conserved(p)
        q[4, 3, 3, 3] = -50.0                            # negative pressure: sqrt -> NaN
        b.upload(q)
        ctx.step(2, dt=1e-3)
        _, nan = ctx.norms()
        assert nan == 1
        with pytest.raises(fvmgpu.FvmError):
            ctx.step(1, cfl=0.5)
