"""The drop-in boundary in the reference's own terms, through the C ABI:

- fvm_upload_planes / fvm_download_planes on ClassBlockData%q as the
  reference allocates it for its 2-D case (src/base/defs_block_data.f90:
  137-141 with gsize(3) = 2, src/base/method_gridgen.f90:98-99): q(ni+6,
  nj+6, 2+6, 5), two identical real k-planes between three ghost planes on
  each side;
- fvm_stage, one Runge-Kutta stage per call (the stage loop of
  src/solve/class_time_discret.f90:172-195), against fvm_step and the oracle.
"""
import numpy as np
import pytest

import _oracle as O
from _parity import NG, reference_q_2d, run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k_plane", [1, 2])
def test_reference_2d_layout_round_trip(k_plane):
    """Upload from the reference's 8-plane array, ten SSP-RK3 steps of the
    reference's Riemann-2D scheme, download into the same layout: every real
    plane equals the oracle bit for bit, the ghost planes carry the real
    vertices of their neighbouring plane, i/j ghosts copy the boundary
    vertex, edges stay zero."""
    n = 33
    p, dx = cases.riemann2d(n, seed=0)
    H = reference_q_2d(p)
    print(f"case reference-layout riemann2d n={n} k_plane={k_plane}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block((n, n, 1), (dx, dx, 1.0))
        blk.upload_planes(H, NG, 2, k_plane)
        ctx.step(10, dt=2e-4)
        out = blk.download_planes(NG, 2)
        plain = blk.download(NG)                     # the one-plane layout
    qo, _ = run_oracle(p, (dx, dx, 1.0), 10, dt=2e-4)   # (5, 1, nj, ni)
    real = out[:, :, NG:NG + n, NG:NG + n]
    for kk in range(2 + 2 * NG):
        assert np.array_equal(real[:, kk], qo[:, 0]), f"plane {kk}"
    for kk in (NG, NG + 1):
        assert np.array_equal(out[:, kk], plain[:, 0]), f"real plane {kk} with its i/j ghosts"
        for g in range(NG):
            assert np.array_equal(out[:, kk, NG:NG + n, g], qo[:, 0, :, 0])
            assert np.array_equal(out[:, kk, g, NG:NG + n], qo[:, 0, 0, :])
    for kk in list(range(NG)) + list(range(NG + 2, 2 + 2 * NG)):
        ring = out[:, kk].copy()
        ring[:, NG:NG + n, NG:NG + n] = 0.0
        assert not ring.any(), f"ghost plane {kk} has values outside its real region"
    for kk in range(2 + 2 * NG):                     # edges (i and j ghost) zero
        assert not out[:, kk, :NG, :NG].any() and not out[:, kk, -NG:, -NG:].any()
    assert not out[3].any()                          # rho*w


def test_reference_2d_layout_rejects_bad_planes():
    n = 17
    p, dx = cases.riemann2d(n, seed=0)
    H = reference_q_2d(p)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block((n, n, 1), (dx, dx, 1.0))
        for nk, kp in ((2, 0), (2, 3), (0, 1)):         # checked by the library
            rc = fvmgpu._lib.fvm_upload_planes(blk._h, H.ctypes.data, NG, nk, kp)
            assert rc == -1, (nk, kp)
        bad = H.copy()
        bad[3, NG + 1, NG + 5, NG + 5] = 1e-3            # rho*w on the plane read
        with pytest.raises(fvmgpu.FvmError, match="rho\\*w"):
            blk.upload_planes(bad, NG, 2, 2)
        blk.upload_planes(bad, NG, 2, 1)                 # the other plane is clean


def test_upload_planes_3d_equals_upload():
    p = cases.smooth_3d((9, 10, 12), seed=5)
    sp = (0.1, 0.11, 0.125)
    q = cases.to_conserved(p)
    H = np.zeros((5, 9 + 2 * NG, 10 + 2 * NG, 12 + 2 * NG))
    H[:, NG:-NG, NG:-NG, NG:-NG] = q
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        a = ctx.block((12, 10, 9), sp)
        b = ctx.block((12, 10, 9), sp)
        a.upload(H, NG)
        b.upload_planes(H, NG, 9)
        with pytest.raises(fvmgpu.FvmError):
            b.upload_planes(H, NG, 8)
        ctx.step(2, dt=1e-3)
        assert np.array_equal(a.download(NG), b.download_planes(NG, 9))


def _stages(ctx, nstage, nsteps, dt):
    for _ in range(nsteps):
        for s in range(1, nstage + 1):
            ctx.stage(s, dt)


@pytest.mark.parametrize("rk,nstage", [(O.RK_SSP3, 3), (O.RK4, 4)])
@pytest.mark.parametrize("dim", [2, 3])
def test_stage_calls_equal_step(rk, nstage, dim):
    """Stage-by-stage calls give fvm_step's state and norms bit for bit, and
    the oracle's."""
    if dim == 2:
        p, dx = cases.riemann2d(29, seed=2)
        sp, dt = (dx, dx, 1.0), 2e-4
    else:
        p = cases.smooth_3d((10, 11, 13), seed=4)
        sp, dt = (1 / 12, 1 / 10, 1 / 9), 1e-3
    nk, nj, ni, _ = p.shape
    print(f"case stage-calls dim={dim} rk={rk}", flush=True)
    res = []
    for per_stage in (False, True):
        with fvmgpu.Context() as ctx:
            ctx.scheme(rk=rk)
            blk = ctx.block((ni, nj, nk), sp)
            blk.upload(cases.to_conserved(p))
            if per_stage:
                _stages(ctx, nstage, 4, dt)
            else:
                ctx.step(4, dt=dt)
            t, steps = ctx.time()
            res.append((blk.download(), ctx.norms()[0], t, steps))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2:] == res[1][2:] and res[1][3] == 4
    qo, _ = run_oracle(p, sp, 4, dt=dt, rk=rk)
    assert np.array_equal(res[1][0], qo)


def test_stage_order_and_dt_enforced():
    p, dx = cases.riemann2d(17, seed=0)
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        blk = ctx.block((17, 17, 1), (dx, dx, 1.0))
        q0 = cases.to_conserved(p)
        blk.upload(q0)
        with pytest.raises(fvmgpu.FvmError, match="stage 1 expected"):
            ctx.stage(2, 2e-4)
        with pytest.raises(fvmgpu.FvmError, match="dt > 0"):
            ctx.stage(1, 0.0)
        ctx.stage(1, 2e-4)
        with pytest.raises(fvmgpu.FvmError, match="differs"):
            ctx.stage(2, 1e-4)
        with pytest.raises(fvmgpu.FvmError, match="fvm_stage"):
            ctx.step(1, dt=2e-4)
        # between stages the state is still u^n
        assert np.array_equal(blk.download(), q0)
        ctx.stage(2, 2e-4)
        ctx.stage(3, 2e-4)
        with pytest.raises(fvmgpu.FvmError, match="out of range"):
            ctx.stage(4, 2e-4)
        ctx.step(1, dt=2e-4)                             # a whole step is allowed again
        # an upload abandons a step in progress
        ctx.stage(1, 2e-4)
        blk.upload(q0)
        ctx.step(2, dt=2e-4)
        a = blk.download()
    qo, _ = run_oracle(p, (dx, dx, 1.0), 2, dt=2e-4)
    assert np.array_equal(a, qo)
