"""Asynchronous snapshots (fvm_snapshot_*) and restart files
(fvm_snapshot_save / fvm_restart_load).

Invariants:
* a snapshot holds u^n at the point of the stream where it was taken, while
  later steps run on: equal bit for bit to a download at that point;
* the output variables are the reference's q -> p conversion
  (src/solve/class_flow_field.f90:383-396, the same operation order as
  cases.to_primitive) in its output layout vdata(ni, nj, nk, nvars)
  (src/solve/class_output.f90:377-410);
* save + restart + n steps = the uninterrupted run, bit for bit, with the
  solution time and step count carried over.
"""
import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

SCHEME = dict(recon=O.MUSCL, limiter=O.LIM_VANLEER, flux=O.ROE, rk=O.RK_SSP3)


def two_blocks(ctx, p, sp):
    """The field p (nk, nj, ni, 5) as two blocks joined across i (shared plane)."""
    nk, nj, ni, _ = p.shape
    xa = ni // 2
    q = cases.to_conserved(p)
    a = ctx.block((xa + 1, nj, nk), sp)
    b = ctx.block((ni - xa, nj, nk), sp)
    a.upload(q[:, :, :, :xa + 1])
    b.upload(q[:, :, :, xa:])
    a.connect(1, b, shared=True)
    return [a, b]


def field(dim):
    if dim == 3:
        return cases.smooth_3d((14, 18, 22), seed=3), (1 / 21, 1 / 17, 1 / 13)
    p, h = cases.riemann2d(40, seed=2)
    return p, (h, h, 1.0)


def to_prim_layout(qblk):
    """Conserved download (5, nk, nj, ni) -> primitive output arrays (5, nk, nj, ni)."""
    return np.moveaxis(cases.to_primitive(qblk), -1, 0)


@pytest.mark.parametrize("dim", [3, 2])
def test_snapshot_is_state_at_begin(dim):
    p, sp = field(dim)
    print(f"case snapshot dim={dim}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**SCHEME)
        ctx.numerics(0)
        blocks = two_blocks(ctx, p, sp)
        ctx.step(3, dt=2e-4)
        sp_id = ctx.snapshot(fvmgpu.PRIMITIVE_VARS)
        sc_id = ctx.snapshot(fvmgpu.CONSERVED_VARS + (fvmgpu.VAR_PRESSURE,))
        ctx.step(4, dt=2e-4)     # runs while the copies land
        after = [b.download() for b in blocks]
        for b in blocks:
            prim, t, n = ctx.snapshot_read(sp_id, b)
            cons, t2, n2 = ctx.snapshot_read(sc_id, b)
            assert n == n2 == 3 and t == t2
            assert ctx.snapshot_ready(sp_id)
            # u^n after 3 steps from a second context
            with fvmgpu.Context() as ref:
                ref.scheme(**SCHEME)
                ref.numerics(0)
                rb = two_blocks(ref, p, sp)
                ref.step(3, dt=2e-4)
                q3 = rb[blocks.index(b)].download()
            assert np.array_equal(cons[:5], q3)
            assert np.array_equal(prim, to_prim_layout(q3))
            assert np.array_equal(cons[5], prim[4])
        ctx.snapshot_release(sp_id)
        ctx.snapshot_release(sc_id)
        # the steps after the snapshot were not disturbed
        with fvmgpu.Context() as ref:
            ref.scheme(**SCHEME)
            ref.numerics(0)
            rb = two_blocks(ref, p, sp)
            ref.step(7, dt=2e-4)
            for x, y in zip(after, rb):
                assert np.array_equal(x, y.download())


def test_snapshot_slots():
    p, sp = field(3)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**SCHEME)
        blocks = two_blocks(ctx, p, sp)
        ids = []
        for s in range(4):
            ctx.step(1, dt=2e-4)
            ids.append(ctx.snapshot([fvmgpu.VAR_DENSITY]))
        with pytest.raises(fvmgpu.FvmError, match="in use"):
            ctx.snapshot([fvmgpu.VAR_DENSITY])
        with pytest.raises(fvmgpu.FvmError):
            ctx.snapshot([999])
        steps = [ctx.snapshot_read(i, blocks[0])[2] for i in ids]
        assert steps == [1, 2, 3, 4]
        ctx.snapshot_release(ids[1])
        with pytest.raises(fvmgpu.FvmError):
            ctx.snapshot_release(ids[1])
        # the freed slot takes a larger snapshot (it grows)
        big = ctx.snapshot(fvmgpu.CONSERVED_VARS + fvmgpu.PRIMITIVE_VARS)
        arr, _, n = ctx.snapshot_read(big, blocks[1])
        assert n == 4 and arr.shape[0] == 10
        assert np.array_equal(arr[:5], blocks[1].download())
        for i in ids[:1] + ids[2:] + [big]:
            ctx.snapshot_release(i)


@pytest.mark.parametrize("cfl_mode", [False, True])
def test_restart_equals_uninterrupted(tmp_path, cfl_mode):
    p, sp = field(3)
    path = tmp_path / "state.fvmsnap"
    kw = dict(cfl=0.4) if cfl_mode else dict(dt=2e-4)
    print(f"case restart cfl={cfl_mode}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**SCHEME)
        ctx.numerics(0)
        blocks = two_blocks(ctx, p, sp)
        ctx.step(5, **kw)
        sid = ctx.snapshot(fvmgpu.CONSERVED_VARS)
        ctx.snapshot_save(sid, path)
        ctx.step(6, **kw)
        ctx.snapshot_release(sid)
        t_ref, n_ref = ctx.time()
        ref = [b.download() for b in blocks]
    f = fvmgpu.read_snapshot_file(path)
    assert f["step"] == 5 and f["codes"] == list(fvmgpu.CONSERVED_VARS)
    assert [b[0] for b in f["blocks"]] == [(12, 18, 14), (11, 18, 14)]
    with fvmgpu.Context() as ctx:
        ctx.scheme(**SCHEME)
        ctx.numerics(0)
        zero = np.zeros_like(p)
        zero[..., 0] = 1.0
        zero[..., 4] = 1.0
        blocks = two_blocks(ctx, zero, sp)
        ctx.restart(path)
        t0, n0 = ctx.time()
        assert n0 == 5 and t0 == f["time"]
        ctx.step(6, **kw)
        t1, n1 = ctx.time()
        assert n1 == n_ref and t1 == t_ref
        for x, b in zip(ref, blocks):
            assert np.array_equal(x, b.download())


def test_restart_rejects_mismatch(tmp_path):
    p, sp = field(3)
    prim_path = tmp_path / "prim.fvmsnap"
    cons_path = tmp_path / "cons.fvmsnap"
    with fvmgpu.Context() as ctx:
        ctx.scheme(**SCHEME)
        two_blocks(ctx, p, sp)
        a = ctx.snapshot(fvmgpu.PRIMITIVE_VARS)
        b = ctx.snapshot(fvmgpu.CONSERVED_VARS)
        ctx.snapshot_save(a, prim_path)
        ctx.snapshot_save(b, cons_path)
        ctx.snapshot_release(a)
        ctx.snapshot_release(b)
        with pytest.raises(fvmgpu.FvmError, match="conserved"):
            ctx.restart(prim_path)
        with pytest.raises(fvmgpu.FvmError):
            ctx.restart(tmp_path / "missing.fvmsnap")
    with fvmgpu.Context() as ctx:
        ctx.scheme(**SCHEME)
        ctx.block((10, 18, 14), sp)
        with pytest.raises(fvmgpu.FvmError, match="blocks"):
            ctx.restart(cons_path)
    with pytest.raises(fvmgpu.FvmError):
        with fvmgpu.Context() as ctx:
            ctx.scheme(**SCHEME)
            two_blocks(ctx, p, sp)
            sid = ctx.snapshot(fvmgpu.CONSERVED_VARS)
            ctx.snapshot_save(sid, tmp_path / "no_such_dir" / "x.fvmsnap")
            ctx.snapshot_release(sid)
