"""Wall / symmetry and inflow / farfield faces (the reference's BC codes,
src/defs/type_defs_bc.f90:11-30, which its flow path never applies).

Invariants, each checked against the oracle on a larger domain:
* symmetry (and inviscid wall) about the boundary vertex plane: a block with
  the face mirrored equals the mirror-symmetric grid twice its size, restricted
  to the block, bit for bit, for a scheme whose oracle is exactly mirror
  symmetric (MUSCL-minmod + AUSM+: the oracle run on the reversed field with
  the normal momentum negated equals the original exactly);
* supersonic inflow / farfield: a block whose face holds the free stream
  equals a longer grid, restricted, whose upstream region stays exactly at
  the free stream (AUSM+ is fully upwind for Mach > 1: the interface mass
  flux and pressure take the upstream side only, and the minmod slope of the
  last uniform vertex is zero, so nothing propagates upstream).
"""
import numpy as np
import pytest

import _oracle as O
from _parity import run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

SYM = dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.AUSMPLUS, rk=O.RK_SSP3)


def mirror_full(half, d, side):
    """half (nk, nj, ni, 5) primitives -> the grid mirrored about its low
    (side 0) or high (side 1) vertex plane in direction d."""
    ax = 2 - d
    h = half.copy()
    plane = [slice(None)] * 4
    plane[ax] = 0 if side == 0 else -1
    plane[3] = 1 + d
    h[tuple(plane)] = 0.0          # no flow through the symmetry plane
    m = np.flip(h, axis=ax).copy()
    m[..., 1 + d] *= -1.0
    cut = [slice(None)] * 4
    if side == 0:
        cut[ax] = slice(0, -1)
        return h, np.concatenate([m[tuple(cut)], h], axis=ax)
    cut[ax] = slice(1, None)
    return h, np.concatenate([h, m[tuple(cut)]], axis=ax)


def run_block(p, sp, nsteps, dt, faces, graph=False, **scheme):
    nk, nj, ni, _ = p.shape
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(0)
        if graph:
            ctx.graph(True)
        blk = ctx.block((ni, nj, nk), sp)
        for f, kind in faces.items():
            if isinstance(kind, tuple):
                blk.set_bc_state(f, kind[0], kind[1])
            else:
                blk.set_bc(f, kind)
        blk.upload(cases.to_conserved(p))
        if graph:   # replays of the captured step, with host work between them
            for _ in range(nsteps):
                ctx.step(1, dt=dt)
        else:
            ctx.step(nsteps, dt=dt)
        return blk.download()


@pytest.mark.parametrize("face", [0, 1, 2, 3, 4, 5])
def test_symmetry_equals_mirrored_grid(face):
    d, side = face // 2, face % 2
    half = cases.smooth_3d((12, 14, 16), seed=30 + face)
    sp = (1 / 15, 1 / 13, 1 / 11)
    h, full = mirror_full(half, d, side)
    kind = fvmgpu.BC_SYMMETRY if face % 2 == 0 else fvmgpu.BC_WALL
    print(f"case symmetry face={face} kind={kind}", flush=True)
    q = run_block(h, sp, 4, 1e-3, {face: kind}, **SYM)
    qo, _ = run_oracle(full, sp, 4, dt=1e-3, **SYM)
    n = h.shape[2 - d]
    sl = [slice(None)] * 4
    sl[3 - d] = slice(qo.shape[3 - d] - n, None) if side == 0 else slice(0, n)
    assert np.array_equal(q, qo[tuple(sl)])


@pytest.mark.parametrize("face", [0, 3])
def test_symmetry_2d(face):
    d, side = face // 2, face % 2
    p, h = cases.riemann2d(40, seed=6)
    hh, full = mirror_full(p, d, side)
    sp = (h, h, 1.0)
    print(f"case symmetry 2-D face={face}", flush=True)
    q = run_block(hh, sp, 6, 2e-4, {face: fvmgpu.BC_SYMMETRY}, **SYM)
    qo, _ = run_oracle(full, sp, 6, dt=2e-4, **SYM)
    sl = [slice(None)] * 4
    sl[3 - d] = slice(qo.shape[3 - d] - 40, None) if side == 0 else slice(0, 40)
    assert np.array_equal(q, qo[tuple(sl)])


@pytest.mark.parametrize("dim,kind", [(3, fvmgpu.BC_FARFIELD), (2, fvmgpu.BC_INFLOW_SUPERSONIC)])
def test_supersonic_inflow_equals_longer_grid(dim, kind):
    pad, ni = 6, 30
    nj, nk = (16, 10) if dim == 3 else (16, 1)
    free = np.array([1.0, 3.0, 0.2, 0.0, 1.0])      # Mach ~2.5 along +x
    big = np.tile(free, (nk, nj, ni + pad, 1))
    # a compact density bump over the short grid's inflow face (zero upstream
    # of it, so extrapolated ghosts would differ from the free stream)
    z, y, x = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni + pad), indexing="ij")
    r2 = (x - (pad + 2.0)) ** 2 + (y - nj / 2) ** 2 + ((z - nk / 2) ** 2 if dim == 3 else 0)
    big[..., 0] += 0.2 * np.maximum(0.0, 1.0 - r2 / 9.0) ** 2
    sp = (0.05, 0.05, 0.05 if dim == 3 else 1.0)
    nsteps, dt = 8, 2e-3
    print(f"case farfield dim={dim} kind={kind}", flush=True)
    q = run_block(big[:, :, pad:].copy(), sp, nsteps, dt, {0: (kind, free)}, **SYM)
    qo, _ = run_oracle(big, sp, nsteps, dt=dt, **SYM)
    fq = cases.to_conserved(np.tile(free, (nk, nj, 3, 1)))
    assert np.array_equal(qo[:, :, :, pad - 3:pad], fq)   # the premise: upstream untouched
    assert np.array_equal(q, qo[:, :, :, pad:])


def test_uniform_free_stream_farfield():
    """Farfield ghosts on every face of a uniform free stream are the
    extrapolated ghosts of the oracle (both equal the free stream)."""
    free = np.array([0.9, 0.4, -0.3, 0.25, 1.3])
    p = np.tile(free, (10, 12, 14, 1))
    faces = {f: (fvmgpu.BC_FARFIELD, free) for f in range(6)}
    for scheme in (dict(recon=O.MUSCL, limiter=O.LIM_VANLEER, flux=O.ROE, rk=O.RK4),
                   dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3)):
        q = run_block(p, (0.1, 0.1, 0.1), 5, 1e-3, faces, **scheme)
        qo, _ = run_oracle(p, (0.1, 0.1, 0.1), 5, dt=1e-3, **scheme)
        assert np.array_equal(q, qo)


def test_bc_kinds_api():
    with fvmgpu.Context() as ctx:
        ctx.scheme(**dict(SYM, rk=O.RK4))       # RK4 leaves a state at rest exactly
        a = ctx.block((10, 8, 6), (0.1, 0.1, 0.1))
        b = ctx.block((10, 8, 6), (0.1, 0.1, 0.1))
        with pytest.raises(fvmgpu.FvmError, match="state"):
            a.set_bc(0, fvmgpu.BC_FARFIELD)
        with pytest.raises(fvmgpu.FvmError):
            a.set_bc_state(0, fvmgpu.BC_FARFIELD, (-1.0, 0, 0, 0, 1.0))
        with pytest.raises(fvmgpu.FvmError):
            a.set_bc_state(0, fvmgpu.BC_SYMMETRY, (1.0, 0, 0, 0, 1.0))
        with pytest.raises(fvmgpu.FvmError, match="isothermal"):
            a.set_bc(0, fvmgpu.BC_WALL_VISCOUS_ISOTHERMAL)
        a.set_bc(0, fvmgpu.BC_WALL_VISCOUS)     # the adiabatic no-slip wall
        a.set_bc(0, fvmgpu.BC_EXTRAP)
        with pytest.raises(fvmgpu.FvmError):
            a.set_bc(0, 999)
        a.set_bc(2, fvmgpu.BC_SYMMETRY)
        ctx.viscous(1e-3)                        # walls and symmetry planes take viscous terms
        with pytest.raises(fvmgpu.FvmError, match="viscous"):
            a.set_bc_state(3, fvmgpu.BC_FARFIELD, (1.0, 0, 0, 0, 1.0))
        ctx.viscous(0.0)
        a.set_bc_state(3, fvmgpu.BC_FARFIELD, (1.0, 0, 0, 0, 1.0))
        with pytest.raises(fvmgpu.FvmError, match="viscous"):
            ctx.viscous(1e-3)                    # a fixed-state face: not with viscous terms
        a.set_bc(3, fvmgpu.BC_OUTFLOW)
        a.set_bc(2, fvmgpu.BC_OUTFLOW)          # replaced: extrapolation again
        a.set_bc(1, fvmgpu.BC_WALL)
        a.connect(1, b)                          # a wall face may become connected
        with pytest.raises(fvmgpu.FvmError, match="connected"):
            a.set_bc(1, fvmgpu.BC_SYMMETRY)
        a.set_bc(4, fvmgpu.BC_PERIODIC)
        a.set_bc(5, fvmgpu.BC_SYMMETRY)          # the k- face falls back to extrapolation
        u = np.zeros((5, 6, 8, 10))
        u[0] = 1.0
        u[4] = 2.5
        a.upload(u)
        b.upload(u)
        ctx.step(2, dt=1e-3)
        assert np.array_equal(a.download(), u)   # a state at rest stays at rest


def test_bc_faces_under_graph_replay():
    """Wall, symmetry and farfield faces refilled inside a replayed hipGraph
    step: every replay equals the eager run (bit for bit)."""
    half = cases.smooth_3d((12, 14, 16), seed=40)
    h, _ = mirror_full(half, 0, 0)
    free = (1.0, 0.2, 0.1, 0.0, 1.0)
    faces = {0: fvmgpu.BC_SYMMETRY, 3: fvmgpu.BC_WALL, 5: (fvmgpu.BC_FARFIELD, free)}
    sp = (1 / 15, 1 / 13, 1 / 11)
    print("case BC faces under graph replay", flush=True)
    eager = run_block(h, sp, 5, 1e-3, faces, **SYM)
    graph = run_block(h, sp, 5, 1e-3, faces, graph=True, **SYM)
    assert np.array_equal(eager, graph)
