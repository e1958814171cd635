"""Multi-block grids: device-side block-to-block ghost exchange and the
batched / persistent / per-block launches.  Invariant: a grid split into
blocks that share their interface vertex planes gives, bit for bit, the same
values as the unsplit grid (the oracle on the union domain)."""
import itertools

import numpy as np
import pytest

import _oracle as O
from _parity import run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


def split_bounds(n, parts):
    """Vertex ranges [lo, hi] of `parts` blocks sharing interface planes."""
    cuts = np.linspace(0, n - 1, parts + 1).round().astype(int)
    return [(int(cuts[i]), int(cuts[i + 1])) for i in range(parts)]


def run_blocks(p, spacing, ranges, nsteps, dt, mode, periodic=(False, False, False),
               numerics=0, **scheme):
    """Run the union state p (nk, nj, ni, 5) as blocks over the given
    per-direction vertex ranges; returns the union array rebuilt from them."""
    rx, ry, rz = ranges
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(numerics)
        ctx.launch_mode(mode)
        blocks = {}
        for (a, (x0, x1)), (b, (y0, y1)), (c, (z0, z1)) in itertools.product(
                enumerate(rx), enumerate(ry), enumerate(rz)):
            blk = ctx.block((x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1), spacing)
            blk.upload(cases.to_conserved(p[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1]))
            blocks[(a, b, c)] = blk
        dims = (len(rx), len(ry), len(rz))
        for (a, b, c), blk in blocks.items():
            idx = [a, b, c]
            for d in range(3):
                if idx[d] + 1 < dims[d]:
                    nb = list(idx); nb[d] += 1
                    blk.connect(2 * d + 1, blocks[tuple(nb)], shared=True)
                elif periodic[d]:
                    nb = list(idx); nb[d] = 0
                    if dims[d] == 1:
                        blk.set_bc(2 * d, fvmgpu.BC_PERIODIC)
                    else:
                        blk.connect(2 * d + 1, blocks[tuple(nb)], shared=False)
        ctx.step(nsteps, dt=dt)
        norms, _ = ctx.norms()
        out = np.zeros((5,) + p.shape[:3])
        for (a, b, c), blk in blocks.items():
            (x0, x1), (y0, y1), (z0, z1) = rx[a], ry[b], rz[c]
            out[:, z0:z1 + 1, y0:y1 + 1, x0:x1 + 1] = blk.download()
        return out, norms


@pytest.mark.parametrize("mode,tile", [(fvmgpu.LAUNCH_BATCHED, ""), (fvmgpu.LAUNCH_PERSISTENT, ""),
                                       (fvmgpu.LAUNCH_PER_BLOCK, ""), (fvmgpu.LAUNCH_BATCHED, "4"),
                                       (fvmgpu.LAUNCH_PERSISTENT, "5")])
def test_split_equals_single_block(mode, tile, monkeypatch):
    p = cases.smooth_3d((26, 30, 40), seed=11)
    sp = (1 / 39, 1 / 29, 1 / 25)
    ranges = (split_bounds(40, 2), split_bounds(30, 2), split_bounds(26, 2))
    if tile:   # the register-reconstruction kernel's connected-face variant
        monkeypatch.setenv("FVM_TILE", tile)
    print(f"case split 2x2x2 of 40x30x26 mode={mode} tile={tile}", flush=True)
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4)
    q, _ = run_blocks(p, sp, ranges, 3, 1e-3, mode, **scheme)
    qo, _ = run_oracle(p, sp, 3, dt=1e-3, **scheme)
    assert np.array_equal(q, qo)


def test_split_periodic_equals_single_block():
    """Periodic wrap across blocks (abutting faces) = single periodic block."""
    p, h, _ = cases.isentropic_vortex(24, seed=2)
    ranges = (split_bounds(24, 3), split_bounds(24, 2), split_bounds(24, 1))
    print("case periodic split 3x2x1 of vortex24", flush=True)
    scheme = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3, ref_compat=0)
    q, _ = run_blocks(p, (h, h, h), ranges, 4, 5e-3, fvmgpu.LAUNCH_BATCHED,
                      periodic=(True, True, True), **scheme)
    qo, _ = run_oracle(p, (h, h, h), 4, dt=5e-3, bc=(1,) * 6, **scheme)
    assert np.array_equal(q, qo)


def test_uneven_3x3x3_weno_ausmplus():
    """27 blocks of seeded uneven sizes (the batched-grid configuration at a
    test size), textbook WENO5 + AUSM+, SSP-RK3."""
    rng = np.random.default_rng(3)
    sizes = [rng.integers(8, 15, 3) for _ in range(3)]
    ranges = []
    for d in range(3):
        lo, r = 0, []
        for s in sizes[d]:
            r.append((lo, lo + int(s) - 1))
            lo += int(s) - 1
        ranges.append(r)
    n = [r[-1][1] + 1 for r in ranges]
    p = cases.smooth_3d((n[2], n[1], n[0]), seed=12)
    sp = (0.05, 0.06, 0.07)
    print(f"case 3x3x3 uneven {n}", flush=True)
    scheme = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3, ref_compat=0)
    q, _ = run_blocks(p, sp, ranges, 2, 1e-3, fvmgpu.LAUNCH_BATCHED, **scheme)
    qo, _ = run_oracle(p, sp, 2, dt=1e-3, **scheme)
    assert np.array_equal(q, qo)


def test_independent_cases_batched_equal_alone():
    """Unconnected blocks of different sizes in one batched launch advance
    exactly as each does alone (fixed dt)."""
    shapes = [(20, 18, 16), (33, 12, 9), (8, 40, 11)]
    ps = [cases.smooth_3d(s[::-1], seed=20 + i) for i, s in enumerate(shapes)]
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.VANLEER, rk=O.RK4)
    print("case independent batch", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        blks = []
        for s, p in zip(shapes, ps):
            b = ctx.block(s, (0.1, 0.1, 0.1))
            b.upload(cases.to_conserved(p))
            blks.append(b)
        ctx.step(3, dt=2e-3)
        together = [b.download() for b in blks]
    for s, p, q in zip(shapes, ps, together):
        with fvmgpu.Context() as ctx:
            ctx.scheme(**scheme)
            b = ctx.block(s, (0.1, 0.1, 0.1))
            b.upload(cases.to_conserved(p))
            ctx.step(3, dt=2e-3)
            assert np.array_equal(b.download(), q)


@pytest.mark.parametrize("tile2", ["0", "1", "2"])
def test_split_2d_equals_single_block(tile2, monkeypatch):
    monkeypatch.setenv("FVM_TILE2", tile2)
    p, dx = cases.riemann2d(61, seed=0)
    ranges = (split_bounds(61, 3), split_bounds(61, 2), [(0, 0)])
    print(f"case 2-D split 3x2 riemann61 tile {tile2}", flush=True)
    q, norms = run_blocks(p, (dx, dx, 1.0), ranges, 5, 2e-4, fvmgpu.LAUNCH_BATCHED)
    qo, r = run_oracle(p, (dx, dx, 1.0), 5, dt=2e-4)
    assert np.array_equal(q, qo)


@pytest.mark.parametrize("recon,limiter,flux", [(O.MUSCL, O.LIM_MINMOD, O.ROE),
                                                (O.WENO, O.LIM_NONE, O.AUSMPLUS)])
def test_independent_cases_2d_batched_equal_alone(recon, limiter, flux):
    """Unconnected 2-D blocks of different sizes (different tile counts and
    row chunks) in one batched launch advance exactly as each does alone."""
    sizes = [(70, 70), (155, 155), (45, 92), (277, 277)]
    ps = []
    for i, (ni, nj) in enumerate(sizes):
        if ni == nj:
            # the reference's quadrant states: drawn states can drive Roe to a
            # negative pressure in strong rarefactions, alone or batched
            p, _ = cases.riemann2d(ni, seed=0)
        else:
            p = cases.smooth_3d((1, nj, ni), seed=30 + i)
            p[..., 3] = 0.0
        ps.append(p)
    scheme = dict(recon=recon, limiter=limiter, flux=flux, rk=O.RK_SSP3, ref_compat=0)
    print(f"case independent 2-D batch recon={recon}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        blks = []
        for (ni, nj), p in zip(sizes, ps):
            b = ctx.block((ni, nj, 1), (1.0 / (ni - 1), 1.0 / (nj - 1), 1.0))
            b.upload(cases.to_conserved(p))
            blks.append(b)
        ctx.step(20, dt=1e-4)
        together = [b.download() for b in blks]
    for (ni, nj), p, q in zip(sizes, ps, together):
        with fvmgpu.Context() as ctx:
            ctx.scheme(**scheme)
            b = ctx.block((ni, nj, 1), (1.0 / (ni - 1), 1.0 / (nj - 1), 1.0))
            b.upload(cases.to_conserved(p))
            ctx.step(20, dt=1e-4)
            assert np.array_equal(b.download(), q), (ni, nj)


@pytest.mark.parametrize("ncut,periodic", [((3, 2, 2), (True, True, True)),
                                           ((2, 3, 1), (False, False, False))])
def test_cut_uniform_equals_single_block(ncut, periodic):
    """The reference's uniform block cut (shared planes, connected faces)
    reproduces the single block bit for bit."""
    if periodic[0]:
        p, h, _ = cases.isentropic_vortex(30, seed=4)
        sp = (h, h, h)
        bc = (O.BC_PERIODIC,) * 6
    else:
        p = cases.smooth_3d((17, 25, 33), seed=12)
        sp = (1 / 32, 1 / 24, 1 / 16)
        bc = None
    scheme = dict(recon=O.MUSCL, limiter=O.LIM_VANLEER, flux=O.ROE, rk=O.RK_SSP3)
    print(f"case cut_uniform {ncut} periodic={periodic}", flush=True)
    with fvmgpu.Context() as ctx:
        ctx.scheme(**scheme)
        ctx.numerics(0)
        blocks = ctx.cut_blocks(cases.to_conserved(p), sp, ncut, periodic=periodic)
        ctx.step(4, dt=2e-3)
        out = np.zeros((5,) + p.shape[:3])
        for blk, (x0, y0, z0) in blocks.values():
            q = blk.download()
            out[:, z0:z0 + q.shape[1], y0:y0 + q.shape[2], x0:x0 + q.shape[3]] = q
    qo, _ = run_oracle(p, sp, 4, dt=2e-3, bc=bc, **scheme)
    assert np.array_equal(out, qo)


def test_connect_cuts_a_periodic_direction_open():
    """A block made periodic in x and then connected on its i+ face: its i-
    face no longer wraps onto itself but extrapolates (the same result as
    setting it so explicitly), so the face table stays consistent."""
    p = cases.smooth_3d((10, 12, 30), seed=8)
    sp = (1 / 29, 1 / 11, 1 / 9)
    q = cases.to_conserved(p)
    out = []
    for periodic_first in (True, False):
        with fvmgpu.Context() as ctx:
            ctx.scheme(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
            a = ctx.block((16, 12, 10), sp)
            b = ctx.block((15, 12, 10), sp)
            a.upload(q[:, :, :, :16])
            b.upload(q[:, :, :, 15:])
            if periodic_first:
                a.set_bc(0, fvmgpu.BC_PERIODIC)
            a.connect(1, b, shared=True)
            ctx.step(3, dt=1e-3)
            out.append((a.download(), b.download()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    qo, _ = run_oracle(p, sp, 3, dt=1e-3, recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE)
    assert np.array_equal(out[0][0], qo[:, :, :, :16])
