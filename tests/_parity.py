"""Shared helpers: run one configuration on the GPU (through the C ABI) and on
the oracle, from the same initial state."""
import numpy as np

import _oracle as O
import fvmgpu
from fvmgpu import cases


def run_gpu(ctx, p, spacing, nsteps, dt=0.0, cfl=0.0, bc=None, numerics=0, **scheme):
    nk, nj, ni, _ = p.shape
    ctx.scheme(**scheme)
    ctx.numerics(numerics)
    blk = ctx.block((ni, nj, nk), spacing)
    if bc is not None:
        for f, kind in enumerate(bc):
            if kind and (nk > 1 or f < 4):
                blk.set_bc(f, kind)
    blk.upload(cases.to_conserved(p))
    dt_last = ctx.step(nsteps, dt=dt, cfl=cfl)
    q = blk.download()
    norms, flag = ctx.norms()
    return q, norms, flag, dt_last


def run_oracle(p, spacing, nsteps, dt=0.0, cfl=0.0, bc=None, recon=O.MUSCL,
               limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3, kappa=0.33333, entropy_fix=0.2,
               gamma=1.4, ref_compat=1):
    nk = p.shape[0]
    dim = 3 if nk > 1 else 2
    r = O.oracle_solve(p, dim, spacing, recon=recon, limiter=limiter, flux=flux, rk=rk,
                       kappa=kappa, entfix=entropy_fix, gamma=gamma, dt=dt if dt > 0 else 0.0,
                       cfl=cfl, nsteps=nsteps, p_from_q=1,
                       bc=tuple(bc) if bc is not None else (0,) * 6,
                       weno_textbook=0 if ref_compat else 1)
    gz = 3 if dim == 3 else 0
    q = r["q"][:, gz:gz + nk, 3:-3, 3:-3]
    return q, r


NG = 3   # ghost layers of the reference's arrays (src/base/defs_ghost_cell.f90:42)


def reference_q_2d(p):
    """The reference's q(ni+6, nj+6, 8, 5) (as a C array [5][8][nj+6][ni+6])
    after flowfield%initial: conserved state on both real planes, face ghosts
    copied from the boundary vertex in i, j and k (update_ghost_cell,
    src/base/defs_block_data.f90:226-256), edges and corners never written."""
    q = cases.to_conserved(p)[:, 0]                      # (5, nj, ni)
    nj, ni = q.shape[1:]
    H = np.zeros((5, 2 + 2 * NG, nj + 2 * NG, ni + 2 * NG))
    for kk in range(2 + 2 * NG):
        H[:, kk, NG:NG + nj, NG:NG + ni] = q
    for kk in (NG, NG + 1):
        for g in range(NG):
            H[:, kk, NG:NG + nj, g] = q[:, :, 0]
            H[:, kk, NG:NG + nj, NG + ni + g] = q[:, :, -1]
            H[:, kk, g, NG:NG + ni] = q[:, 0, :]
            H[:, kk, NG + nj + g, NG:NG + ni] = q[:, -1, :]
    return H


def rel_linf(a, b):
    """Relative L-inf per conserved component, normalised by max|b| with an
    absolute floor of 1 (rho*w is identically zero in 2-D)."""
    out = []
    for c in range(a.shape[0]):
        scale = max(np.abs(b[c]).max(), 1.0)
        out.append(np.abs(a[c] - b[c]).max() / scale)
    return max(out)
