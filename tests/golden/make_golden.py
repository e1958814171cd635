"""Writes the golden fixtures in this directory from the reference's own
compiled scheme modules (oracle/_ref/libfvmref.so, built by
`make -C oracle ref` from the sources under /root/reference).

The reference ships no test vectors, known-answer tests or fixtures; these
files are inputs and outputs of its own FP64 arithmetic (-fdefault-real-8,
no FMA contraction) on seeded and hand-picked edge-case inputs:

  scheme_vectors.npz  per-function vectors: Simple/MUSCL (every limiter)/WENO
                      reconstruction on pencils, Roe/Van Leer/AUSM face fluxes
                      in every direction, SSP-RK3 combine stages
  block_states.npz    whole-block runs of the reference driver loop
                      (ref_harness/ref_driver.f90 around the compiled scheme
                      modules): Riemann-2D 21x21 and 64x64, a smooth 3-D block

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402


def prim_states(rng, n, supersonic=False):
    s = np.empty((n, 5))
    s[:, 0] = rng.uniform(0.1, 2.0, n)
    vmax = 4.0 if supersonic else 1.5
    s[:, 1:4] = rng.uniform(-vmax, vmax, (n, 3))
    s[:, 4] = rng.uniform(0.02, 1.5, n)
    return s


def edge_pencil():
    """Pencil with flat runs (zero slopes), extrema, equal slopes, signed
    zeros and a jump: the limiter branches of scheme_reconstruct_MUSCL.f90."""
    base = np.array([1.0, 1.0, 1.0, 2.0, 3.0, 4.0, 4.0, 3.0, 2.0, 2.0, 5.0, 0.5, 0.5, 0.5,
                     1.5, 2.5, 3.5, 3.5, 1.0, 1.0, 1.0, 0.2, 0.2, 0.2])
    pg = np.empty((base.size, 5))
    pg[:, 0] = base
    pg[:, 1] = np.where(np.arange(base.size) % 5 == 0, -0.0, np.sin(base))
    pg[:, 2] = -0.0
    pg[:, 3] = 0.0
    pg[:, 4] = base * 0.5
    return pg


def flux_edge_states():
    """Face pairs on the branch boundaries of the flux routines: equal states,
    |M| = 1 exactly, supersonic both ways, stagnation (Harten fix active)."""
    c = np.sqrt(1.4 * 1.0 / 1.0)
    L, R = [], []
    L.append([1.0, 0.0, 0.0, 0.0, 1.0]); R.append([1.0, 0.0, 0.0, 0.0, 1.0])
    L.append([1.0, c, 0.1, -0.2, 1.0]); R.append([1.0, c, 0.1, -0.2, 1.0])
    L.append([1.0, -c, 0.0, 0.0, 1.0]); R.append([0.9, -c, 0.0, 0.0, 0.9])
    L.append([1.0, 3.0, 0.5, 0.1, 1.0]); R.append([0.8, 2.8, 0.4, 0.0, 0.7])
    L.append([1.0, -3.0, 0.5, 0.1, 1.0]); R.append([0.8, -2.8, 0.4, 0.0, 0.7])
    L.append([1.5, 0.0, 0.0, 0.0, 1.5]); R.append([0.138, 1.206, 1.206, 0.0, 0.029])
    L.append([1.0, 1e-12, 0.0, 0.0, 1.0]); R.append([1.0, -1e-12, 0.0, 0.0, 1.0])
    L.append([0.5323, 1.206, 0.0, 0.0, 0.3]); R.append([1.5, 0.0, 0.0, 0.0, 1.5])
    return np.array(L), np.array(R)


def scheme_vectors(ref):
    rng = np.random.default_rng(20261203)
    out = {}
    pencils = {"rand": prim_states(rng, 30), "edge": edge_pencil()}
    for pname, pg in pencils.items():
        npg = pg.shape[0]
        nface = npg - 6 + 1
        out[f"pencil_{pname}"] = pg
        for w in (-1, 1):
            out[f"simple_{pname}_{w}"] = O.recon(ref, O.SIMPLE, pg, nface, 3, w)
            for lim in (O.LIM_NONE, O.LIM_MINMOD, O.LIM_VANLEER, O.LIM_VANALBADA):
                out[f"muscl_{pname}_{lim}_{w}"] = O.recon(ref, O.MUSCL, pg, nface, 3, w,
                                                          kappa=0.33333, limiter=lim)
            qg = np.abs(pg) + 0.1
            out[f"wenoin_{pname}"] = qg
            out[f"weno_{pname}_{w}"] = O.recon(ref, O.WENO, qg, nface, 3, w)
    ql = np.vstack([prim_states(rng, 200), prim_states(rng, 100, supersonic=True),
                    flux_edge_states()[0]])
    qr = np.vstack([prim_states(rng, 200), prim_states(rng, 100, supersonic=True),
                    flux_edge_states()[1]])
    out["flux_ql"], out["flux_qr"] = ql, qr
    for kind, name in ((O.ROE, "roe"), (O.VANLEER, "vanleer"), (O.AUSM, "ausm")):
        for d in range(3):
            out[f"flux_{name}_{d}"] = O.flux(ref, kind, ql, qr, d, gamma=1.4, entfix=0.2)
    u0, um, r = rng.normal(size=64), rng.normal(size=64), rng.normal(size=64)
    out["rk_u0"], out["rk_um"], out["rk_r"] = u0, um, r
    for m in (1, 2, 3):
        out[f"rk_{m}"] = O.runge_kutta(ref, m, 2e-4, u0, um, r)[1]
    return out


def riemann_ic(n):
    """The reference's hard-coded quadrants (src/solve/class_flow_field.f90:
    207-243) on the nodal grid x = dx*(i-1) (src/base/method_gridgen.f90:
    120-131), regions applied in order with inclusive bounds."""
    dx = 1.0 / (n - 1)
    x = dx * np.arange(n)
    p = np.zeros((2, n, n, 5))   # the reference's 2-D grid has two k-planes
    X, Y = np.meshgrid(x, x, indexing="xy")
    regs = [((0.8, 1.0), (0.8, 1.0), (1.5, 0.0, 0.0, 0.0, 1.5)),
            ((0.0, 0.8), (0.8, 1.0), (0.5323, 1.206, 0.0, 0.0, 0.3)),
            ((0.0, 0.8), (0.0, 0.8), (0.138, 1.206, 1.206, 0.0, 0.029)),
            ((0.8, 1.0), (0.0, 0.8), (0.5323, 0.0, 1.206, 0.0, 0.3))]
    for (xa, xb), (ya, yb), st in regs:
        m = (X >= xa) & (X <= xb) & (Y >= ya) & (Y <= yb)
        for k in range(2):
            p[k][m] = st
    return p, dx


COMBOS = [(O.SIMPLE, O.LIM_NONE), (O.MUSCL, O.LIM_MINMOD), (O.MUSCL, O.LIM_VANLEER),
          (O.MUSCL, O.LIM_VANALBADA), (O.WENO, O.LIM_NONE)]
FLUXES = [O.ROE, O.VANLEER, O.AUSM]


def block_states(ref):
    out = {}
    p21, dx21 = riemann_ic(21)
    out["riemann21_p"] = p21
    out["riemann21_dx"] = np.array(dx21)
    r0 = O.ref_solve(p21, 2, (dx21, dx21, 1.0), nsteps=0)
    out["riemann21_q0"] = r0["q"][:, 3, 3:-3, 3:-3]     # initial conserved state
    for rc, lim in COMBOS:
        for fl in FLUXES:
            for steps in (1, 10):
                for pfq in ((0, 1) if (rc, lim, fl) == (O.MUSCL, O.LIM_MINMOD, O.ROE) else (0,)):
                    r = O.ref_solve(p21, 2, (dx21, dx21, 1.0), recon=rc, limiter=lim, flux=fl,
                                    kappa=0.33333, entfix=0.2, dt=2e-4, nsteps=steps,
                                    p_from_q=pfq)
                    key = f"riemann21_{rc}_{lim}_{fl}_{steps}_{pfq}"
                    out[key + "_q"] = r["q"][:, 3, 3:-3, 3:-3]   # plane k = 1, real region
                    out[key + "_res"] = r["res"]
    p64, dx64 = riemann_ic(64)
    out["riemann64_p"] = p64
    r = O.ref_solve(p64, 2, (dx64, dx64, 1.0), dt=2e-4, nsteps=10, p_from_q=0)
    out["riemann64_q"] = r["q"][:, 3, 3:-3, 3:-3]
    out["riemann64_res"] = r["res"]
    # smooth 3-D block, every combination, 3 steps
    n = (10, 9, 8)
    z, y, x = np.meshgrid(*(np.linspace(0.0, 1.0, m) for m in n[::-1]), indexing="ij")
    p3 = np.empty(n[::-1] + (5,))
    p3[..., 0] = 1 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.1 * z
    p3[..., 1] = 0.3 * np.sin(2 * np.pi * y)
    p3[..., 2] = 0.2 * np.cos(2 * np.pi * z)
    p3[..., 3] = 0.25 * np.sin(2 * np.pi * x)
    p3[..., 4] = 1 + 0.2 * np.cos(2 * np.pi * (x + y + z))
    dr = (1 / 9, 1 / 8, 1 / 7)
    out["smooth3d_p"] = p3
    out["smooth3d_dr"] = np.array(dr)
    for rc, lim in COMBOS:
        for fl in FLUXES:
            r = O.ref_solve(p3, 3, dr, recon=rc, limiter=lim, flux=fl, dt=1e-3, nsteps=3,
                            p_from_q=1)
            key = f"smooth3d_{rc}_{lim}_{fl}"
            out[key + "_q"] = r["q"][:, 3:-3, 3:-3, 3:-3]
            out[key + "_res"] = r["res"]
    return out


def main():
    ref = O.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libfvmref.so is not built (make -C oracle ref)")
    np.savez_compressed(os.path.join(HERE, "scheme_vectors.npz"), **scheme_vectors(ref))
    np.savez_compressed(os.path.join(HERE, "block_states.npz"), **block_states(ref))
    for f in ("scheme_vectors.npz", "block_states.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
