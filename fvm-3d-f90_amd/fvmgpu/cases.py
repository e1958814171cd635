"""Initial states of the benchmark cases (host side, numpy).

All generators return primitive variables p of shape (nk, nj, ni, 5) in the
order (rho, u, v, w, p) on the real vertices, and `to_conserved` gives the
(5, nk, nj, ni) conserved array in the memory image of the Fortran
q(i, j, k, eq) that fvm_upload expects (host_ng = 0).
"""
import numpy as np


def to_conserved(p, gamma=1.4):
    """p -> q, src/solve/class_flow_field.f90:353-377 (same operation order)."""
    rho, u, v, w, pr = (p[..., m] for m in range(5))
    q = np.empty((5,) + p.shape[:-1])
    q[0] = rho
    q[1] = rho * u
    q[2] = rho * v
    q[3] = rho * w
    q[4] = pr / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return q


def to_primitive(q, gamma=1.4):
    """q -> p, src/solve/class_flow_field.f90:383-396."""
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    p = np.empty(q.shape[1:] + (5,))
    p[..., 0] = rho
    p[..., 1] = u
    p[..., 2] = v
    p[..., 3] = w
    p[..., 4] = (gamma - 1.0) * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
    return p


def nodal_coords(n, length):
    """Vertices x = dx * (i - 1), dx = L / (N - 1) (src/base/method_gridgen.f90:120-131)."""
    dx = length / float(n - 1)
    return dx, dx * np.arange(n, dtype=np.float64)


# Four-quadrant states of the reference's hard-coded case, in the order the
# regions are written (src/solve/class_flow_field.f90:207-243): region 1 is
# the upper-right quadrant x, y in [0.8, 1].
REF_QUADRANTS = (
    ((0.8, 1.0), (0.8, 1.0), (1.5, 0.0, 0.0, 0.0, 1.5)),
    ((0.0, 0.8), (0.8, 1.0), (0.5323, 1.206, 0.0, 0.0, 0.3)),
    ((0.0, 0.8), (0.0, 0.8), (0.138, 1.206, 1.206, 0.0, 0.029)),
    ((0.8, 1.0), (0.0, 0.8), (0.5323, 0.0, 1.206, 0.0, 0.3)),
)


def riemann2d(n, seed=0, mid=0.8):
    """2-D four-quadrant Riemann problem on the unit square, n x n vertices.

    seed 0: the reference's states (Lax-Liu configuration 3).  Regions are
    applied in order with inclusive bounds (src/base/defs_block_data.f90:
    166-201), so later regions win on the x = 0.8 / y = 0.8 lines.
    seed >= 1: each quadrant draws rho in U[0.1, 2], p in U[0.02, 1.5],
    u, v in U[-1.3, 1.3] from numpy's PCG64 stream of that seed.
    Returns (p, dx)."""
    dx, x = nodal_coords(n, 1.0)
    if seed == 0:
        quads = REF_QUADRANTS
        xm = ym = mid
    else:
        rng = np.random.default_rng(seed)
        xm = ym = mid
        quads = []
        for (xr, yr) in (((xm, 1.0), (ym, 1.0)), ((0.0, xm), (ym, 1.0)),
                         ((0.0, xm), (0.0, ym)), ((xm, 1.0), (0.0, ym))):
            rho = rng.uniform(0.1, 2.0)
            pr = rng.uniform(0.02, 1.5)
            u, v = rng.uniform(-1.3, 1.3, 2)
            quads.append((xr, yr, (rho, u, v, 0.0, pr)))
    p = np.zeros((1, n, n, 5))
    X, Y = np.meshgrid(x, x, indexing="xy")
    for (xa, xb), (ya, yb), state in quads:
        m = (X >= xa) & (X <= xb) & (Y >= ya) & (Y <= yb)
        p[0][m] = state
    return p, dx


def isentropic_vortex(n, seed=0, length=10.0, beta=5.0, gamma=1.4, u_inf=(1.0, 1.0, 0.0),
                      center=None):
    """Yee's isentropic vortex extruded along z on a periodic [0, L]^3 box of
    n^3 vertices, x = i h with h = L / n (vertex n is vertex 0).  Free stream
    rho = p = 1.  Returns (p, h, centre)."""
    h = length / n
    x = h * np.arange(n, dtype=np.float64)
    if center is None:
        if seed == 0:
            center = (0.5 * length, 0.5 * length)
        else:
            rng = np.random.default_rng(seed)
            center = tuple(rng.uniform(0.3 * length, 0.7 * length, 2))
    p = np.empty((n, n, n, 5))
    prim2 = vortex_2d(x, x, center, length, beta, gamma, u_inf)
    p[:] = prim2[None]
    return p, h, center


def vortex_2d(x, y, center, length, beta, gamma, u_inf, t=0.0):
    """Exact vortex state at time t (translated with the free stream, wrapped
    periodically), shape (ny, nx, 5)."""
    X, Y = np.meshgrid(x, y, indexing="xy")
    xc = (center[0] + u_inf[0] * t) % length
    yc = (center[1] + u_inf[1] * t) % length
    dxv = (X - xc + 0.5 * length) % length - 0.5 * length
    dyv = (Y - yc + 0.5 * length) % length - 0.5 * length
    r2 = dxv * dxv + dyv * dyv
    e = np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * np.pi * np.pi) * e * e
    rho = T ** (1.0 / (gamma - 1.0))
    out = np.empty(X.shape + (5,))
    out[..., 0] = rho
    out[..., 1] = u_inf[0] - beta / (2.0 * np.pi) * e * dyv
    out[..., 2] = u_inf[1] + beta / (2.0 * np.pi) * e * dxv
    out[..., 3] = u_inf[2]
    out[..., 4] = rho ** gamma
    return out


def smooth_3d(shape, seed=0, amp=0.2):
    """Smooth positive random field for parity tests: (nk, nj, ni, 5)."""
    nk, nj, ni = shape
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.linspace(0.0, 1.0, m) for m in (nk, nj, ni)), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 8)
    p = np.empty((nk, nj, ni, 5))
    p[..., 0] = 1.0 + amp * np.sin(2 * np.pi * x + ph[0]) * np.cos(2 * np.pi * y + ph[1]) + 0.1 * z
    p[..., 1] = 0.3 * np.sin(2 * np.pi * y + ph[2])
    p[..., 2] = 0.2 * np.cos(2 * np.pi * z + ph[3])
    p[..., 3] = 0.25 * np.sin(2 * np.pi * x + ph[4])
    p[..., 4] = 1.0 + amp * np.cos(2 * np.pi * (x + y + z) + ph[5])
    return p


def blob(x, y, z, seed=0, nblob=8, amp=0.2, umax=0.5):
    """Density-blob state on the unit cube at the coordinate vectors x, y, z:
    rho = 1 + amp * sum of nblob Gaussians (centres U[0.15, 0.85]^3, sigma
    U[0.05, 0.2]), p = 1, and a smooth divergence-free-ish velocity built from
    a few sine modes scaled so |u| < umax.  Returns (nk, nj, ni, 5)."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0.15, 0.85, (nblob, 3))
    sig = rng.uniform(0.05, 0.2, nblob)
    ph = rng.uniform(0.0, 2.0 * np.pi, 6)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.ones(X.shape)
    for (cx, cy, cz), s in zip(cen, sig):
        r2 = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2
        rho += amp * np.exp(-0.5 * r2 / (s * s))
    a = umax / np.sqrt(3.0) * 0.99
    p = np.empty(X.shape + (5,))
    p[..., 0] = rho
    p[..., 1] = a * np.sin(2 * np.pi * Y + ph[0]) * np.cos(2 * np.pi * Z + ph[1])
    p[..., 2] = a * np.sin(2 * np.pi * Z + ph[2]) * np.cos(2 * np.pi * X + ph[3])
    p[..., 3] = a * np.sin(2 * np.pi * X + ph[4]) * np.cos(2 * np.pi * Y + ph[5])
    p[..., 4] = 1.0
    return p


def blob_blocks(seed=0, layout=(3, 3, 3), lo=32, hi=96):
    """Conforming multi-block layout of the unit cube: block (a, b, c) has
    nx[a] x ny[b] x nz[c] vertices, sizes drawn from U{lo..hi} with the seed;
    neighbouring blocks share their interface vertex plane, one spacing per
    direction.  Returns (sizes per direction, global vertex offsets per
    direction, spacing, per-block primitive states keyed (a, b, c))."""
    rng = np.random.default_rng(seed)
    sizes = [list(int(v) for v in rng.integers(lo, hi + 1, layout[d])) for d in range(3)]
    offs, coords, spacing = [], [], []
    for d in range(3):
        o = np.concatenate(([0], np.cumsum(np.array(sizes[d]) - 1)[:-1]))
        ntot = int(sum(sizes[d]) - (layout[d] - 1))
        h = 1.0 / (ntot - 1)
        offs.append([int(v) for v in o])
        coords.append(h * np.arange(ntot, dtype=np.float64))
        spacing.append(h)
    full = blob(coords[0], coords[1], coords[2], seed=seed)
    states = {}
    for a in range(layout[0]):
        for b in range(layout[1]):
            for c in range(layout[2]):
                x0, y0, z0 = offs[0][a], offs[1][b], offs[2][c]
                states[(a, b, c)] = full[z0:z0 + sizes[2][c], y0:y0 + sizes[1][b],
                                         x0:x0 + sizes[0][a]].copy()
    return sizes, offs, tuple(spacing), states


# Taylor-Green vortex (Brachet et al.; the laminar-transition benchmark
# case): non-dimensional with rho0 = V0 = L = 1, T0 = 1, gas constant
# R = 1 / (gamma Ma^2) so p0 = rho0 R T0 and the Mach number is Ma;
# mu_ref = rho0 V0 L / Re at T0; Sutherland constant 110.4 K over 273.15 K.
TGV_SUTHERLAND_S = 110.4 / 273.15


def tgv_params(re=1600.0, ma=0.1, pr=0.71, gamma=1.4):
    """(mu_ref, t_ref, sutherland_s, prandtl, r_gas) of the TGV case."""
    return (1.0 / re, 1.0, TGV_SUTHERLAND_S, pr, 1.0 / (gamma * ma * ma))


def tgv(n, ma=0.1, gamma=1.4):
    """TGV on the periodic box [0, 2 pi)^3, n^3 vertices x = i h, h = 2 pi / n
    (vertex n is vertex 0):
      u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
      p = p0 + (cos 2x + cos 2y)(cos 2z + 2) / 16,  rho = p / (R T0).
    Returns (p (n, n, n, 5), h)."""
    r = 1.0 / (gamma * ma * ma)
    h = 2.0 * np.pi / n
    x = h * np.arange(n, dtype=np.float64)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    pr = r + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0) / 16.0
    p = np.empty(X.shape + (5,))
    p[..., 0] = pr / r
    p[..., 1] = np.sin(X) * np.cos(Y) * np.cos(Z)
    p[..., 2] = -np.cos(X) * np.sin(Y) * np.cos(Z)
    p[..., 3] = 0.0
    p[..., 4] = pr
    return p, h


def kinetic_energy(q):
    """Volume-mean kinetic energy 0.5 rho |u|^2 / rho0 of a conserved state
    (5, nk, nj, ni) on a periodic box (each vertex one cell)."""
    rho = q[0]
    return float(np.mean(0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / rho))


# ---------------------------------------------------------------------------
# The mixed serving workload: Poisson arrivals of independent solves, 2-D
# four-quadrant Riemann cases of 64^2-512^2 vertices and 3-D vortex / blob
# cases of 32^3-128^3, sizes log-uniform, scheme MUSCL-minmod + Roe or textbook
# WENO5 + AUSM+ (SSP-RK3, CFL 0.5, FP64), each to a fixed end time.
SERVE_T_END = {"riemann2d": 0.1, "vortex3d": 1.0, "blob3d": 0.1}

# scheme codes (src/defs/type_defs_scheme.f90:11-36; 1303: AUSM+)
_MUSCL_ROE = dict(recon=201, limiter=251, flux=300, rk=101, ref_compat=1)
_WENO_AUSMPLUS = dict(recon=202, limiter=250, flux=1303, rk=101, ref_compat=0)


def serving_requests(seed, nreq, rate, t_end=None):
    """(arrival times, requests) of the serving workload: each request a dict
    of Server.submit arguments plus a `label` and the primitive state `p0`
    that q0 was converted from.  t_end: a dict of per-kind end
    times replacing SERVE_T_END (e.g. a short run for tests)."""
    ends = dict(SERVE_T_END)
    ends.update(t_end or {})
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    arrivals -= arrivals[0]
    reqs = []
    for i in range(nreq):
        kind = rng.integers(3)            # 0: Riemann 2-D, 1: vortex 3-D, 2: blob 3-D
        sch = dict((_MUSCL_ROE, _WENO_AUSMPLUS)[rng.integers(2)])
        if kind == 0:
            n = int(round(np.exp(rng.uniform(np.log(64), np.log(512)))))
            # the reference's quadrant states (Lax-Liu configuration 3): the
            # drawn states of other seeds include near-vacuum rarefactions in
            # which Roe's linearisation loses positivity
            p, dx = riemann2d(n, seed=0)
            r = dict(q0=to_conserved(p), p0=p, spacing=(dx, dx, 1.0), t_end=ends["riemann2d"],
                     cfl=0.5, label=f"riemann2d {n}^2")
        else:
            n = int(round(np.exp(rng.uniform(np.log(32), np.log(128)))))
            if kind == 1:
                p, h, _ = isentropic_vortex(n, seed=1000 + i)
                r = dict(q0=to_conserved(p), p0=p, spacing=(h, h, h), t_end=ends["vortex3d"],
                         cfl=0.5, periodic=(True, True, True), label=f"vortex3d {n}^3")
            else:
                x = np.linspace(0.0, 1.0, n)
                p = blob(x, x, x, seed=1000 + i)
                h = 1.0 / (n - 1)
                r = dict(q0=to_conserved(p), p0=p, spacing=(h, h, h), t_end=ends["blob3d"],
                         cfl=0.5, label=f"blob3d {n}^3")
        r.update(sch)
        reqs.append(r)
    return arrivals, reqs
