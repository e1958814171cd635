"""ctypes bindings for the oracle (oracle/liboracle.so) and, where it has been
built in this checkout, the reference harness (oracle/_ref/libfvmref.so).

Test infrastructure only: these libraries check results, never produce them.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfvmref.so")

i64, ci, cd, P = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p

# scheme codes, src/defs/type_defs_scheme.f90:11-36 (>= 1000: additions)
RK_SSP3, RK4 = 101, 1001
SIMPLE, MUSCL, WENO = 200, 201, 202
LIM_NONE, LIM_MINMOD, LIM_VANLEER, LIM_VANALBADA = 250, 251, 252, 253
ROE, VANLEER, AUSM, AUSMPLUS = 300, 301, 303, 1303
BC_EXTRAP, BC_PERIODIC = 0, 1
BC_WALL, BC_NOSLIP, BC_SYMMETRY = 30, 33, 50      # slip wall, adiabatic no-slip wall


class OrcConfig(ctypes.Structure):
    _fields_ = [("dim", ci), ("recon", ci), ("limiter", ci), ("flux", ci), ("rk", ci),
                ("weno_textbook", ci), ("p_from_q", ci), ("bc", ci * 6),
                ("kappa", cd), ("entfix", cd), ("gamma", cd), ("dr", cd * 3),
                ("visc_mu_ref", cd), ("visc_t_ref", cd), ("visc_s", cd), ("visc_pr", cd),
                ("visc_r", cd)]


def _ptr(a):
    return a.ctypes.data_as(P)


def _bind_common(lib, prefix):
    getattr(lib, prefix + "recon_simple").argtypes = [i64, i64, i64, P, P, ci]
    getattr(lib, prefix + "recon_muscl").argtypes = [i64, i64, cd, ci, i64, P, P, ci]
    for f in ("flux_vanleer", "flux_ausm"):
        getattr(lib, prefix + f).argtypes = [i64, ci, cd, P, P, P, P]
    getattr(lib, prefix + "flux_roe").argtypes = [i64, ci, cd, P, P, P, P, cd]
    getattr(lib, prefix + "runge_kutta").argtypes = [ci, cd, i64, P, P, P]
    for name in ("recon_simple", "recon_muscl", "flux_vanleer", "flux_ausm", "flux_roe",
                 "runge_kutta"):
        getattr(lib, prefix + name).restype = None


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        _bind_common(lib, "orc_")
        lib.orc_recon_weno.argtypes = [i64, ci, i64, i64, P, P, ci, ci]
        lib.orc_flux_ausmplus.argtypes = [i64, ci, cd, P, P, P, P]
        lib.orc_solve.argtypes = [P, ctypes.POINTER(OrcConfig), cd, cd, ci, P, P, P, P, P]
        lib.orc_solve.restype = ci
        lib.orc_solve_q.argtypes = [P, ctypes.POINTER(OrcConfig), cd, cd, ci, P, P, P, P, P]
        lib.orc_solve_q.restype = ci
        lib.orc_solve_to.argtypes = [P, ctypes.POINTER(OrcConfig), cd, cd, ci, P, P,
                                     ctypes.POINTER(ci), ctypes.POINTER(cd)]
        lib.orc_solve_to.restype = ci
        lib.orc_cfl_dt.argtypes = [P, ci, P, cd, P, cd]
        lib.orc_cfl_dt.restype = cd
        lib.orc_rk4_stage.argtypes = [ci, cd, i64, P, P, P, P]
        lib.orc_rk4_stage.restype = None
        lib.orc_set_threads.argtypes = [ci]
        lib.orc_set_threads.restype = ci
        lib.orc_sutherland.argtypes = [cd, cd, cd, cd]
        lib.orc_sutherland.restype = cd
        lib.orc_viscous_residual.argtypes = [P, ctypes.c_void_p, P, P]
        lib.orc_viscous_residual.restype = None
        _oracle = lib
    return _oracle


def ref():
    """The reference harness, or None where it is not built (no reference sources)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        lib = ctypes.CDLL(REF_SO)
        _bind_common(lib, "ref_")
        lib.ref_recon_weno.argtypes = [i64, ci, i64, i64, P, P, ci]
        lib.ref_solve.argtypes = [i64, i64, i64, ci, cd, cd, cd, ci, ci, ci, cd, cd, cd, cd,
                                  ci, ci, P, P, P, P]
        lib.ref_solve.restype = None
        _ref = lib
    return _ref


# ---------------------------------------------------------------- per-function
def recon(lib, kind, pg, nface, dghost, wind, kappa=0.33333, limiter=LIM_MINMOD,
          textbook=0):
    """Face states of one pencil.  pg: (npg, 5) AoS primitives (or (npg, 5)
    conserved for WENO).  Returns (nface, 5)."""
    pg = np.ascontiguousarray(pg, dtype=np.float64)
    npg = pg.shape[0]
    out = np.zeros((nface, 5))
    pre = "orc_" if lib is oracle() else "ref_"
    if kind == SIMPLE:
        getattr(lib, pre + "recon_simple")(nface, dghost, npg, _ptr(pg), _ptr(out), wind)
    elif kind == MUSCL:
        getattr(lib, pre + "recon_muscl")(nface, dghost, kappa, limiter, npg, _ptr(pg),
                                          _ptr(out), wind)
    elif kind == WENO:
        qg = np.asfortranarray(pg)            # (npg, neq) column-major
        qr = np.zeros((nface, 5), order="F")
        if pre == "orc_":
            lib.orc_recon_weno(nface, 5, dghost, npg, _ptr(qg), _ptr(qr), wind, textbook)
        else:
            lib.ref_recon_weno(nface, 5, dghost, npg, _ptr(qg), _ptr(qr), wind)
        out = np.ascontiguousarray(qr)
    return out


def flux(lib, kind, ql, qr, direction, gamma=1.4, entfix=0.2):
    """Face fluxes for AoS primitive states ql, qr of shape (nface, 5)."""
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    nface = ql.shape[0]
    norm = np.zeros((nface, 3), order="F")
    norm[:, direction] = 1.0
    out = np.zeros((nface, 5), order="F")
    pre = "orc_" if lib is oracle() else "ref_"
    if kind == ROE:
        getattr(lib, pre + "flux_roe")(nface, 5, gamma, _ptr(norm), _ptr(ql), _ptr(qr),
                                       _ptr(out), entfix)
    elif kind == VANLEER:
        getattr(lib, pre + "flux_vanleer")(nface, 5, gamma, _ptr(norm), _ptr(ql), _ptr(qr),
                                           _ptr(out))
    elif kind == AUSM:
        getattr(lib, pre + "flux_ausm")(nface, 5, gamma, _ptr(norm), _ptr(ql), _ptr(qr),
                                        _ptr(out))
    elif kind == AUSMPLUS:
        lib.orc_flux_ausmplus(nface, 5, gamma, _ptr(norm), _ptr(ql), _ptr(qr), _ptr(out))
    return np.ascontiguousarray(out)


def runge_kutta(lib, m, dt, u0, um, rm):
    u0 = np.array(u0, dtype=np.float64)
    um = np.array(um, dtype=np.float64)
    rm = np.ascontiguousarray(rm, dtype=np.float64)
    pre = "orc_" if lib is oracle() else "ref_"
    getattr(lib, pre + "runge_kutta")(m, dt, u0.size, _ptr(u0), _ptr(um), _ptr(rm))
    return u0, um


# ---------------------------------------------------------------- whole block
def oracle_solve(p_real, dim, dr, recon=MUSCL, limiter=LIM_MINMOD, flux=ROE, rk=RK_SSP3,
                 kappa=0.33333, entfix=0.2, gamma=1.4, dt=2e-4, cfl=0.0, nsteps=1,
                 p_from_q=1, bc=(BC_EXTRAP,) * 6, weno_textbook=0, visc=None):
    """p_real: (nk, nj, ni, 5) primitives on real vertices.  visc: None or
    (mu_ref, t_ref, sutherland_s, prandtl, r_gas) for the viscous terms.
    Returns dict with q (5, nk+2gz, nj+6, ni+6), rhs (5, nk, nj, ni),
    res (nsteps, 6), dt (nsteps)."""
    p_real = np.ascontiguousarray(p_real, dtype=np.float64)
    nk, nj, ni, _ = p_real.shape
    gz = 3 if dim == 3 else 0
    n = np.array([ni, nj, nk], dtype=np.int64)
    cfg = OrcConfig(dim, recon, limiter, flux, rk, weno_textbook, p_from_q,
                    (ci * 6)(*bc), kappa, entfix, gamma, (cd * 3)(*dr), *(visc or (0.0,) * 5))
    q = np.zeros((5, nk + 2 * gz, nj + 6, ni + 6))
    rhs = np.zeros((5, nk, nj, ni))
    res = np.zeros((max(nsteps, 1), 6))
    dts = np.zeros(max(nsteps, 1))
    rc = oracle().orc_solve(_ptr(n), ctypes.byref(cfg), dt, cfl, nsteps, _ptr(p_real),
                            _ptr(q), _ptr(rhs), _ptr(res), _ptr(dts))
    assert rc == 0
    return {"q": q, "rhs": rhs, "res": res[:nsteps], "dt": dts[:nsteps]}


def oracle_solve_q(q_real, dim, dr, recon=MUSCL, limiter=LIM_MINMOD, flux=ROE, rk=RK_SSP3,
                   kappa=0.33333, entfix=0.2, gamma=1.4, dt=2e-4, cfl=0.0, nsteps=1,
                   bc=(BC_EXTRAP,) * 6, weno_textbook=0):
    """oracle_solve from a conserved state q_real (5, nk, nj, ni) on real
    vertices (primitives derived from it): nsteps from an earlier run's final
    q continue that run bit for bit.  Returns dict as oracle_solve's, plus
    q_real, the final real vertices."""
    q_real = np.ascontiguousarray(q_real, dtype=np.float64)
    _, nk, nj, ni = q_real.shape
    gz = 3 if dim == 3 else 0
    n = np.array([ni, nj, nk], dtype=np.int64)
    cfg = OrcConfig(dim, recon, limiter, flux, rk, weno_textbook, 1, (ci * 6)(*bc), kappa,
                    entfix, gamma, (cd * 3)(*dr), 0.0, 0.0, 0.0, 0.0, 0.0)
    q = np.zeros((5, nk + 2 * gz, nj + 6, ni + 6))
    rhs = np.zeros((5, nk, nj, ni))
    res = np.zeros((max(nsteps, 1), 6))
    dts = np.zeros(max(nsteps, 1))
    rc = oracle().orc_solve_q(_ptr(n), ctypes.byref(cfg), dt, cfl, nsteps, _ptr(q_real),
                              _ptr(q), _ptr(rhs), _ptr(res), _ptr(dts))
    assert rc == 0
    return {"q": q, "q_real": q[:, gz:gz + nk, 3:3 + nj, 3:3 + ni].copy(), "rhs": rhs,
            "res": res[:nsteps], "dt": dts[:nsteps]}


def oracle_solve_to(p_real, dim, dr, cfl, t_end, max_steps=100000, recon=MUSCL,
                    limiter=LIM_MINMOD, flux=ROE, rk=RK_SSP3, kappa=0.33333, entfix=0.2,
                    gamma=1.4, bc=(BC_EXTRAP,) * 6, weno_textbook=0):
    """CFL steps to t_end (the last one clipped), as a served solve takes
    them.  Returns (q (5, nk+2gz, nj+6, ni+6), steps, final time)."""
    p_real = np.ascontiguousarray(p_real, dtype=np.float64)
    nk, nj, ni, _ = p_real.shape
    gz = 3 if dim == 3 else 0
    n = np.array([ni, nj, nk], dtype=np.int64)
    cfg = OrcConfig(dim, recon, limiter, flux, rk, weno_textbook, 1, (ci * 6)(*bc), kappa,
                    entfix, gamma, (cd * 3)(*dr), 0.0, 0.0, 0.0, 0.0, 0.0)
    q = np.zeros((5, nk + 2 * gz, nj + 6, ni + 6))
    steps, t = ci(0), cd(0.0)
    rc = oracle().orc_solve_to(_ptr(n), ctypes.byref(cfg), cfl, t_end, max_steps, _ptr(p_real),
                               _ptr(q), ctypes.byref(steps), ctypes.byref(t))
    assert rc == 0
    return q, steps.value, t.value


def ref_solve(p_real, idim, dr, recon=MUSCL, limiter=LIM_MINMOD, flux=ROE, kappa=0.33333,
              entfix=0.2, gamma=1.4, dt=2e-4, nsteps=1, p_from_q=0):
    """The reference harness (3 ghost layers in every direction, as the
    reference allocates them).  Returns q (5, nk+6, nj+6, ni+6)."""
    lib = ref()
    p_real = np.ascontiguousarray(p_real, dtype=np.float64)
    nk, nj, ni, _ = p_real.shape
    q = np.zeros((5, nk + 6, nj + 6, ni + 6))
    rhs = np.zeros((5, nk, nj, ni))
    res = np.zeros((max(nsteps, 1), 6))
    lib.ref_solve(ni, nj, nk, idim, dr[0], dr[1], dr[2], recon, limiter, flux, kappa, entfix,
                  gamma, dt, nsteps, p_from_q, _ptr(p_real), _ptr(q), _ptr(rhs), _ptr(res))
    return {"q": q, "rhs": rhs, "res": res[:nsteps]}


def rk4_stage(s, dt, u0, um, acc, rm):
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    um = np.array(um, dtype=np.float64)
    acc = np.array(acc, dtype=np.float64)
    rm = np.ascontiguousarray(rm, dtype=np.float64)
    oracle().orc_rk4_stage(s, dt, u0.size, _ptr(u0), _ptr(um), _ptr(acc), _ptr(rm))
    return um, acc


def cfl_dt(q_real, dim, dr, gamma=1.4, cfl=1.0):
    """q_real: (5, nk, nj, ni)."""
    q_real = np.ascontiguousarray(q_real, dtype=np.float64)
    n = np.array(q_real.shape[1:][::-1], dtype=np.int64)
    dr = np.ascontiguousarray(dr, dtype=np.float64)
    return oracle().orc_cfl_dt(_ptr(n), dim, _ptr(q_real), gamma, _ptr(dr), cfl)


def sutherland(t, mu_ref, t_ref, s):
    return oracle().orc_sutherland(t, mu_ref, t_ref, s)


def viscous_residual(p_real, dim, dr, visc, gamma=1.4, bc=(BC_EXTRAP,) * 6):
    """Viscous residual alone of a primitive field p_real (nk, nj, ni, 5);
    returns (5, nk, nj, ni) (component 0 is zero)."""
    p_real = np.ascontiguousarray(p_real, dtype=np.float64)
    nk, nj, ni, _ = p_real.shape
    n = np.array([ni, nj, nk], dtype=np.int64)
    cfg = OrcConfig(dim, MUSCL, LIM_MINMOD, ROE, RK_SSP3, 0, 1, (ci * 6)(*bc), 0.33333, 0.2,
                    gamma, (cd * 3)(*dr), *visc)
    rv = np.zeros((5, nk, nj, ni))
    oracle().orc_viscous_residual(_ptr(n), ctypes.byref(cfg), _ptr(p_real), _ptr(rv))
    return rv


def set_threads(n):
    """Threads of the oracle's pencil loops; returns the count in use."""
    return oracle().orc_set_threads(int(n))
