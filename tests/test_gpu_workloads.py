"""The batched and served workloads at their stated sizes, against the oracle.

- 27-block density blob: a 3 x 3 x 3 layout of blocks of 32-96 vertices per
  direction (seed 3: 5.45 M vertices), joined by device-side ghost exchange,
  textbook WENO5 + AUSM+, SSP-RK3, CFL 0.5, in one batched launch per stage;
  against the oracle on the union grid (the blocks share their interface
  vertex planes, src/base/defs_block_geom.f90:610-611, as the reference's
  cut does).
- The serving mix: requests drawn from the benchmark's own seeded stream
  (2-D Riemann 64^2-512^2, 3-D vortex / blob 32^3-128^3, MUSCL + Roe or
  WENO5 + AUSM+) plus its extreme sizes, served together with continuous
  batching; each result equals the oracle run with the same CFL steps to the
  same clipped end time.
- Admission by device-memory budget and the refusal of malformed requests.
"""
import numpy as np
import pytest

import _oracle as O
from _parity import rel_linf, run_oracle
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

WENO_AUSMPLUS = dict(recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3,
                     ref_compat=0)


def blob27_union(seed=3):
    sizes, offs, sp, states = cases.blob_blocks(seed)
    ntot = [offs[d][-1] + sizes[d][-1] for d in range(3)]
    p = np.zeros((ntot[2], ntot[1], ntot[0], 5))
    for (a, b, c), pb in states.items():
        x0, y0, z0 = offs[0][a], offs[1][b], offs[2][c]
        nk, nj, ni, _ = pb.shape
        p[z0:z0 + nk, y0:y0 + nj, x0:x0 + ni] = pb
    return sizes, offs, sp, states, p


def run_blob27(states, offs, sp, shape, nsteps, numerics):
    with fvmgpu.Context() as ctx:
        ctx.scheme(**WENO_AUSMPLUS)
        ctx.numerics(numerics)
        blocks = {}
        for key, pb in states.items():
            nk, nj, ni, _ = pb.shape
            blk = ctx.block((ni, nj, nk), sp)
            blk.upload(cases.to_conserved(pb))
            blocks[key] = blk
        for (a, b, c), blk in blocks.items():
            idx = [a, b, c]
            for d in range(3):
                if idx[d] + 1 < 3:
                    nb = list(idx)
                    nb[d] += 1
                    blk.connect(2 * d + 1, blocks[tuple(nb)], shared=True)
        dt = ctx.step(nsteps, cfl=0.5)
        out = np.zeros((5,) + shape)
        for (a, b, c), blk in blocks.items():
            q = blk.download()
            x0, y0, z0 = offs[0][a], offs[1][b], offs[2][c]
            out[:, z0:z0 + q.shape[1], y0:y0 + q.shape[2], x0:x0 + q.shape[3]] = q
        return out, dt, ctx.norms()[0]


def test_blob27_full_size_equals_oracle():
    """The stated configuration itself: seed 3, blocks of 32-96 vertices,
    5.45 M vertices, two CFL 0.5 steps; exact family bit for bit, fast family
    within 1e-12 relative L-inf."""
    sizes, offs, sp, states, p = blob27_union(3)
    nv = sum(pb.shape[0] * pb.shape[1] * pb.shape[2] for pb in states.values())
    assert all(32 <= s <= 96 for d in sizes for s in d) and nv > 5.4e6
    print(f"case blob27 seed 3 sizes {sizes} ({nv} vertices in blocks, union {p.shape[:3]})",
          flush=True)
    O.set_threads(16)
    qo, r = run_oracle(p, sp, 2, cfl=0.5, **WENO_AUSMPLUS)
    q, dt, norms = run_blob27(states, offs, sp, p.shape[:3], 2, fvmgpu.NUMERICS_EXACT)
    assert dt == r["dt"][-1]
    print("exact rel linf", rel_linf(q, qo), flush=True)
    assert np.array_equal(q, qo)
    qf, _, _ = run_blob27(states, offs, sp, p.shape[:3], 2, fvmgpu.NUMERICS_FAST)
    err = rel_linf(qf, qo)
    print("fast rel linf", err)
    assert err <= 1e-12


def short_mix():
    """Requests of the benchmark's stream (seed 0, 96 arrivals at 50 / s) with
    short end times: the largest and smallest of each kind and scheme drawn,
    plus requests at the distribution's edges (512^2, 128^3)."""
    ends = {"riemann2d": 5e-4, "vortex3d": 0.02, "blob3d": 2e-3}
    _, reqs = cases.serving_requests(0, 96, 50.0, t_end=ends)
    pick = {}
    for r in reqs:
        key = (r["label"].split()[0], r["recon"])
        n = r["q0"].shape[-1]
        lo, hi = pick.get(key, (None, None))
        if lo is None or n < lo["q0"].shape[-1]:
            lo = r
        if hi is None or n > hi["q0"].shape[-1]:
            hi = r
        pick[key] = (lo, hi)
    out = []
    for lo, hi in pick.values():
        out.append(lo)
        if hi is not lo:
            out.append(hi)
    p, dx = cases.riemann2d(512, seed=0)
    out.append(dict(q0=cases.to_conserved(p), p0=p, spacing=(dx, dx, 1.0),
                    t_end=ends["riemann2d"], cfl=0.5, label="riemann2d 512^2", **WENO_AUSMPLUS))
    p, h, _ = cases.isentropic_vortex(128, seed=1234)
    out.append(dict(q0=cases.to_conserved(p), p0=p, spacing=(h, h, h), t_end=ends["vortex3d"],
                    cfl=0.5, periodic=(True, True, True), label="vortex3d 128^3", recon=O.MUSCL,
                    limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3, ref_compat=1))
    x = np.linspace(0.0, 1.0, 128)
    p = cases.blob(x, x, x, seed=1235)
    out.append(dict(q0=cases.to_conserved(p), p0=p, spacing=(1 / 127,) * 3, t_end=ends["blob3d"],
                    cfl=0.5, label="blob3d 128^3", **WENO_AUSMPLUS))
    return out


def test_served_mix_equals_oracle():
    reqs = short_mix()
    labels = [r["label"] + (" weno" if r["recon"] == O.WENO else " muscl") for r in reqs]
    print("case serving mix:", ", ".join(labels), flush=True)
    with fvmgpu.Server(max_active=64) as srv:
        ids = [srv.submit(**{k: v for k, v in r.items() if k not in ("label", "p0")})
               for r in reqs]
        while srv.poll(20, 2)[2] < len(ids):
            pass
        got = [srv.result(i) for i in ids]
    O.set_threads(16)
    for r, lab, (q, info) in zip(reqs, labels, got):
        assert info["status"] == 2, lab
        p = r["p0"]                       # the state q0 was converted from
        nk = r["q0"].shape[1]
        dim = 3 if nk > 1 else 2
        bc = (1,) * 6 if any(r.get("periodic", ())) else (0,) * 6
        qo, steps, t = O.oracle_solve_to(p, dim, r["spacing"], r["cfl"], r["t_end"],
                                         recon=r["recon"], limiter=r["limiter"],
                                         flux=r["flux"], rk=r["rk"], bc=bc,
                                         weno_textbook=0 if r["ref_compat"] else 1)
        gz = 3 if dim == 3 else 0
        qo = qo[:, gz:gz + nk, 3:-3, 3:-3]
        print(f"{lab}: {steps} steps to t = {t!r}", flush=True)
        assert info["steps"] == steps and info["t_final"] == t, lab
        if dim == 2:
            qo = qo.copy()
            qo[3] = 0.0
        assert np.array_equal(q, qo), f"{lab}: rel linf {rel_linf(q, qo)}"


def test_admission_held_back_by_memory_budget():
    """A device-memory budget that fits one of two solves but not both: the
    second is admitted only after the first retires; results equal the
    unconstrained run."""
    p1 = cases.smooth_3d((40, 40, 40), seed=1)
    p2 = cases.smooth_3d((36, 36, 36), seed=2)
    reqs = [dict(q0=cases.to_conserved(p), spacing=(0.02,) * 3, t_end=0.01, cfl=0.5,
                 recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4) for p in (p1, p2)]

    def serve(budget):
        with fvmgpu.Server(hbm_budget=budget, max_active=8) as srv:
            ids = [srv.submit(**r) for r in reqs]
            while srv.poll(10, 1)[2] < 2:
                pass
            return [srv.result(i) for i in ids]

    free = serve(0.0)
    b1, b2 = free[0][1]["bytes"], free[1][1]["bytes"]
    assert free[1][1]["t_admit"] < free[0][1]["t_done"]          # both ran together
    held = serve(max(b1, b2) + 0.5 * min(b1, b2))
    assert held[1][1]["t_admit"] >= held[0][1]["t_done"]         # waited for the memory
    for (qa, ia), (qb, ib) in zip(free, held):
        assert ib["status"] == 2 and ia["steps"] == ib["steps"]
        assert np.array_equal(qa, qb)
    with fvmgpu.Server(hbm_budget=0.5 * min(b1, b2)) as srv:      # never fits: refused
        with pytest.raises(fvmgpu.FvmError, match="budget"):
            srv.submit(**reqs[0])


def test_malformed_requests_refused_at_submit():
    """Requests a block or scheme would refuse are refused by submit, so
    admission never meets them; a valid request submitted before them is
    served exactly as alone."""
    p, dx = cases.riemann2d(32, seed=0)
    ok = dict(q0=cases.to_conserved(p), spacing=(dx, dx, 1.0), t_end=2e-3, cfl=0.5)
    with fvmgpu.Server() as srv:
        i = srv.submit(**ok)
        while srv.poll(100, 1)[2] < 1:
            pass
        q_alone, _ = srv.result(i)
    q3 = np.zeros((5, 2, 32, 32))
    q3[0] = 1.0
    q3[4] = 2.5
    with fvmgpu.Server() as srv:
        rid = srv.submit(**ok)
        bad = [dict(ok, q0=q3, spacing=(dx, dx, dx)),                  # nk = 2
               dict(ok, spacing=(dx, 0.0, 1.0)),                      # spacing <= 0
               dict(ok, recon=999),                                   # unknown scheme
               dict(ok, flux=302),
               dict(ok, gamma=1.0),
               dict(ok, numerics=7)]
        for b in bad:
            with pytest.raises(fvmgpu.FvmError):
                srv.submit(**b)
        while srv.poll(100, 1)[2] < 1:
            pass
        q, info = srv.result(rid)
    assert info["status"] == 2
    assert np.array_equal(q, q_alone)
