"""The serving loop: continuous batching of independent solves.  Each
result must be the same alone and in the mix (bit for bit: same kernels,
per-solve CFL step), and the same as the oracle run to the same final time."""
import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu


def mixed_requests():
    reqs = []
    for n, seed in ((48, 1), (64, 2), (33, 3)):
        p, dx = cases.riemann2d(n, seed=seed)
        reqs.append(dict(q0=cases.to_conserved(p), spacing=(dx, dx, 1.0), t_end=0.01, cfl=0.5,
                         recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3))
    for n, seed in ((20, 4), (28, 5)):
        p, h, _ = cases.isentropic_vortex(n, seed=seed)
        reqs.append(dict(q0=cases.to_conserved(p), spacing=(h, h, h), t_end=0.05, cfl=0.6,
                         recon=O.WENO, limiter=O.LIM_NONE, flux=O.AUSMPLUS, rk=O.RK_SSP3,
                         ref_compat=0, periodic=(True, True, True)))
    p = cases.smooth_3d((18, 22, 26), seed=6)
    reqs.append(dict(q0=cases.to_conserved(p), spacing=(0.05, 0.05, 0.05), t_end=0.02, cfl=0.5,
                     recon=O.MUSCL, limiter=O.LIM_VANALBADA, flux=O.ROE, rk=O.RK4))
    return reqs


def run(reqs, max_active=64, steps_per_sync=1):
    out = []
    with fvmgpu.Server(max_active=max_active) as srv:
        ids = [srv.submit(**r) for r in reqs]
        while True:
            pend, act, done = srv.poll(50, steps_per_sync)
            if done == len(ids):
                break
        for i in ids:
            out.append(srv.result(i))
    return out


def test_alone_equals_mix():
    reqs = mixed_requests()
    print("case serving mix of", len(reqs), flush=True)
    mix = run(reqs)
    for r, (qm, im) in zip(reqs, mix):
        (qa, ia), = run([r])
        assert im["status"] == 2 and ia["status"] == 2
        assert im["steps"] == ia["steps"]
        assert np.array_equal(qm, qa)
        assert abs(im["t_final"] - r["t_end"]) < 1e-14


def test_admission_by_slots_and_memory():
    """Two slots: requests wait in arrival order and are admitted as slots
    free up; results do not change."""
    reqs = mixed_requests()
    few = run(reqs, max_active=2)
    many = run(reqs, max_active=64)
    for (qa, ia), (qb, ib) in zip(few, many):
        assert np.array_equal(qa, qb)
    admits = [i["t_admit"] for _, i in few]
    assert max(admits) > min(i["t_done"] for _, i in few)   # later ones waited for a slot


def plain_solve(r):
    """The same solve through a plain context: CFL steps, the last one clipped
    to t_end with a fixed dt, time summed as the serving loop sums it."""
    q0 = r["q0"]
    _, nk, nj, ni = q0.shape
    with fvmgpu.Context(0) as ctx:
        ctx.scheme(recon=r["recon"], limiter=r["limiter"], flux=r["flux"], rk=r["rk"],
                   ref_compat=r.get("ref_compat", 1))
        b = ctx.block((ni, nj, nk), r["spacing"])
        if any(r.get("periodic", ())):
            b.set_periodic([d for d in range(3) if r["periodic"][d]])
        b.upload(q0)
        t, steps = 0.0, 0
        while t < r["t_end"] * (1.0 - 1e-14):
            # a CFL step that would pass t_end is replaced by the clipped one
            q_before = b.download()
            dt = ctx.step(1, cfl=r["cfl"])
            if t + dt > r["t_end"]:
                b.upload(q_before)
                dt = r["t_end"] - t
                ctx.step(1, dt=dt)
            t = t + dt
            steps += 1
        return b.download(), steps, t


def test_served_equals_plain_context():
    """Served results equal a plain context's (whose CFL path the oracle pins)."""
    for r in mixed_requests()[::2]:
        (q, info), = run([r])
        qp, steps, t = plain_solve(r)
        assert info["steps"] == steps
        assert info["t_final"] == t
        assert np.array_equal(q, qp)


def test_arrivals_during_work():
    """Requests submitted one by one between polls join while others run
    (batches queued ahead of the host, slots and device memory reused once
    the queued batches are done); each result equals its alone run.  Every
    other request is copied at submit instead of borrowed."""
    reqs = mixed_requests()
    print("case serving arrivals between polls", flush=True)
    alone = [run([r])[0] for r in reqs]
    with fvmgpu.Server(max_active=3) as srv:
        ids = []
        for i, r in enumerate(reqs):
            ids.append(srv.submit(**r, borrow=i % 2 == 0))
            srv.poll(1, 2)
        while srv.poll(1, 2)[2] < len(ids):
            pass
        got = [srv.result(i) for i in ids]
    for (qa, ia), (qb, ib) in zip(alone, got):
        assert ib["status"] == 2
        assert ia["steps"] == ib["steps"]
        assert np.array_equal(qa, qb)


def test_steps_per_sync():
    """Several steps per host synchronisation: solves that finish mid-run
    halt on the device; results are unchanged."""
    reqs = mixed_requests()
    one = run(reqs)
    many = run(reqs, steps_per_sync=7)
    for (qa, ia), (qb, ib) in zip(one, many):
        assert ia["steps"] == ib["steps"]
        assert np.array_equal(qa, qb)


def test_plain_end_time():
    """A plain context with an end time clips the last CFL step and skips
    the steps past t_end on the device: same result as the served solve."""
    for r in mixed_requests()[1::2]:
        (q, info), = run([r])
        q0 = r["q0"]
        _, nk, nj, ni = q0.shape
        with fvmgpu.Context(0) as ctx:
            ctx.scheme(recon=r["recon"], limiter=r["limiter"], flux=r["flux"], rk=r["rk"],
                       ref_compat=r.get("ref_compat", 1))
            b = ctx.block((ni, nj, nk), r["spacing"])
            if any(r.get("periodic", ())):
                b.set_periodic([d for d in range(3) if r["periodic"][d]])
            b.upload(q0)
            ctx.set_time(0.0, r["t_end"])
            for _ in range(1000):
                ctx.step(5, cfl=r["cfl"])
                t, steps = ctx.time()
                if t >= r["t_end"] * (1 - 1e-14):
                    break
            norms, nan = ctx.norms()
            assert steps == info["steps"] and t == info["t_final"]
            assert np.array_equal(b.download(), q)
            assert not nan and norms[5] > 0.0


def test_device_initial_state_requests():
    """Requests whose initial state is computed on the device at admission
    (no host field crosses PCIe) equal the same solves submitted with that
    state uploaded: the reference's quadrants (the host generator is bit for
    bit the device regions) and a vortex (its state taken from a plain
    context's device init, the same kernel)."""
    p, dx = cases.riemann2d(41, seed=0)
    quads = fvmgpu.region_params(cases.REF_QUADRANTS)
    n = 20
    _, h, centre = cases.isentropic_vortex(n, seed=4)
    vparams = [10.0, 5.0, centre[0], centre[1], 1.0, 1.0, 0.0]
    with fvmgpu.Context() as ctx:
        ctx.scheme()
        b = ctx.block((n, n, n), (h, h, h))
        b.init(fvmgpu.IC_VORTEX, vparams)
        qv = b.download()
    two = dict(spacing=(dx, dx, 1.0), t_end=0.01, cfl=0.5, recon=O.MUSCL, limiter=O.LIM_MINMOD,
               flux=O.ROE, rk=O.RK_SSP3)
    vort = dict(spacing=(h, h, h), t_end=0.05, cfl=0.6, recon=O.WENO, limiter=O.LIM_NONE,
                flux=O.AUSMPLUS, rk=O.RK_SSP3, ref_compat=0, periodic=(True, True, True))
    print("case serving device initial states", flush=True)
    dev = run([dict(q0=None, ic=((41, 41, 1), fvmgpu.IC_REGIONS, quads), **two),
               dict(q0=None, ic=((n, n, n), fvmgpu.IC_VORTEX, vparams), **vort)])
    host = run([dict(q0=cases.to_conserved(p), **two), dict(q0=qv, **vort)])
    for (qd, idv), (qh, ih) in zip(dev, host):
        assert idv["status"] == 2 and ih["status"] == 2
        assert np.array_equal(qd, qh)
    with fvmgpu.Server() as srv:
        with pytest.raises(fvmgpu.FvmError):
            srv.submit(None, (0.1, 0.1, 0.1), 0.1, 0.5, ic=((8, 8, 8), fvmgpu.IC_TGV, [-1.0]))
