"""Stage timing with HIP events (fvm_timing_enable / fvm_timing_read): the
counts of timed stage launches and cell updates, and the exclusion of steps
skipped past the end time.  Timing never changes results."""
import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

MUSCL_ROE = dict(recon=O.MUSCL, limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)


def _run(p, dx, nsteps, every, t_end=None):
    with fvmgpu.Context() as ctx:
        ctx.scheme(**MUSCL_ROE)
        n = p.shape[1]
        blk = ctx.block((n, n, 1), (dx, dx, 1.0))
        blk.upload(cases.to_conserved(p))
        if t_end is not None:
            ctx.set_time(0.0, t_end)
        if every:
            ctx.timing(every)
        ctx.step(nsteps, cfl=0.5)
        out = ctx.timing_read() if every else None
        _, steps = ctx.time()
        return blk.download(), out, steps


def test_timed_launch_counts():
    p, dx = cases.riemann2d(48, seed=0)
    print("case riemann2d 48^2 seed 0, 10 steps, timing every 3rd", flush=True)
    q_plain, _, _ = _run(p, dx, 10, 0)
    q, (ms, launches, cu), steps = _run(p, dx, 10, 3)
    assert steps == 10
    # steps 0, 3, 6, 9 are timed, three stage launches each
    assert launches == 4 * 3
    assert cu == 48 * 48 * launches
    assert ms > 0.0
    assert np.array_equal(q, q_plain)


def test_skipped_steps_not_counted():
    p, dx = cases.riemann2d(48, seed=0)
    _, _, ran = _run(p, dx, 40, 0, t_end=2e-3)
    assert 0 < ran < 40
    print(f"case riemann2d 48^2 seed 0, t_end 2e-3: {ran} of 40 steps run", flush=True)
    _, (ms, launches, cu), steps = _run(p, dx, 40, 1, t_end=2e-3)
    assert steps == ran
    assert launches == 3 * ran
    assert cu == 48 * 48 * launches
