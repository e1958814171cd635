"""Host-side case setup against the reference's initial state."""
import os

import numpy as np

import fvmgpu.cases as cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_riemann_seed0_is_reference_ic():
    bs = np.load(os.path.join(GOLDEN, "block_states.npz"))
    p, dx = cases.riemann2d(21, seed=0)
    assert dx == float(bs["riemann21_dx"])
    assert np.array_equal(p[0], bs["riemann21_p"][0])
    # p -> q with the reference's operation order (class_flow_field.f90:353-377)
    assert np.array_equal(cases.to_conserved(p)[:, 0], bs["riemann21_q0"])


def test_riemann_seeded_ranges():
    p, _ = cases.riemann2d(33, seed=7)
    rho, u, v, w, pr = (p[0, ..., m] for m in range(5))
    assert rho.min() >= 0.1 and rho.max() <= 2.0
    assert pr.min() >= 0.02 and pr.max() <= 1.5
    assert np.abs(u).max() <= 1.3 and np.abs(v).max() <= 1.3 and np.all(w == 0)
    assert len(np.unique(rho)) == 4


def test_conversions_round_trip():
    p = cases.smooth_3d((4, 5, 6), seed=1)
    back = cases.to_primitive(cases.to_conserved(p))
    np.testing.assert_allclose(back, p, rtol=1e-14, atol=1e-15)


def test_vortex_is_isentropic_and_periodic():
    p, h, c = cases.isentropic_vortex(32, seed=0)
    np.testing.assert_allclose(p[..., 4], p[..., 0] ** 1.4, rtol=1e-14)
    assert h == 10.0 / 32 and c == (5.0, 5.0)
    # far field tends to the free stream
    assert abs(p[0, 0, 0, 0] - 1.0) < 1e-3 and abs(p[0, 0, 0, 1] - 1.0) < 1e-2
