"""Known-answer tests of the oracle's AUSM+ flux (orc_flux_ausmplus), against
values derived by hand from Liou, "A sequel to AUSM: AUSM+", J. Comput. Phys.
129 (1996) 364-382: split Mach polynomials eq. (19)-(20), split pressures
eq. (21) with beta = 1/8 and alpha = 3/16, interface flux eq. (26)/(32), and
the interface speed of sound eq. (40):

    c*^2 = 2 (gamma - 1) / (gamma + 1) H          (= H / 3 at gamma = 1.4)
    c~_L = c*_L^2 / max(c*_L,  V_L)
    c~_R = c*_R^2 / max(c*_R, -V_R)
    c_1/2 = min(c~_L, c~_R),   M_L = V_L / c_1/2,   M_R = V_R / c_1/2
    |M| < 1:  M+-(M) = +-(M +- 1)^2 / 4 +- beta (M^2 - 1)^2
              P+-(M) = (M +- 1)^2 (2 -+ M) / 4 +- alpha M (M^2 - 1)^2
    |M| >= 1: M+-(M) = (M +- |M|) / 2,  P+-(M) = (1 +- sign M) / 2
    m = M+(M_L) + M-(M_R),  p_1/2 = P+(M_L) p_L + P-(M_R) p_R
    F = c_1/2 (max(m, 0) rho_L (1, u_L, H_L) + min(m, 0) rho_R (1, u_R, H_R))
        + p_1/2 (0, n, 0)

Hand values used below (exact binary fractions):
    M+(0) = 3/8        P+(0) = 1/2
    M+(1/2) = 9/16 + 9/128 = 81/128 = 0.6328125
    P+(1/2) = 27/32 + 27/512 = 459/512 = 0.896484375
    M-(1/2) = -1/16 - 9/128 = -17/128 = -0.1328125
    P-(1/2) = 5/32 - 27/512 = 53/512 = 0.103515625

The reference has no AUSM+ (its AUSM variant, src/scheme/scheme_AUSM.f90:
27-171, has beta = alpha = 0 and per-side sound speeds, and is pinned by
the reference's own compiled module).  Consistency F(q, q) = f(q) and
supersonic upwinding cancel beta, alpha and c_1/2, so each case below is
built where they change the answer.  For every case the docstring gives the
mass flux that a plausible slip would produce; the test checks all five
components against the hand values to 1e-14 and that each slip's value is
rejected.
"""
from fractions import Fraction as Fr
import math

import numpy as np
import pytest

import _oracle as O

G = 1.4


def _state(rho, V, H, d):
    """AoS primitive state (rho, u, v, w, p) with normal velocity V along e_d
    and total enthalpy H at gamma = 1.4: p = (H - V^2 / 2) rho (gamma - 1) / gamma."""
    p = (H - 0.5 * V * V) * rho * 2.0 / 7.0
    s = [rho, 0.0, 0.0, 0.0, p]
    s[1 + d] = V
    return s, p


def _expect(c12, m, rhoL, VL, HL, rhoR, VR, HR, p12, d):
    mp, mm = max(m, 0.0), min(m, 0.0)
    f = np.zeros(5)
    f[0] = c12 * (mp * rhoL + mm * rhoR)
    f[1 + d] = c12 * (mp * rhoL * VL + mm * rhoR * VR) + p12
    f[4] = c12 * (mp * rhoL * HL + mm * rhoR * HR)
    return f


def _run(L, R, d):
    return O.flux(O.oracle(), O.AUSMPLUS, np.array([L]), np.array([R]), d)[0]


@pytest.mark.parametrize("d", [0, 1, 2])
def test_subsonic_split_polynomials(d):
    """Case 1: L = (rho 2, V 1/2, H 6), R = (rho 1, V 0, H 3).
    c*_L = sqrt 2 > V_L and c*_R = 1 > -V_R, so c~ = c* on both sides and
    c_1/2 = min(sqrt 2, 1) = 1; M_L = 1/2, M_R = 0 (both subsonic).
    m = M+(1/2) + M-(0) = 81/128 - 3/8 = 33/128;
    p_1/2 = 459/512 p_L + 1/2 p_R, p_L = 47/14, p_R = 6/7.
    Mass flux 1 * 33/128 * 2 = 0.515625.  Slips:
      beta = 1/6:   m = (9/16 + 3/32) - (1/4 + 1/6) = 23/96   -> 0.479166...
      beta = 0 (the reference AUSM's value): m = 9/16 - 1/4   -> 0.625
      alpha = 0:    p_1/2 = 27/32 p_L + 1/2 p_R (momentum off by 0.17703)
      max instead of min in c_1/2: c_1/2 = sqrt 2, M_L = 0.35355 -> 0.5058...
    """
    L, pL = _state(2.0, 0.5, 6.0, d)
    R, pR = _state(1.0, 0.0, 3.0, d)
    assert pL == pytest.approx(47 / 14, rel=1e-15) and pR == pytest.approx(6 / 7, rel=1e-15)
    m = float(Fr(81, 128) - Fr(3, 8))
    p12 = 459 / 512 * pL + 0.5 * pR
    want = _expect(1.0, m, 2.0, 0.5, 6.0, 1.0, 0.0, 3.0, p12, d)
    got = _run(L, R, d)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
    assert got[0] == pytest.approx(0.515625, rel=1e-15)
    # the slips' mass and momentum fluxes, derived above, are rejected
    for slip_mass in (2.0 * 23 / 96, 2.0 * 0.3125):
        assert abs(got[0] - slip_mass) > 1e-6
    assert abs(got[1 + d] - (m * 2.0 * 0.5 + 27 / 32 * pL + 0.5 * pR)) > 1e-6
    c_max = math.sqrt(2.0)
    ML = 0.5 / c_max
    m_max = 0.25 * (ML + 1) ** 2 + 0.125 * (ML * ML - 1) ** 2 - 3 / 8
    assert abs(got[0] - c_max * m_max * 2.0) > 1e-6


@pytest.mark.parametrize("d", [0, 1, 2])
def test_interface_speed_from_velocity_branch(d):
    """Case 2: L = (rho 1, V 2, H 3), R = (rho 1, V 1/4, H 3).
    c*_L = 1 < V_L: c~_L = 1 / 2 (the velocity branch); c*_R = 1 > -V_R:
    c~_R = 1 (the c* branch); c_1/2 = 1/2.  M_L = 4 (supersonic: M+ = 4,
    P+ = 1), M_R = 1/2 (subsonic: M- = -17/128, P- = 53/512).
    m = 4 - 17/128 = 495/128; p_1/2 = p_L + 53/512 p_R, p_L = 2/7,
    p_R = 95/112.  Mass flux 1/2 * 495/128 = 1.93359375.  Slips:
      c~ = c* on both sides (no velocity branch): c_1/2 = 1, M_L = 2,
        M_R = 1/4                                            -> 1.7495...
      max instead of min in c_1/2: c_1/2 = 1, same as above
      beta = 1/6 at M_R = 1/2: m = 4 - (1/16 + 3/32)       -> 1.921875
    """
    L, pL = _state(1.0, 2.0, 3.0, d)
    R, pR = _state(1.0, 0.25, 3.0, d)
    assert pL == pytest.approx(2 / 7, rel=1e-15) and pR == pytest.approx(95 / 112, rel=1e-15)
    m = float(Fr(4) - Fr(17, 128))
    p12 = pL + 53 / 512 * pR
    want = _expect(0.5, m, 1.0, 2.0, 3.0, 1.0, 0.25, 3.0, p12, d)
    got = _run(L, R, d)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
    assert got[0] == pytest.approx(1.93359375, rel=1e-15)
    MR = 0.25
    m_nobranch = 2.0 - 0.25 * (MR - 1) ** 2 - 0.125 * (MR * MR - 1) ** 2
    for slip_mass in (1.0 * m_nobranch, 0.5 * (4.0 - (1 / 16 + 3 / 32))):
        assert abs(got[0] - slip_mass) > 1e-6


@pytest.mark.parametrize("d", [0, 1, 2])
def test_interface_speed_sign_of_right_velocity(d):
    """Case 3: L = (rho 1, V 0, H 12), R = (rho 1, V 2, H 3).
    c~_L = c*_L = 2; c*_R = 1 and -V_R = -2, so c~_R = 1 / max(1, -2) = 1;
    c_1/2 = 1.  M_L = 0 (M+ = 3/8, P+ = 1/2), M_R = 2 (M- = 0, P- = 0).
    m = 3/8; p_1/2 = p_L / 2 = 12/7 (p_L = 24/7).  Mass flux 0.375, energy
    flux 4.5.  Slips:
      |V_R| or +V_R in c~_R: c~_R = 1/2, c_1/2 = 1/2              -> 0.1875
      max instead of min: c_1/2 = 2                               -> 0.75
    """
    L, pL = _state(1.0, 0.0, 12.0, d)
    R, pR = _state(1.0, 2.0, 3.0, d)
    assert pL == pytest.approx(24 / 7, rel=1e-15) and pR == pytest.approx(2 / 7, rel=1e-15)
    want = _expect(1.0, 0.375, 1.0, 0.0, 12.0, 1.0, 2.0, 3.0, 0.5 * pL, d)
    got = _run(L, R, d)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
    assert got[0] == pytest.approx(0.375, rel=1e-15)
    assert got[4] == pytest.approx(4.5, rel=1e-15)
    for slip_mass in (0.1875, 0.75):
        assert abs(got[0] - slip_mass) > 1e-6


@pytest.mark.parametrize("d", [0, 1, 2])
def test_left_velocity_sign_and_negative_mass_flux(d):
    """Case 4, the mirror image of case 1 (flow towards the left):
    L = (rho 1, V 0, H 3), R = (rho 2, V -1/2, H 6).
    c~_L = c*_L = 1 (V_L = 0); c~_R = 2 / max(sqrt 2, 1/2) = sqrt 2; c_1/2 = 1.
    M_L = 0 (M+ = 3/8, P+ = 1/2), M_R = -1/2 (M- = -M+(1/2) = -81/128,
    P- = P+(1/2) = 459/512).  m = 3/8 - 81/128 = -33/128 < 0: the flux is
    carried by the right state; mass flux -0.515625, the negative of case 1.
    Slip: -V_L replaced by |V| on the left side would not matter here (V_L =
    0), but a swapped state (max(m, 0) paired with rho_R) gives +0.515625 or
    -0.2578125."""
    L, pL = _state(1.0, 0.0, 3.0, d)
    R, pR = _state(2.0, -0.5, 6.0, d)
    m = float(Fr(3, 8) - Fr(81, 128))
    p12 = 0.5 * pL + 459 / 512 * pR
    want = _expect(1.0, m, 1.0, 0.0, 3.0, 2.0, -0.5, 6.0, p12, d)
    got = _run(L, R, d)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)
    assert got[0] == pytest.approx(-0.515625, rel=1e-15)
    for slip_mass in (0.515625, -0.2578125):
        assert abs(got[0] - slip_mass) > 1e-6
    # mirror symmetry with case 1: mass and energy flip sign, normal momentum keeps it
    L1, _ = _state(2.0, 0.5, 6.0, d)
    R1, _ = _state(1.0, 0.0, 3.0, d)
    f1 = _run(L1, R1, d)
    np.testing.assert_allclose([got[0], got[1 + d], got[4]], [-f1[0], f1[1 + d], -f1[4]],
                               rtol=1e-14)
