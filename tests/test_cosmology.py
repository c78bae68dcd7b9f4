"""Expansion table and step look-ups (cube_nu_amd.cosmology) against the closed
forms of an Einstein-de Sitter universe, an ODE solver for flat LCDM and the
stopping/landing rules of the reference (Pk.py:260-283, power_nu.f90:8-71,
timestep.f90:20-89).  CAMB's H(z) is not available, so the H and tau columns
are pinned by these closed forms, not by the reference's own table."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from cube_nu_amd import cosmology as cz

H0 = 0.6817


def eds_a(t):
    """Omega_m = 1: da/dt = omhsq0 H a^3 = (2/3) a^(3/2), a(0) = 1."""
    return (1.0 - np.asarray(t) / 3.0) ** -2


def test_table_stops_below_a_min_and_pads_with_ones():
    t, a, tau, H, ie = cz.expansion_table(H0, 0.301)
    assert a[ie + 1] <= 1 / 301 < a[ie]
    assert np.all(a[ie + 2:] == 1.0)
    assert np.all(np.diff(a[: ie + 2]) < 0)
    np.testing.assert_array_equal(t, -np.arange(len(t)) * cz.DT0)
    assert np.all(H == H[0])  # every row is H(a_nu), Pk.py:280


def test_eds_table_is_first_order_euler_of_closed_form():
    errs = []
    for dt0 in (1e-3, 5e-4, 2.5e-4):
        t, a, *_ , ie = cz.expansion_table(H0, 1.0, dt0=dt0, n_a=int(20 / dt0))
        sel = slice(1, ie + 1)
        errs.append(np.max(np.abs(a[sel] / eds_a(t[sel]) - 1)))
    assert errs[1] < 2e-3
    # halving dt0 halves the error: the scheme is explicit Euler, no more, no less
    assert 1.8 < errs[0] / errs[1] < 2.2 and 1.8 < errs[1] / errs[2] < 2.2


def test_lcdm_table_matches_ode_solver():
    om = 0.301
    t, a, *_ , ie = cz.expansion_table(H0, om)
    ts = t[: ie + 1]
    sol = solve_ivp(lambda tt, y: [cz.omhsq0(H0, om) * cz.hubble(y[0], H0, om) * y[0] ** 3],
                    (0.0, ts[-1]), [1.0], t_eval=ts, rtol=1e-11, atol=1e-14)
    rel = np.abs(a[: ie + 1] / sol.y[0] - 1)
    assert rel.max() < 3e-3
    # and the error is O(dt0): a tenth of dt0 brings it down tenfold
    t2, a2, *_ , ie2 = cz.expansion_table(H0, om, dt0=cz.DT0 / 10, n_a=2_000_000)
    idx = np.arange(0, ie + 1, 100)
    rel2 = np.abs(a2[idx * 10] / sol.y[0][idx] - 1)
    assert rel2.max() < rel[idx].max() / 8


def test_conformal_time_closed_forms():
    a = np.array([1e-3, 0.01, 0.1, 0.5, 1.0])
    # EdS: tau = 2 sqrt(a) / H0
    np.testing.assert_allclose(cz.conformal_time(a, H0, 1.0), 2 * np.sqrt(a) / (100 * H0), rtol=1e-7)
    # radiation only: H = H0 a^-2, tau = a / H0
    np.testing.assert_allclose(cz.conformal_time(a, H0, 0.0, 1.0), a / (100 * H0), rtol=1e-7)


def eds_history(z_start):
    t, a, tau, H, _ = cz.expansion_table(H0, 1.0)
    return cz.ExpansionHistory(t, a, tau, H, np.float32(1 / (1 + z_start)), H0, 1.0)


def test_expansion_matches_eds_closed_form():
    h = eds_history(9.0)
    a0 = float(h.a)
    for dt in (1e-3, 0.05, 0.7):
        t0 = 3 * (1 - a0 ** -0.5)
        exact = eds_a(t0 + dt) - a0
        da = float(h.expansion(dt))
        # table Euler error (<1e-3 relative in a) bounds the look-up
        assert abs(da - exact) < 1.5e-3 * a0 + 1e-6 * abs(exact), (dt, da, exact)
    # a zero step lands one table segment back on the neighbouring secant
    # (the i2 walk of power_nu.f90:26-29), so it gives the segment's
    # curvature, O(dt0^2 a''), not zero
    assert abs(float(h.expansion(0.0))) < 1e-6 * a0


def test_cursor_walks_toward_a_equal_one():
    h = cz.ExpansionHistory.lcdm(99.0)
    ia0 = h.ia
    assert h.s2a[h.ia] >= h.a > h.s2a[h.ia + 1]
    prev = h.ia
    for _ in range(40):
        da = h.expansion(0.3)
        assert h.ia <= prev
        prev = h.ia
        # the cursor found by the downward walk is the global bracket
        assert h.s2a[h.ia] > h.a
        assert h.ia == 1 or h.s2a[h.ia + 1] <= h.a
        h.advance(da, h.dtau_a(da))
    assert h.ia < ia0 and h.a > 0.025


def test_dtau_and_sf_a_match_eds_closed_forms():
    h = eds_history(49.0)
    a_start = float(h.a)
    # stime <= 0 everywhere, so sf_a's clamp (power_nu.f90:68) makes s_fi
    # and every later sf_a zero, as in the reference
    assert h.s_fi == 0.0 and h.sf_a(a_start) == 0.0 and h.sf_a(0.9) == 0.0
    # sim%tau starts at 0 (ic.f90:126, initialize.f90:68), so the first dtau
    # is the whole conformal time to a+da and sim%tau is absolute afterwards
    for step in range(20):
        da = h.expansion(0.4)
        dtau = h.dtau_a(da)
        a0, a1 = float(h.a), float(h.a) + float(da)
        exact = 2 / (100 * H0) * (math.sqrt(a1) - (math.sqrt(a0) if step else 0.0))
        assert abs(float(dtau) - exact) < 2e-3 * exact + 2e-7, step
        h.advance(da, dtau)
    exact_tau = 2 / (100 * H0) * math.sqrt(float(h.a))
    assert abs(float(h.tau) - exact_tau) < 1e-3 * exact_tau


def test_choose_step_limits_and_landing():
    h = cz.ExpansionHistory.lcdm(49.0)
    # unconstrained: the expansion limit da/(a+da) <= 0.1 binds at late times
    h.a = np.float32(0.5)
    h.ia = 1
    while h.s2a[h.ia + 1] >= h.a:
        h.ia += 1
    dt, da, dtau, a_mid, landed = cz.choose_step(h, 9, 9, 9, 9, 9, a_next=1.0)
    assert not landed and da / (h.a + da) <= cz.RA_MAX * (1 + 1e-5)
    assert a_mid == np.float32(h.a + np.float32(da / 2))
    # kick limits bind: dt is their minimum
    dt, *_ = cz.choose_step(h, 0.03, 0.02, 0.05, 0.011, 0.04, a_next=1.0)
    assert dt == np.float32(0.011)
    # a step past a_next is shortened to land on it (|a+da|/a_next - 1 < 1e-6)
    a_next = np.float32(0.52)
    dt, da, dtau, a_mid, landed = cz.choose_step(h, 9, 9, 9, 9, 9, a_next=a_next)
    assert landed and abs(np.float32(h.a + da) / a_next - 1) < 1e-6
    assert dtau > 0
