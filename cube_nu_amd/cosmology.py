"""Expansion history and the step look-ups that turn a code-time step into
da and dtau, without CAMB.

The reference reads a(t) from s_a_tau_H.txt, which neutrinos/Pk.py:260-283
writes by integrating da/dt = omhsq0 * H(a) * a^3 backwards from a = 1 with
explicit Euler steps of dt0 = 5e-4 (omhsq0 = 2/3 sqrt(1/Omega_m)/H0,
parameters.f90:163), stopping once a <= 1/301; t = -i*dt0; tau is CAMB's
conformal time / c; every H entry is overwritten with H(a_nu) (Pk.py:280).
CAMB is not available here, so H(a) is the flat FLRW rate
H0 sqrt(Om a^-3 + Or a^-4 + OL) with OL = 1 - Om - Or (Pk.py:146) and tau is
integrated from it.  The table scheme, its stopping rule and the look-ups are
the reference's; the H(a) and tau columns differ from CAMB's by the massive
neutrino and exact radiation terms CAMB carries.

ExpansionHistory holds the tables as the Fortran host does (real(4) arrays of
istep_max entries, variables.f90:10, read in initialize.f90:84-89, entries past
the stopping point keep a = 1 and tau, H of the last row) and the cursor ia
(initialize.f90:93-96); expansion(), dtau_a() and sf_a() follow
neutrinos/power_nu.f90:8-71 with their 1-based indices kept.
"""
from __future__ import annotations

import math

import numpy as np

C_KMS = 299792.458
ISTEP_MAX = 100000  # parameters.f90:127
DT0 = 5e-4          # Pk.py:264


def hubble(a, h0=0.6817, omega_m=0.301, omega_r=0.0):
    """Flat FLRW expansion rate in km/s/Mpc."""
    omega_l = 1.0 - omega_m - omega_r
    a = np.asarray(a, np.float64)
    return 100.0 * h0 * np.sqrt(omega_m * a ** -3 + omega_r * a ** -4 + omega_l)


def omhsq0(h0, omega_m):
    """parameters.f90:163 / Pk.py:153."""
    return 2.0 / 3.0 * np.sqrt(1.0 / omega_m) / h0 / 100.0


def a_nu_default(mass_nu):
    """Pk.py:151 (used when the parameter file has no a_nu)."""
    return 1.0 / (595.0 / 5.47 * (mass_nu / 3) / 0.1 + 0.01)


def conformal_time(a, h0=0.6817, omega_m=0.301, omega_r=0.0):
    """tau(a) = int_0^a da' / (a'^2 H(a')) in Mpc/(km/s): the reference's
    CAMB conformal_time(z)/c (Pk.py:281).  Trapezoid in ln a' on 2e5 points
    from a_c = 1e-10, plus the exact integral over [0, a_c] where the
    constant term of H^2 is negligible:
    2 (sqrt(Om a_c + Or) - sqrt(Or)) / (H0 Om)."""
    a_c = 1e-10
    H0 = 100.0 * h0
    if omega_m > 0:
        head = 2.0 * (math.sqrt(omega_m * a_c + omega_r) - math.sqrt(omega_r)) / (H0 * omega_m)
    else:
        head = a_c / (H0 * math.sqrt(omega_r))
    la = np.linspace(np.log(a_c), 0.0, 200001)
    aa = np.exp(la)
    f = 1.0 / (aa * hubble(aa, h0, omega_m, omega_r))
    cum = head + np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(la))])
    return np.interp(np.log(np.asarray(a, np.float64)), la, cum)


def expansion_table(h0=0.6817, omega_m=0.301, omega_r=0.0, mass_nu=0.0, a_nu=None,
                    dt0=DT0, n_a=ISTEP_MAX):
    """The four rows of s_a_tau_H.txt (Pk.py:262-283), float64, length n_a:
    (t, a, tau, H).  Returns also i_end, the last Euler step taken."""
    om = omhsq0(h0, omega_m)
    omega_l = 1.0 - omega_m - omega_r
    H0 = 100.0 * h0
    t = -np.arange(n_a) * dt0
    a = np.ones(n_a)
    i_end = n_a - 2
    ai = 1.0
    for i in range(n_a - 1):
        Hai = H0 * math.sqrt(omega_m / ai ** 3 + omega_r / ai ** 4 + omega_l)
        ai = -om * Hai * ai ** 3 * dt0 + ai
        a[i + 1] = ai
        if ai <= 1.0 / 301 or math.isnan(ai):
            i_end = i
            break
    if a_nu is None:
        a_nu = a_nu_default(mass_nu)
    H = np.full(n_a, float(hubble(a_nu, h0, omega_m, omega_r)))
    tau = conformal_time(a, h0, omega_m, omega_r)
    return t, a, tau, H, i_end


class ExpansionHistory:
    """The host's view of s_a_tau_H.txt plus the cursor ia.

    `a` and `tau` are the run's sim%a and sim%tau; a step calls expansion(dt)
    for da and dtau_a(da) for dtau (timestep.f90:25-71), then advances
    a += da, tau += dtau (timestep.f90:87-89)."""

    def __init__(self, t, a_tab, tau_tab, H_tab, a_start, h0, omega_m):
        n = len(t)
        # 1-based: element 0 is unused so indices read as in the Fortran
        self.stime = np.concatenate([[np.nan], np.asarray(t, np.float32)])
        self.s2a = np.concatenate([[np.nan], np.asarray(a_tab, np.float32)])
        self.s2tau = np.concatenate([[np.nan], np.asarray(tau_tab, np.float32)])
        self.s2H = np.concatenate([[np.nan], np.asarray(H_tab, np.float32)])
        self.istep_max = n
        self.omhsq0 = omhsq0(h0, omega_m)
        self.a = np.float32(a_start)
        ia = 1
        while ia < n and self.s2a[ia + 1] >= self.a:  # initialize.f90:93-96
            ia += 1
        self.ia = ia
        self.tau = np.float32(0.0)
        self.s_fi = np.float32(0.0)
        self.s_fi = self.sf_a(a_start)  # initialize.f90:100

    @classmethod
    def lcdm(cls, z_start, h0=0.6817, omega_m=0.301, omega_r=0.0, mass_nu=0.0, n_a=ISTEP_MAX):
        t, a, tau, H, _ = expansion_table(h0, omega_m, omega_r, mass_nu, n_a=n_a)
        return cls(t, a, tau, H, np.float32(1.0 / (1 + z_start)), h0, omega_m)

    def expansion(self, dt0) -> np.float32:
        """power_nu.f90:8-36: da for code time dt0 from the current a."""
        stime, s2a = self.stime, self.s2a
        a8 = float(self.a)
        i1 = self.ia
        while s2a[i1] <= a8 and i1 > 1:
            i1 -= 1
        self.ia = i1
        tdoa = (float(stime[i1 + 1]) - float(stime[i1])) / (float(s2a[i1 + 1]) - float(s2a[i1]))
        t_x = float(stime[i1]) + tdoa * (a8 - float(s2a[i1])) + float(np.float32(dt0))
        i2 = i1
        while stime[i2 + 1] < t_x and i2 > 1:
            i2 -= 1
        adot = (float(s2a[i2 + 1]) - float(s2a[i2])) / (float(stime[i2 + 1]) - float(stime[i2]))
        a_x = float(s2a[i2]) + adot * (t_x - float(stime[i2]))
        return np.float32(a_x - a8)

    def dtau_a(self, da0) -> np.float32:
        """power_nu.f90:38-54, in real(4) as the reference declares it."""
        f = np.float32
        stime, s2a, s2tau = self.stime, self.s2a, self.s2tau
        a0 = f(self.a + f(da0))
        i1 = self.ia - 1
        while s2a[i1] <= a0 and i1 > 1:
            i1 -= 1
        i2 = i1 + 1
        tdoa = f((stime[i2] - stime[i1]) / (s2a[i2] - s2a[i1]))
        t_x = f(stime[i1] + tdoa * f(a0 - s2a[i1]))
        taudot = f((s2tau[i2] - s2tau[i1]) / (stime[i2] - stime[i1]))
        return f(s2tau[i1] + taudot * f(t_x - stime[i1]) - self.tau)

    def sf_a(self, a0) -> np.float32:
        """power_nu.f90:56-70: t(a0)*omhsq0 - s_fi clamped at zero (zero for
        every a0 <= 1, since the table's t is <= 0)."""
        f = np.float32
        stime, s2a = self.stime, self.s2a
        a0 = f(a0)
        i1 = 1
        while s2a[i1 + 1] >= a0 and i1 < self.istep_max - 1:
            i1 += 1
        i2 = i1 + 1
        tdoa = f((stime[i2] - stime[i1]) / (s2a[i2] - s2a[i1]))
        v = f(f(stime[i1] + tdoa * f(a0 - s2a[i1])) * self.omhsq0 - self.s_fi)
        return f(0.0) if v < 0 else v

    def advance(self, da, dtau) -> None:
        self.a = np.float32(self.a + np.float32(da))
        self.tau = np.float32(self.tau + np.float32(dtau))


DT_MAX = 1.0     # variables.f90:27
RA_MAX = 0.1     # parameters.f90:128
DT_REFINE = 1.0  # parameters.f90:45


def choose_step(hist: ExpansionHistory, dt_pm1, dt_pm2, dt_pm3, dt_pp, dt_vmax, a_next,
                dt_max=DT_MAX, ra_max=RA_MAX, dt_refine=DT_REFINE):
    """timestep.f90:20-71 on the master image: the largest dt allowed by the
    expansion (da/(a+da) <= ra_max, at most 10 tries) and by the kick's
    limits (cube_kick_out dt_*), shortened to land on a_next when the step
    would pass it.  Returns (dt, da, dtau, a_mid, landed); the caller then
    calls hist.advance(da, dtau)."""
    f = np.float32
    dt_e = f(dt_max)
    ntry = 0
    while True:
        ntry += 1
        da = hist.expansion(dt_e)
        ra = f(da / f(hist.a + da))
        if ra > ra_max:
            dt_e = f(dt_e * f(f(ra_max) / ra))
        else:
            break
        if ntry > 10:
            break
    dt = f(min(dt_e, f(dt_pm1), f(dt_pm2), f(dt_pm3), f(f(dt_refine) * f(dt_pp)), f(dt_vmax)))
    da = hist.expansion(dt)
    a_next = f(a_next)
    landed = bool(da >= f(a_next - hist.a))
    if landed:
        while abs(f(hist.a + da) / a_next - 1) >= 1e-6 or f(hist.a + da) > 1:
            dt = f(dt * f(f(a_next - hist.a) / da))
            da = hist.expansion(dt)
    a_mid = f(hist.a + f(da / 2))
    dtau = hist.dtau_a(da)
    return dt, da, dtau, a_mid, landed


def fixed_step_schedule(hist: ExpansionHistory, z_end, nsteps):
    """A fixed code-time step dt that takes `nsteps` steps from hist.a to
    a_end = 1/(1+z_end), and the per-step (dt, a, da, a_mid, dtau) of the
    main loop with that dt (timestep.f90:25-89 with dt fixed instead of the
    minimum of the kick's criteria).  The table's a(t) is piecewise linear, so
    dt = (t(a_end) - t(a)) / nsteps; the last step lands on a_end by the
    reference's rescaling rule (timestep.f90:57-61).  hist is advanced to
    a_end."""
    f = np.float32
    a_end = f(1.0 / (1.0 + z_end))

    def code_time(a8):
        stime, s2a = hist.stime, hist.s2a
        i1 = 1
        while s2a[i1 + 1] >= a8 and i1 < hist.istep_max - 1:
            i1 += 1
        tdoa = (float(stime[i1 + 1]) - float(stime[i1])) / (float(s2a[i1 + 1]) - float(s2a[i1]))
        return float(stime[i1]) + tdoa * (float(a8) - float(s2a[i1]))

    dt = f((code_time(float(a_end)) - code_time(float(hist.a))) / nsteps)
    steps = []
    for n in range(nsteps):
        d = f(dt)
        da = hist.expansion(d)
        if n == nsteps - 1 or da >= f(a_end - hist.a):
            while abs(f(hist.a + da) / a_end - 1) >= 1e-6:
                d = f(d * f(f(a_end - hist.a) / da))
                da = hist.expansion(d)
        a_mid = f(hist.a + f(da / 2))
        dtau = hist.dtau_a(da)
        steps.append({"dt": float(d), "a": float(hist.a), "da": float(da), "a_mid": float(a_mid),
                      "dtau": float(dtau)})
        hist.advance(da, dtau)
    return steps


# Neutrino particles (no reference counterpart: the reference treats massive
# neutrinos by linear response, neutrinos/power_nu.f90).
K_B_EV = 8.617333262e-5      # Boltzmann constant, eV/K
T_NU0 = 1.9454               # neutrino temperature today, K ((4/11)^(1/3) T_CMB, T_CMB = 2.7255 K)
C_KM_S = 299792.458
NU_EV_PER_OMEGA_H2 = 93.14   # sum m_nu / (Omega_nu h^2), eV


def f_nu(sum_m_nu_ev, h0=0.6817, omega_m=0.301):
    """Neutrino fraction of the matter density, Omega_nu / Omega_m."""
    return sum_m_nu_ev / NU_EV_PER_OMEGA_H2 / h0 ** 2 / omega_m


def nu_thermal_velocity_scale(m_nu_ev, ng, box, omega_m=0.301):
    """Code velocity of u = pc/(k T_nu) = 1 for neutrinos of mass m_nu (eV).

    Non-relativistic, the peculiar velocity is u k T_nu / (m_nu c^2) c with
    T_nu = T_nu0 (1+z).  The code velocity is dx/dtau in fine cells (box/ng
    Mpc/h each) per unit of the time tau with dtau = dt / (a^2 t_u) and
    t_u = 2 / (3 H0 sqrt(Omega_m)) (the time unit of H = 2/(3 a^1.5)
    sqrt(1 + ...) in the Zel'dovich velocities' vfactor a^2 H, ic.f90:937-945,
    and of omhsq0, parameters.f90:163), so v_pec = 150 sqrt(Omega_m)
    (box/ng) v / a km/s; the factors a (1+z) = 1 cancel and the scale is the
    same at every redshift."""
    return K_B_EV * T_NU0 / m_nu_ev * C_KM_S * ng / (150.0 * math.sqrt(omega_m) * box)
