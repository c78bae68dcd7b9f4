"""Neutrino particles: a PM-only second species in a context of its own,
coupled to the CDM context's kick (cube_p3m_kick_resident_nu), and seeded
Fermi-Dirac thermal velocities (cube_p3m_add_thermal_velocities).

The reference has no neutrino particles (it scales the CDM transfer functions
by 1 - f_nu, initialize.f90:316-334), so the checks are: the coupled PM1+PM2
kick with equal particle masses against the oracle's kick of the merged
particle set (the deposit is linear in the particles, the mesh is shared);
a zero neutrino mass leaves the CDM kick bit for bit as it is alone; the
neutrinos' drift after the coupled kick against the oracle's drift, bit for
bit; the thermal velocity sampler against the Fermi-Dirac moments
<u> = 7 pi^4 / (180 zeta(3)) and <u^2> = 15 zeta(5) / zeta(3) and its CDF."""
import numpy as np
import pytest
from scipy import integrate, special

from cube_nu_amd import P3M
from cube_nu_amd.p3m import Geometry
from inputs import buffered_set, oparams, rel_rms
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

NG = 64
A_MID, DT = 0.5, 0.3


def _merge(geo, a, b):
    """Per extended cell, the particles of a then those of b."""
    xa, va, pa, ra = a
    xb, vb, pb, rb = b
    ncell = ra.size
    ca = np.repeat(np.arange(ncell), ra.reshape(-1))
    cb = np.repeat(np.arange(ncell), rb.reshape(-1))
    key = np.concatenate([2 * ca, 2 * cb + 1])
    order = np.argsort(key, kind="stable")
    x = np.concatenate([xa, xb])[order]
    v = np.concatenate([va, vb])[order]
    p = np.concatenate([pa, pb])[order]
    src = np.concatenate([np.zeros(ca.size, np.int8), np.ones(cb.size, np.int8)])[order]
    return (np.ascontiguousarray(x), np.ascontiguousarray(v), np.ascontiguousarray(p), (ra + rb).astype(np.int32),
            src)


def test_coupled_kick_equal_masses_matches_merged_oracle():
    geo = Geometry(NG)
    A = buffered_set(geo, seed=11, mean=6.0)
    B = buffered_set(geo, seed=12, mean=3.0)
    B = (B[0], B[1], B[2] + 10 ** 7, B[3])
    pc, pn = P3M(NG), P3M(NG)
    try:
        pc.upload(*A)
        pn.upload(*B)
        oc, on = pc.kick_resident_nu(pn, a_mid=A_MID, dt=DT, nu_mass=1.0, phase_mask=3)
        vc = pc.download()[1]
        vn = pn.download()[1]
    finally:
        pc.close()
        pn.close()
    xm, vm, pm, rm, src = _merge(geo, A, B)
    P = oparams(geo)
    tfs = orc.tf_default(P, levels=(1, 2))
    v_o = vm.copy()
    oo, _, _ = orc.kick(P, xm, v_o, rm, tfs, A_MID, DT, phase_mask=3)
    dv_o = v_o - vm
    ec = rel_rms(vc - A[1], dv_o[src == 0])
    en = rel_rms(vn - B[1], dv_o[src == 1])
    print(f"coupled PM1+PM2 kick vs merged oracle: CDM {ec:.3e}, neutrinos {en:.3e}")
    assert ec <= 1e-5 and en <= 1e-5
    assert abs(oc["testrho"] + on["testrho"] - oo["testrho"]) <= 1e-6 * oo["testrho"]


def test_zero_neutrino_mass_leaves_cdm_kick_bit_exact_and_drift_matches_oracle():
    geo = Geometry(NG)
    A = buffered_set(geo, seed=21, mean=6.0)
    xb, vb, pb, rb = buffered_set(geo, seed=22, mean=4.0)
    vb = vb * np.float32(3.0)  # faster than the CDM, as thermal neutrinos are
    pc, pn, pr = P3M(NG), P3M(NG), P3M(NG)
    try:
        pc.upload(*A)
        pr.upload(*A)
        pn.upload(xb, vb, pb + 10 ** 7, rb)
        oc, on = pc.kick_resident_nu(pn, a_mid=A_MID, dt=DT, nu_mass=0.0)
        orf = pr.kick_resident(a_mid=A_MID, dt=DT)
        sc, sr = pc.download(), pr.download()
        assert np.array_equal(sc[1], sr[1])
        for k in ("dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "vz_max", "testrho"):
            assert oc[k] == orf[k], k
        # the neutrinos' drift after the coupled kick and their buffer_v: the
        # fused rank pass takes the kick's neutrino velocity bound
        pn.buffer_v_resident()
        xe, ve, pe, re = pn.download()
        assert abs(on["vz_max"] - np.abs(ve[:, 2]).max()) <= 1e-6 * on["vz_max"]
        d = pn.drift_resident(dt_old=DT, dt=DT, vz_max=on["vz_max"])
        xg, vg, pg, rg = pn.download()
    finally:
        for p in (pc, pn, pr):
            p.close()
    st, xo, vo, po, ro, oo = orc.drift(oparams(geo), xe, ve, pe, re, DT, DT, on["vz_max"])
    assert st == 0 and d["nplocal"] == oo["nplocal"] and d["nlayer"] == oo["nlayer"]
    assert np.array_equal(rg, ro)
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(pg, po)


def test_thermal_velocities_follow_fermi_dirac():
    p = P3M(NG, box=600.0 * NG / 512)
    try:
        p.make_ic(seed=5, z=0.5, nic=2)
        x0, v0, pid0, r0 = p.download()
        vs = 2.0
        vz = p.add_thermal_velocities(vs, seed=99)
        x1, v1, pid1, r1 = p.download()
        xb = v1.copy()
        orc.buffer_v(oparams(p.geo), xb, r1)
    finally:
        p.close()
    assert np.array_equal(pid0, pid1) and np.array_equal(x0, x1)
    assert np.array_equal(xb, v1)  # buffer copies drew their originals' velocities
    assert vz == np.abs(v1[:, 2]).max()
    dv = (v1.astype(np.float64) - v0) / vs
    phys = np.unique(pid1, return_index=True)[1]  # one copy per particle
    dv = dv[phys]
    u = np.sqrt((dv * dv).sum(axis=1))
    n = u.size
    z3, z5 = special.zeta(3), special.zeta(5)
    mean_u, mean_u2 = 7 * np.pi ** 4 / (180 * z3), 15 * z5 / z3
    sd = np.sqrt(mean_u2 - mean_u ** 2)
    print(f"{n} particles: <u> {u.mean():.4f} (FD {mean_u:.4f}), <u^2> {np.mean(u * u):.4f} (FD {mean_u2:.4f})")
    assert abs(u.mean() - mean_u) <= 6 * sd / np.sqrt(n)
    assert abs(np.mean(u * u) - mean_u2) <= 0.02 * mean_u2
    assert np.all(np.abs(dv.mean(axis=0)) <= 6 * np.sqrt(mean_u2 / 3 / n))  # isotropic
    # Kolmogorov-Smirnov distance to the Fermi-Dirac CDF
    norm = 1.5 * z3
    grid = np.linspace(0.0, 20.0, 201)
    cdf = np.array([integrate.quad(lambda t: t * t / (np.exp(t) + 1), 0, g)[0] for g in grid]) / norm
    emp = np.searchsorted(np.sort(u), grid) / n
    ks = np.abs(emp - cdf).max()
    print(f"KS distance {ks:.4f}")
    assert ks <= 2.0 / np.sqrt(n)
