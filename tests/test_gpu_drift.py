"""GPU drift parity at large colour counts: the within-cell order after a
drift depends on the pass pass(k) = (k - (1 - ncb)) mod nlayer of the source
cell (drift.f90:26, 46-136), nlayer = 2 ceil(dt_mid vz_max / ratio_cs) + 2.
These tests drive nlayer to 6, 8, 10, 14 and 18 with cell shifts up to
2 ncb = 8 coarse cells, for int8 and int16 coordinates, through the host
drift (cube_p3m_drift) and through the fused rank pass a resident drift takes
after a full kick, and compare xp, vp, pid, rhoc and nplocal with the oracle
bit for bit.
"""
import numpy as np
import pytest

from cube_nu_amd import P3M
from cube_nu_amd.p3m import Geometry
from inputs import buffered_set, oparams, physical_mask
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

NG = 64


def _fast_set(geo, seed, v_phys, v_buf):
    """Buffered Poisson set whose physical particles move by up to
    v_phys/ratio_cs coarse cells per unit dt_mid (uniform per axis) and
    buffer particles by up to v_buf/ratio_cs (a buffer particle may move by
    at most ncb outward and stay in the drift's range, drift.f90:64-73)."""
    xp, vp, pid, rhoc = buffered_set(geo, seed=seed, mean=6.0)
    rng = np.random.default_rng(seed + 1000)
    phys = physical_mask(geo, rhoc)
    v = rng.uniform(-1.0, 1.0, size=vp.shape).astype(np.float32)
    v[phys] *= np.float32(v_phys)
    v[~phys] *= np.float32(v_buf)
    return xp, np.ascontiguousarray(v), pid, rhoc


@pytest.mark.parametrize("izipx", [1, 2])
@pytest.mark.parametrize("vz,nlayer,v_phys", [(6.0, 6, 6.0), (10.0, 8, 10.0), (14.0, 10, 14.0),
                                              (22.0, 14, 22.0), (30.0, 18, 30.0)])
def test_host_drift_large_colour_counts(izipx, vz, nlayer, v_phys):
    geo = Geometry(NG, izipx=izipx)
    xp, vp, pid, rhoc = _fast_set(geo, 40 + int(vz), v_phys, min(v_phys, 16.0))
    st, xo, vo, po, ro, oo = orc.drift(oparams(geo), xp, vp, pid, rhoc, 1.0, 1.0, vz)
    assert st == 0 and oo["nlayer"] == nlayer
    p = P3M(NG, izipx=izipx)
    try:
        xg, vg, pg, rg = xp.copy(), vp.copy(), pid.copy(), rhoc.copy()
        og = p.drift(xg, vg, pg, rg, dt_old=1.0, dt=1.0, vz_max=vz)
    finally:
        p.close()
    n = og["nplocal"]
    shift = np.abs(np.ceil(vp.astype(np.float64) / geo.ratio_cs)).max()
    print(f"izipx {izipx}: nlayer {og['nlayer']}, largest shift {shift:.0f} coarse cells, nplocal {n}")
    assert n == oo["nplocal"] and og["nlayer"] == nlayer
    assert np.array_equal(rg, ro)
    assert np.array_equal(xg[:n], xo) and np.array_equal(vg[:n], vo) and np.array_equal(pg[:n], po)


@pytest.mark.parametrize("izipx", [1, 2])
@pytest.mark.parametrize("dt,nlayer_min", [(1.0, 6), (2.0, 8), (3.0, 10)])
def test_fused_resident_drift_large_colour_counts(izipx, dt, nlayer_min):
    """After a full resident kick and buffer_v, the drift's range check and
    destination counts ride on the rank pass with the kick's max |v| as the
    shift bound; with velocities up to 5 and dt_mid up to 3 the shifts reach
    4 coarse cells and nlayer 10."""
    geo = Geometry(NG, izipx=izipx)
    xp, vp, pid, rhoc = _fast_set(geo, 70 + int(dt), 5.0, 5.0)
    p = P3M(NG, izipx=izipx)
    try:
        p.upload(xp, vp, pid, rhoc)
        k = p.kick_resident(a_mid=0.5, dt=0.05)
        p.buffer_v_resident()
        xe, ve, pe, re = p.download()
        p.timing(True)
        p.timing_reset()
        d = p.drift_resident(dt_old=dt, dt=dt, vz_max=k["vz_max"])
        rep = p.timing_report()
        xg, vg, pg, rg = p.download()
    finally:
        p.close()
    assert rep.get("drift_shift", {"calls": 0})["calls"] == 0 and rep["drift_rank"]["calls"] == 1
    st, xo, vo, po, ro, oo = orc.drift(oparams(geo), xe, ve, pe, re, dt, dt, k["vz_max"])
    print(f"izipx {izipx}: dt {dt}: nlayer {d['nlayer']} (oracle {oo['nlayer']}), vz_max {k['vz_max']:.3f}")
    assert st == 0 and d["nlayer"] == oo["nlayer"] >= nlayer_min
    assert d["nplocal"] == oo["nplocal"]
    assert np.array_equal(rg, ro)
    assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(pg, po)
