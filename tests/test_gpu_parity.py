"""GPU parity: the HIP path through the C-ABI against the CPU oracle on
identical seeded inputs (ng=64: 262k particles, 64 tiles of 4^3 coarse cells,
4096 sub-tiles).

Tolerances: rms relative error of the velocity change (i.e. of
the force) <= 1e-5 per phase; integer outputs (napm, drift order, counts,
re-encoded coordinates, PID permutation) bit-exact.
"""
import numpy as np
import pytest

from cube_nu_amd import P3M
from cube_nu_amd.p3m import CubeError, CUBE_ERR_OUT_OF_RANGE
from inputs import buffered_set, oparams, physical_mask, physical_set, rel_rms
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

NG = 64
A_MID, DT = 0.5, 0.3


@pytest.fixture(scope="module")
def gpu():
    p = P3M(NG, izipx=2, precision=32)
    yield p
    p.close()


@pytest.fixture(scope="module")
def data():
    # one dense coarse cell (2000 particles) puts its sub-tile at ires=3
    dense = [((5, 9, 2), 2000)]
    return buffered_set(_geo(), seed=11, mean=6.0, dense=dense)


def _geo():
    from cube_nu_amd.p3m import Geometry

    return Geometry(NG)


@pytest.fixture(scope="module")
def tfs():
    return orc.tf_default(oparams(_geo()))


def test_green_matches_oracle(gpu, tfs):
    for level in range(1, 7):
        g = gpu.get_tf(level)
        o = tfs[level - 1]
        scale = np.abs(o).max()
        assert np.max(np.abs(g - o)) <= 2e-6 * scale, level


@pytest.mark.parametrize("mask", [1, 2, 4, 8])
def test_phase_parity(gpu, data, tfs, mask):
    xp, vp, pid, rhoc = data
    v_gpu = vp.copy()
    v_orc = vp.copy()
    out_g = gpu.kick(xp, v_gpu, rhoc, a_mid=A_MID, dt=DT, phase_mask=mask)
    out_o, _, _ = orc.kick(oparams(_geo()), xp, v_orc, rhoc, tfs, A_MID, DT, phase_mask=mask)
    dv_g = v_gpu - vp
    dv_o = v_orc - vp
    err = rel_rms(dv_g, dv_o)
    print(f"mask {mask}: rms rel dv err {err:.3e}")
    assert err <= 1e-5
    key = {1: "f2max_pm1", 2: "f2max_pm2", 4: "f2max_pm3", 8: "f2max_pp"}[mask]
    assert abs(out_g[key] - out_o[key]) <= 1e-5 * abs(out_o[key])
    if mask & 12:
        assert out_g["napm"] == out_o["napm"]


@pytest.fixture(scope="module")
def data_fine():
    # 8000 and 21000 particles in one coarse cell: npc_max 125 and 328 put
    # their sub-tiles at ires = 4 (ratio_sf 6: 24 mesh cells per coarse cell,
    # the divided, non-power-of-two mesh scale) and ires = 5 (ratio_sf 8)
    dense = [((5, 9, 2), 8000), ((10, 3, 11), 21000)]
    return buffered_set(_geo(), seed=23, mean=6.0, dense=dense)


@pytest.mark.parametrize("mask", [4, 8, 0])
def test_fine_pm3_levels_parity(gpu, data_fine, tfs, mask):
    """PM3 at ires = 4 and 5, PP in their sub-tiles, and the full kick."""
    xp, vp, pid, rhoc = data_fine
    v_gpu = vp.copy()
    v_orc = vp.copy()
    out_g = gpu.kick(xp, v_gpu, rhoc, a_mid=A_MID, dt=DT, phase_mask=mask)
    out_o, _, _ = orc.kick(oparams(_geo()), xp, v_orc, rhoc, tfs, A_MID, DT, phase_mask=mask)
    assert out_o["napm"][3] == 1 and out_o["napm"][4] == 1
    assert out_g["napm"] == out_o["napm"]
    err = rel_rms(v_gpu - vp, v_orc - vp)
    print(f"mask {mask}: rms rel dv err {err:.3e}")
    assert err <= 1e-5
    for k in {4: ["f2max_pm3"], 8: ["f2max_pp"], 0: ["dt_pm3", "dt_pp", "dt_vmax", "vz_max"]}[mask]:
        assert abs(out_g[k] - out_o[k]) <= 1e-5 * abs(out_o[k]), k


def test_pp_pair_set_is_the_references(gpu, data, tfs):
    """The PP kernel tests only the x window and the y/z columns that can hold
    partners; the number of interacting pairs (physical-physical counted once,
    physical-ghost once, ghost-ghost never, kick.f90:516-566) must equal the
    oracle's full 27-cell count exactly."""
    xp, vp, pid, rhoc = data
    gpu.kick(xp, vp.copy(), rhoc, a_mid=A_MID, dt=DT, phase_mask=8)
    n_gpu = gpu.timing_report()["pp_pairs_last"]["calls"]
    _, _, n_orc = orc.kick(oparams(_geo()), xp, vp.copy(), rhoc, tfs, A_MID, DT, phase_mask=8)
    print(f"pp pairs gpu {n_gpu} oracle {n_orc}")
    assert n_orc > 0 and n_gpu == n_orc


def test_autotune_picks_keep_results(gpu, data, tfs):
    """The first PP call of a level times the kernel's load depth (6, 4 or 8
    candidates in flight) and the first PM3 list deposit its workgroup size
    (512, 256 or 1024 threads); every trial's result is checksummed and the
    tuner stops with an error if any candidate's differs, so a pick changes
    only the time.  The report lists each choice's candidate times and pick,
    and the kick still matches the oracle."""
    xp, vp, pid, rhoc = data
    v1 = vp.copy()
    gpu.kick(xp, v1, rhoc, a_mid=A_MID, dt=DT)
    rep = gpu.timing_report()
    picks = {}
    for k, v in rep.items():
        if k.startswith("autotune:"):
            _, key, cand = k.split(":")
            picks.setdefault(key, {})[int(cand)] = (v["calls"], v["ms"])
    print("autotuned", picks)
    assert any(k.startswith("pp_depth") for k in picks)
    for key, c in picks.items():
        assert sum(flag for flag, _ in c.values()) == 1, key
        assert all(ms > 0 for _, ms in c.values()), key
    v2 = vp.copy()
    orc.kick(oparams(_geo()), xp, v2, rhoc, tfs, A_MID, DT)
    assert rel_rms(v1 - vp, v2 - vp) <= 1e-5


def test_pm1_density_and_testrho(gpu, data, tfs):
    xp, vp, pid, rhoc = data
    v = vp.copy()
    dens, out = gpu.particle_mesh(xp, v, rhoc, a_mid=A_MID, dt=DT)
    v2 = vp.copy()
    out_o, rho1, _ = orc.kick(oparams(_geo()), xp, v2, rhoc, tfs, A_MID, DT, phase_mask=1, want_rho1=True)
    assert rel_rms(dens, rho1) <= 1e-6
    assert abs(out["testrho"] - out_o["testrho"]) <= 1e-6 * out_o["testrho"]


def test_full_kick(gpu, data, tfs):
    xp, vp, pid, rhoc = data
    v_gpu = vp.copy()
    v_orc = vp.copy()
    out_g = gpu.kick(xp, v_gpu, rhoc, a_mid=A_MID, dt=DT)
    out_o, _, _ = orc.kick(oparams(_geo()), xp, v_orc, rhoc, tfs, A_MID, DT)
    assert rel_rms(v_gpu - vp, v_orc - vp) <= 1e-5
    for k in ("dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "dt_vmax", "vz_max"):
        assert abs(out_g[k] - out_o[k]) <= 1e-5 * abs(out_o[k]), k


def test_drift_bit_exact(gpu, data):
    xp, vp, pid, rhoc = data
    vz = float(np.abs(vp[:, 2]).max())
    st, xo, vo, po, ro, oo = orc.drift(oparams(_geo()), xp, vp, pid, rhoc, 0.4, 0.6, vz)
    assert st == 0
    xg, vg, pg, rg = xp.copy(), vp.copy(), pid.copy(), rhoc.copy()
    og = gpu.drift(xg, vg, pg, rg, dt_old=0.4, dt=0.6, vz_max=vz)
    n = og["nplocal"]
    assert n == oo["nplocal"] and og["nlayer"] == oo["nlayer"]
    assert og["overhead_tile"] == oo["overhead_tile"]
    assert np.array_equal(rg, ro)
    assert np.array_equal(xg[:n], xo)
    assert np.array_equal(vg[:n], vo)
    assert np.array_equal(pg[:n], po)
    assert not xg[n:].any() and not pg[n:].any()


def test_buffer_bit_exact(gpu):
    geo = _geo()
    xp, vp, pid, rhoc = physical_set(geo, seed=5, mean=5.0)
    n = xp.shape[0]
    cap = int(n * (geo.nte / geo.nt) ** 3 * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(oparams(geo), xp, vp, pid, rhoc, cap)
    xg = np.zeros((cap, 3), np.int16)
    vg = np.zeros((cap, 3), np.float32)
    pg = np.zeros(cap, np.int64)
    xg[:n], vg[:n], pg[:n] = xp, vp, pid
    rg = rhoc.copy()
    m = gpu.buffer(xg, vg, pg, rg, n)
    assert m == xe.shape[0]
    assert np.array_equal(rg, re)
    assert np.array_equal(xg[:m], xe) and np.array_equal(vg[:m], ve) and np.array_equal(pg[:m], pe)


def test_buffer_v_bit_exact(gpu):
    geo = _geo()
    xp, vp, pid, rhoc = physical_set(geo, seed=6, mean=5.0)
    cap = int(xp.shape[0] * (geo.nte / geo.nt) ** 3 * 1.2) + 1024
    v_new = (vp * np.float32(1.5) + np.float32(0.25)).astype(np.float32)
    xe, ve, pe, re = orc.buffer(oparams(geo), xp, v_new, pid, rhoc, cap)
    phys = physical_mask(geo, re)
    v = ve.copy()
    p = pe.copy()
    v[~phys] = -7.0
    p[~phys] = 0
    gpu.buffer_v(v, re, p)
    assert np.array_equal(v, ve) and np.array_equal(p, pe)
    v = ve.copy()
    v[~phys] = 2.0
    gpu.buffer_v(v, re)
    assert np.array_equal(v, ve)


def test_resident_steps_conserve_particles():
    """Several resident steps (drift, buffer, kick, buffer_v) of a clustered
    Zel'dovich start keep every particle exactly once: nplocal and the PID
    set are preserved, and the buffered layout's buffer cells carry their
    physical cells' velocities after every step."""
    p = P3M(NG, izipx=2, precision=32, box=600.0 * NG / 512)
    try:
        ic = p.make_ic(seed=7, z=1.0, nic=2)
        a = 0.5
        k0 = p.kick_resident(a_mid=a, dt=0.0)
        dt = min(k0["dt_pm1"], k0["dt_pm2"], k0["dt_pm3"], k0["dt_pp"], 1.0) * 2
        vz = ic["vz_max"]
        for _ in range(4):
            d, k = p.step_resident(dt_old=dt, dt=dt, vz_max=vz, a_mid=a)
            vz = k["vz_max"]
            assert d["nplocal"] == NG ** 3
        xe, ve, pe, re = p.download()
        v = ve.copy()
        orc.buffer_v(oparams(_geo()), v, re)
        assert np.array_equal(v, ve)
        p.drift_resident(dt_old=dt, dt=dt, vz_max=vz)
        xp, vp, pid, rhoc = p.download()
        assert xp.shape[0] == NG ** 3
        assert np.array_equal(np.sort(pid), np.arange(1, NG ** 3 + 1))
    finally:
        p.close()


def test_resident_drift_after_kick_bit_exact():
    """A resident drift right after a full kick and buffer_v takes the fused
    count + rank pass (its cell-shift bound comes from the kick's max |v|)
    and gives the oracle's drift bit for bit."""
    geo = _geo()
    p = P3M(NG, izipx=2, precision=32)
    try:
        xp, vp, pid, rhoc = buffered_set(geo, seed=13, mean=6.0)
        p.upload(xp, vp, pid, rhoc)
        k = p.kick_resident(a_mid=A_MID, dt=DT)
        p.buffer_v_resident()
        xe, ve, pe, re = p.download()
        p.timing(True)
        p.timing_reset()
        d = p.drift_resident(dt_old=0.4, dt=0.6, vz_max=k["vz_max"])
        rep = p.timing_report()
        assert rep.get("drift_shift", {"calls": 0})["calls"] == 0 and rep["drift_rank"]["calls"] == 1
        xg, vg, pg, rg = p.download()
        st, xo, vo, po, ro, oo = orc.drift(oparams(geo), xe, ve, pe, re, 0.4, 0.6, k["vz_max"])
        assert st == 0 and d["nplocal"] == oo["nplocal"] and d["nlayer"] == oo["nlayer"]
        assert np.array_equal(rg, ro)
        assert np.array_equal(xg, xo) and np.array_equal(vg, vo) and np.array_equal(pg, po)
    finally:
        p.close()


def test_drift_tile_overflow_error(data):
    """np_tile_max (tile_buffer) smaller than an extended tile's particle count
    stops the drift with the tile overflow status (drift.f90:87-93) before any
    particle moves."""
    from cube_nu_amd.p3m import CUBE_ERR_TILE_OVERFLOW

    xp, vp, pid, rhoc = data
    p = P3M(NG, izipx=2, precision=32, tile_buffer=0.05)
    try:
        with pytest.raises(CubeError) as e:
            p.drift(xp.copy(), vp.copy(), pid.copy(), rhoc.copy(), dt_old=0.4, dt=0.6, vz_max=1.0)
        assert e.value.code == CUBE_ERR_TILE_OVERFLOW
    finally:
        p.close()


def test_drift_out_of_range_error(gpu, data):
    xp, vp, pid, rhoc = data
    v = vp.copy()
    v[5, 0] = 1e6  # moves far beyond 2*ncb coarse cells (drift.f90:64-73)
    with pytest.raises(CubeError) as e:
        gpu.drift(xp.copy(), v, pid.copy(), rhoc.copy(), dt_old=0.5, dt=0.5, vz_max=1.0)
    assert e.value.code == CUBE_ERR_OUT_OF_RANGE


@pytest.mark.parametrize("mask", [1, 2, 4])
def test_fft_plane_path_matches_generic(data, mask, monkeypatch):
    """The size-specialised 3-pass FFT convolution agrees with the generic
    five-pass line kernels (both FP32) on the same deposit."""
    xp, vp, pid, rhoc = data
    gpu = P3M(NG)
    v1 = vp.copy()
    gpu.kick(xp, v1, rhoc, a_mid=A_MID, dt=DT, phase_mask=mask)
    monkeypatch.setenv("CUBE_FFT_GENERIC", "1")
    v2 = vp.copy()
    gpu.kick(xp, v2, rhoc, a_mid=A_MID, dt=DT, phase_mask=mask)
    gpu.close()
    assert rel_rms(v1 - vp, v2 - vp) <= 2e-6


@pytest.mark.parametrize("nw", [32, 128])
def test_cicpower_matches_oracle(gpu, nw):
    """CIC P(k) (cicpower.f90 + powerspectrum.f90 auto_power): counts exact,
    P(k) within 1e-4 up to half-Nyquist, the density contrast to FP32
    rounding."""
    geo = _geo()
    xp, vp, pid, rhoc = physical_set(geo, seed=23, mean=12.0)
    cap = int(xp.shape[0] * (geo.nte / geo.nt) ** 3 * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(oparams(geo), xp, vp, pid, rhoc, cap)
    gpu.upload(np.ascontiguousarray(xe), np.ascontiguousarray(ve), np.ascontiguousarray(pe), re)
    xi_g, mean_g, d_g = gpu.cicpower(nw, want_delta=True)
    xi_o, mean_o, d_o = orc.cicpower(oparams(geo), xp, rhoc, nw, gpu.cfg.box, want_delta=True)
    assert abs(mean_g - mean_o) <= 1e-7 * mean_o
    assert rel_rms(d_g, d_o) <= 1e-6
    assert np.array_equal(xi_g[:, 0], xi_o[:, 0])
    b = slice(1, nw // 4 + 1)  # up to half-Nyquist
    assert np.allclose(xi_g[b, 1], xi_o[b, 1], rtol=1e-6)
    for c in (2, 3, 4):
        err = np.abs(xi_g[b, c] - xi_o[b, c]) / np.abs(xi_o[b, 2])
        print(f"nw {nw} column {c + 1}: max rel err {err.max():.2e}")
        assert err.max() <= 1e-4, c
