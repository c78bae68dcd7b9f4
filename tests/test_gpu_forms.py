"""GPU parity of the two other forms of the step against the CPU oracle:
int8 compressed coordinates (izipx = 1, the reference's ZIPX int8 mode,
parameters.f90:91-92) in the FP32 form, and the FP64 form (the same kernels
in double) with int16 coordinates.  Same seeded inputs and tolerances as
test_gpu_parity.py: velocity change per phase rms relative <= 1e-5, integer
outputs (napm, drift order, counts, re-encoded coordinates, PIDs, buffer
layout) bit-exact.
"""
import numpy as np
import pytest

from cube_nu_amd import P3M
from cube_nu_amd.p3m import Geometry
from inputs import buffered_set, oparams, physical_set, rel_rms
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

NG = 64
A_MID, DT = 0.5, 0.3
FORMS = [(1, 32), (2, 64)]  # (izipx, precision)


@pytest.fixture(scope="module", params=FORMS, ids=lambda f: f"izipx{f[0]}-fp{f[1]}")
def form(request):
    izipx, prec = request.param
    geo = Geometry(NG, izipx=izipx)
    dense = [((5, 9, 2), 2000)]
    data = buffered_set(geo, seed=11, mean=6.0, dense=dense)
    P = oparams(geo)
    tfs = orc.tf_default(P)
    p = P3M(NG, izipx=izipx, precision=prec)
    yield geo, p, data, P, tfs, prec
    p.close()


@pytest.mark.parametrize("mask", [1, 2, 4, 8, 0])
def test_kick_parity(form, mask):
    geo, p, data, P, tfs, prec = form
    xp, vp, pid, rhoc = data
    v_gpu = vp.copy()
    v_orc = vp.copy()
    out_g = p.kick(xp, v_gpu, rhoc, a_mid=A_MID, dt=DT, phase_mask=mask)
    out_o, _, _ = orc.kick(P, xp, v_orc, rhoc, tfs, A_MID, DT, phase_mask=mask)
    err = rel_rms(v_gpu - vp, v_orc - vp)
    print(f"izipx {geo.izipx} fp{prec} mask {mask}: rms rel dv err {err:.3e}")
    assert err <= 1e-5
    keys = {1: ["f2max_pm1"], 2: ["f2max_pm2"], 4: ["f2max_pm3"], 8: ["f2max_pp"],
            0: ["dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "dt_vmax", "vz_max"]}[mask]
    for k in keys:
        assert abs(out_g[k] - out_o[k]) <= 1e-5 * abs(out_o[k]), k
    if mask == 0 or mask & 12:
        assert out_g["napm"] == out_o["napm"]


def test_drift_bit_exact(form):
    geo, p, data, P, tfs, prec = form
    xp, vp, pid, rhoc = data
    vz = float(np.abs(vp[:, 2]).max())
    st, xo, vo, po, ro, oo = orc.drift(P, xp, vp, pid, rhoc, 0.4, 0.6, vz)
    assert st == 0
    xg, vg, pg, rg = xp.copy(), vp.copy(), pid.copy(), rhoc.copy()
    og = p.drift(xg, vg, pg, rg, dt_old=0.4, dt=0.6, vz_max=vz)
    n = og["nplocal"]
    assert n == oo["nplocal"] and og["nlayer"] == oo["nlayer"]
    assert np.array_equal(rg, ro)
    assert np.array_equal(xg[:n], xo) and np.array_equal(vg[:n], vo) and np.array_equal(pg[:n], po)


def test_buffer_bit_exact(form):
    geo, p, data, P, tfs, prec = form
    xp, vp, pid, rhoc = physical_set(geo, seed=5, mean=5.0)
    n = xp.shape[0]
    cap = int(n * (geo.nte / geo.nt) ** 3 * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(P, xp, vp, pid, rhoc, cap)
    xg = np.zeros((cap, 3), geo.xp_dtype)
    vg = np.zeros((cap, 3), np.float32)
    pg = np.zeros(cap, np.int64)
    xg[:n], vg[:n], pg[:n] = xp, vp, pid
    rg = rhoc.copy()
    m = p.buffer(xg, vg, pg, rg, n)
    assert m == xe.shape[0] and np.array_equal(rg, re)
    assert np.array_equal(xg[:m], xe) and np.array_equal(vg[:m], ve) and np.array_equal(pg[:m], pe)


@pytest.mark.parametrize("ng,n,mean", [(192, 80, 6.0), (448, 144, 1.0)])
def test_pm2_with_80_and_144_point_tiles_matches_oracle(ng, n, mean):
    """The PM2 tile meshes are ngt = ng/4 + 2 ngb points: 80^3 at ng = 192
    (the PM3 size of the 512^3 step, two z slices per workgroup) and 144^3
    at ng = 448 (the PM3 level-2 size of the 1024^3 step), both on the
    size-specialised FFT passes, against the oracle; the 64^3 cases (meshes
    of 24 to 96 points) reach neither size."""
    geo = Geometry(ng, izipx=2)
    assert geo.ngt == n
    data = buffered_set(geo, seed=17, mean=mean)
    P = oparams(geo)
    tfs = orc.tf_default(P, levels=(2,))  # the PM2 phase reads tf2 only
    p = P3M(ng, izipx=2, precision=32)
    try:
        xp, vp, pid, rhoc = data
        v_gpu = vp.copy()
        v_orc = vp.copy()
        out_g = p.kick(xp, v_gpu, rhoc, a_mid=A_MID, dt=DT, phase_mask=2)
        out_o, _, _ = orc.kick(P, xp, v_orc, rhoc, tfs, A_MID, DT, phase_mask=2)
        err = rel_rms(v_gpu - vp, v_orc - vp)
        print(f"PM2 at {ng}^3 ({n}-point tiles): rms rel dv err {err:.3e}")
        assert err <= 1e-5
        assert abs(out_g["f2max_pm2"] - out_o["f2max_pm2"]) <= 1e-5 * abs(out_o["f2max_pm2"])
    finally:
        p.close()


def test_other_tile_decomposition_matches_oracle():
    """nnt = nns = 2 (2^3 tiles of 8^3 coarse cells, 2^3 sub-tiles each): other
    tile, sub-tile, mesh (64^3 PM2, 48^3 PM3) and PP grid (8^3) sizes than the
    default decomposition; full kick within the tolerance, drift bit-exact."""
    geo = Geometry(64, izipx=2, nnt=2, nns=2)
    data = buffered_set(geo, seed=29, mean=6.0, dense=[((3, 7, 12), 2000)])
    P = oparams(geo)
    tfs = orc.tf_default(P)
    p = P3M(64, izipx=2, precision=32, nnt=2, nns=2)
    try:
        xp, vp, pid, rhoc = data
        v_gpu = vp.copy()
        v_orc = vp.copy()
        out_g = p.kick(xp, v_gpu, rhoc, a_mid=A_MID, dt=DT)
        out_o, _, _ = orc.kick(P, xp, v_orc, rhoc, tfs, A_MID, DT)
        err = rel_rms(v_gpu - vp, v_orc - vp)
        print(f"nnt = nns = 2: rms rel dv err {err:.3e}, napm {out_g['napm']}")
        assert err <= 1e-5 and out_g["napm"] == out_o["napm"]
        for k in ("dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "dt_vmax", "vz_max"):
            assert abs(out_g[k] - out_o[k]) <= 1e-5 * abs(out_o[k]), k
        vz = float(np.abs(vp[:, 2]).max())
        st, xo, vo, po, ro, oo = orc.drift(P, xp, vp, pid, rhoc, 0.4, 0.6, vz)
        assert st == 0
        xg, vg, pg, rg = xp.copy(), vp.copy(), pid.copy(), rhoc.copy()
        og = p.drift(xg, vg, pg, rg, dt_old=0.4, dt=0.6, vz_max=vz)
        n = og["nplocal"]
        assert n == oo["nplocal"] and np.array_equal(rg, ro)
        assert np.array_equal(xg[:n], xo) and np.array_equal(vg[:n], vo) and np.array_equal(pg[:n], po)
    finally:
        p.close()
