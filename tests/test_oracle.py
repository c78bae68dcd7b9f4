"""Pins the CPU oracle (oracle/cube_oracle.cpp) without the reference being
runnable here: against golden fixtures computed by independent numpy
statements (tests/golden/make_golden.py), against numpy.fft (pocketfft, an
independent FFT library), against closed forms (TSC B-spline moments, the
FD stencil's transfer function on plane waves, the S2 force-law continuity,
the softened point-mass force), against brute force (O(N^2) PP sums) and
against invariants (momentum conservation of PM, count/PID conservation and
re-decoded positions of drift, periodic images of the buffer).

No GPU is involved.
"""
import os

import numpy as np
import pytest

from cube_nu_amd.p3m import Geometry
from inputs import buffered_explicit, buffered_set, oparams, physical_mask, physical_set, rel_rms
from oracle import pyoracle as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold(name):
    return np.load(os.path.join(GOLD, name))


# --------------------------------------------------------------------- decode
def test_decode_every_value_matches_golden():
    g = _gold("decode.npz")
    assert np.array_equal(orc.decode(np.repeat(g["xp8"], 3).reshape(-1, 3), 1)[:, 0], g["frac8"])
    assert np.array_equal(orc.decode(np.repeat(g["xp16"], 3).reshape(-1, 3), 2)[:, 0], g["frac16"])
    # cell fractions are the 2^B bin centres, strictly inside (0, 1)
    assert g["frac16"].min() == 0.5 / 65536 and g["frac16"].max() == 1 - 0.5 / 65536


# ------------------------------------------------------------------------ TSC
def test_tsc_weights_are_the_quadratic_bspline():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-40, 40, 4000), [0.0, 1.0, -1.0, 0.5, 2.999999999, -1e-12, 7.25]])
    idx, w = orc.tsc(x)
    assert np.array_equal(idx, np.floor(x).astype(np.int64))
    l = (x - np.floor(x)).astype(np.float32).astype(np.float64)
    w = w.astype(np.float64)
    assert (w >= 0).all()
    assert np.abs(w.sum(1) - 1).max() <= 2.5e-7  # partition of unity
    # mean offset of the taps {-1, 0, +1} about floor(x)+1 is l - 1/2
    assert np.abs((w[:, 2] - w[:, 0]) - (l - 0.5)).max() <= 3e-7
    # second moment of a quadratic B-spline: ((1-l)^2 + l^2) / 2
    assert np.abs((w[:, 2] + w[:, 0]) - ((1 - l) ** 2 + l ** 2) / 2).max() <= 3e-7


# ---------------------------------------------------------------------- Green
@pytest.mark.parametrize("case", ["coarse16", "tile48", "subtile24"])
def test_green_matches_independent_numpy(case):
    g = _gold("green.npz")
    n, soft, rng_, gs = g[case + "_args"]
    G = orc.green(int(n), float(soft), float(rng_), float(gs))
    ref = g[case]
    assert np.max(np.abs(G - ref)) <= 2e-6 * np.abs(ref).max()


def test_green_tends_to_coulomb_at_small_k():
    """G -> 4 pi / (k^2 gridsize^2) as k -> 0 (Poisson kernel of the force),
    with an O(k^2) deviation, isotropic along the three axes."""
    dev = {}
    for n in (32, 64):
        G = orc.green(n, 1.0, 0.0, 1.0)
        k = 2 * np.pi / n
        r = np.array([G[0, 0, 1], G[0, 1, 0], G[1, 0, 0]]) * k * k / (4 * np.pi)
        assert np.ptp(r) <= 1e-6
        dev[n] = r[0] - 1
    assert 0 < dev[64] < 3e-3 and 3.5 < dev[32] / dev[64] < 4.5
    G4 = orc.green(64, 1.0, 0.0, 4.0)
    assert abs(G4[0, 0, 1] * 16 / orc.green(64, 1.0, 0.0, 1.0)[0, 0, 1] - 1) < 1e-6


# ---------------------------------------------------------------------- F_ra
def test_f_ra_matches_closed_form_and_is_continuous():
    g = _gold("f_ra.npz")
    r = g["r"]
    assert np.allclose(orc.f_ra(r, 0.06), g["app"], rtol=3e-6, atol=0)
    a = orc.f_ra(r, 1.75).astype(np.float64)
    assert np.max(np.abs(a - g["range175"]) / np.maximum(np.abs(g["range175"]), 1e-3)) <= 2e-4
    for apm in (0.06, 1.75, 0.875):
        for ep in (1.0, 2.0):
            r0 = ep * apm / 2
            lo, hi = orc.f_ra(np.float32([r0 * (1 - 1e-5), r0 * (1 + 1e-5)]), apm)
            assert abs(lo - hi) <= 2e-3 * abs(hi), (apm, ep)
    # beyond the softening the law is Newtonian
    assert orc.f_ra(np.float32([3.0]), 1.75)[0] == np.float32(1.0) / np.float32(9.0)


# ----------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [6, 16, 24, 30, 40, 48, 80, 96, 160, 240, 320])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(orc.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    assert np.allclose(orc.fft1d(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)
    if n == 6:  # the DFT definition itself
        j = np.arange(n)
        dft = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x
        assert np.allclose(orc.fft1d(x, -1), dft, atol=1e-12)


@pytest.mark.parametrize("n", [12, 16, 20])
def test_r2c_c2r_match_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n)).astype(np.float32)
    c = orc.r2c3d(x)
    ref = np.fft.rfftn(x.astype(np.float64))
    assert np.allclose(c, ref, rtol=0, atol=1e-9 * n ** 3)
    back = orc.c2r3d(c, n)
    assert np.allclose(back, x.astype(np.float64) * n ** 3, rtol=0, atol=1e-8 * n ** 3)


def test_poisson_matches_numpy():
    n = 16
    rng = np.random.default_rng(7)
    rho = rng.standard_normal((n, n, n)).astype(np.float32)
    tf = orc.green(n, 3.5, 0.0, 4.0)
    phi = orc.poisson(rho, tf)
    ref = np.fft.irfftn(np.fft.rfftn(rho.astype(np.float64)) * tf, s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    assert rel_rms(phi, ref) <= 1e-6


# ---------------------------------------------------------------- FD gradient
@pytest.mark.parametrize("m", [1, 3, 5])
def test_fd_gradient_plane_wave(m):
    """sum_m w_m phi(x+m) on phi = A sin(kx) is A cos(kx) (a sin k + (1-a) sin 2k / 2)."""
    n, A = 16, 2.5
    k = 2 * np.pi * m / n
    x = np.arange(n)
    for axis in range(3):
        shape = [1, 1, 1]
        shape[2 - axis] = n  # arrays are [z][y][x]
        phi = np.broadcast_to((A * np.sin(k * x)).reshape(shape), (n, n, n)).astype(np.float32)
        f, f2max = orc.fd_gradient(phi)
        alpha = 4.0 / 3.0
        D = alpha * np.sin(k) + (1 - alpha) * np.sin(2 * k) / 2
        exp = np.broadcast_to((A * D * np.cos(k * x)).reshape(shape), (n, n, n))
        assert np.abs(f[..., axis] - exp).max() <= 2e-6 * A
        for other in range(3):
            if other != axis:
                assert np.abs(f[..., other]).max() <= 1e-6 * A
        assert abs(f2max - (A * D) ** 2) <= 1e-5 * (A * D) ** 2


# ------------------------------------------------------------ classification
def test_classify_thresholds_and_subtile_order():
    geo = Geometry(64)  # ntt = 1: each sub-tile is one coarse cell
    P = oparams(geo)
    rhoc = np.zeros(geo.rhoc_shape, np.int32)
    b = geo.ncb
    vals = [1535, 1536, 7295, 7296, 20479, 20480]  # npc_max = max/64 around 24, 114, 320
    cells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (3, 2, 1), (0, 0, 3), (2, 3, 3)]
    T = 5  # tile (itx=2, ity=2, itz=1)
    for v, (i, j, k) in zip(vals, cells):
        rhoc.reshape(-1, geo.nte, geo.nte, geo.nte)[T, k + b, j + b, i + b] = v
    ires, napm = orc.classify(P, rhoc)
    ns = geo.nns
    exp = np.full(ires.shape, 2)
    for (i, j, k), r in zip(cells, [2, 3, 3, 4, 4, 5]):
        exp[T * ns ** 3 + (k * ns + j) * ns + i] = r  # ifx fastest, then ify, ifz, tiles
    assert np.array_equal(ires, exp)
    assert list(napm) == [0, int((exp == 2).sum()), 2, 2, 1, 0, 0]


# ---------------------------------------------------------------------- spine
def test_spine_matches_cumulative_sums():
    geo = Geometry(64)
    P = oparams(geo)
    _, _, _, rhoc = buffered_set(geo, seed=4, mean=3.0)
    s = orc.spine_image(P, rhoc)
    nt, b = geo.nt, geo.ncb
    r = rhoc.reshape(geo.nnt ** 3, geo.nte, geo.nte, geo.nte).astype(np.int64)
    rows = r.sum(axis=3).reshape(-1)  # [T][k][j] row sums in memory order
    incl = np.cumsum(rows).reshape(geo.rhoc_shape[:5])
    assert np.array_equal(s["idx_b_r"], incl)
    assert np.array_equal(s["idx_b_l"], incl - rows.reshape(incl.shape))
    phys = r[:, b:b + nt, b:b + nt, b:b + nt]
    prow = phys.sum(axis=3).reshape(-1)
    pin = np.cumsum(prow).reshape(s["ppr"].shape)
    assert np.array_equal(s["ppr"], pin)
    assert np.array_equal(s["ppl"], pin - prow.reshape(pin.shape))
    inner = (slice(None), slice(None), slice(None), slice(b, b + nt), slice(b, b + nt))
    rr = rhoc.astype(np.int64)[..., b:b + nt, b:b + nt, :]
    assert np.array_equal(s["ppr0"], s["idx_b_r"][inner] - rr[..., b + nt:].sum(-1))
    assert np.array_equal(s["pprl"], s["ppr0"] - rr[..., nt:b + nt].sum(-1))
    assert np.array_equal(s["ppl0"], s["idx_b_l"][inner] + rr[..., :b].sum(-1))
    assert np.array_equal(s["pplr"], s["ppl0"] + rr[..., b:2 * b].sum(-1))


# --------------------------------------------------------------------- buffer
def test_buffer_is_the_periodic_image():
    geo = Geometry(64)
    xp, vp, pid, rhoc = physical_set(geo, seed=9, mean=2.0)
    P = oparams(geo)
    cap = int(xp.shape[0] * (geo.nte / geo.nt) ** 3 * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(P, xp, vp, pid, rhoc, cap)
    nt, nc, b, nnt, nte = geo.nt, geo.nc, geo.ncb, geo.nnt, geo.nte
    # global physical counts and CSR
    cnt = np.zeros((nc, nc, nc), np.int64)
    phys = rhoc.reshape(nnt, nnt, nnt, nte, nte, nte)[..., b:b + nt, b:b + nt, b:b + nt]
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                cnt[tz * nt:(tz + 1) * nt, ty * nt:(ty + 1) * nt, tx * nt:(tx + 1) * nt] = phys[tz, ty, tx]
    # physical particle order: tiles (itx fastest), cells i fastest
    gcell = []
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                kk, jj, ii = np.meshgrid(np.arange(nt) + tz * nt, np.arange(nt) + ty * nt, np.arange(nt) + tx * nt,
                                         indexing="ij")
                gcell.append(((kk * nc + jj) * nc + ii).reshape(-1))
    gcell = np.concatenate(gcell)
    start = np.zeros(nc ** 3, np.int64)
    start[gcell] = np.concatenate([[0], np.cumsum(cnt.reshape(-1)[gcell])[:-1]])
    # expected buffered layout: every extended cell copies its periodic source
    exp_r = np.zeros(geo.rhoc_shape, np.int32)
    src = []
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                e = np.arange(nte) - b
                kk, jj, ii = np.meshgrid((e + tz * nt) % nc, (e + ty * nt) % nc, (e + tx * nt) % nc, indexing="ij")
                g = ((kk * nc + jj) * nc + ii).reshape(-1)
                exp_r[tz, ty, tx] = cnt.reshape(-1)[g].reshape(nte, nte, nte)
                src.append(g)
    src = np.concatenate(src)
    n_per = cnt.reshape(-1)[src]
    idx = np.repeat(start[src], n_per) + (np.arange(n_per.sum()) - np.repeat(np.cumsum(n_per) - n_per, n_per))
    assert np.array_equal(re, exp_r)
    assert np.array_equal(pe, pid[idx]) and np.array_equal(xe, xp[idx]) and np.array_equal(ve, vp[idx])


def test_buffer_v_restores_the_periodic_copies():
    """buffer_v on a buffered layout whose physical velocities changed (as
    after a kick) and whose buffer velocities and IDs are stale gives exactly
    the buffer step's layout of the new physical velocities; physical cells
    are not touched."""
    geo = Geometry(64)
    xp, vp, pid, rhoc = physical_set(geo, seed=21, mean=3.0)
    P = oparams(geo)
    cap = int(xp.shape[0] * (geo.nte / geo.nt) ** 3 * 1.2) + 1024
    v_new = (vp + np.random.default_rng(3).normal(size=vp.shape)).astype(np.float32)
    xe, ve, pe, re = orc.buffer(P, xp, v_new, pid, rhoc, cap)
    phys = physical_mask(geo, re)
    v = ve.copy()
    p = pe.copy()
    v[~phys] = -7.0
    p[~phys] = 0
    orc.buffer_v(P, v, re, p)
    assert np.array_equal(v, ve) and np.array_equal(p, pe)
    # velocities only
    v = ve.copy()
    v[~phys] = 3.0
    orc.buffer_v(P, v, re)
    assert np.array_equal(v, ve)


# ---------------------------------------------------------------------- drift
def _drift_spec(geo, xp, vp, pid, rhoc, dt_old, dt, vz_max):
    """Drift from the written spec (drift.f90:21-136): destination cell,
    re-encode and the within-cell order key, computed with numpy sorts.  The
    coefficient dt_mid and the displacement dt_mid*v are FP64 (the reference
    rounds both to real(4))."""
    B = 8 * geo.izipx
    dt_mid = (float(dt_old) + float(dt)) / 2.0
    nlayer = 2 * int(np.ceil(dt_mid * float(np.float32(vz_max)) / 4.0)) + 2
    nte, b, nt = geo.nte, geo.ncb, geo.nt
    r = rhoc.reshape(-1)
    n = int(r.sum())
    cell = np.repeat(np.arange(r.size), r)
    l = np.arange(n) - np.repeat(np.cumsum(r) - r, r)
    T = cell // nte ** 3
    c = cell % nte ** 3
    ijk = np.stack([c % nte, (c // nte) % nte, c // nte ** 2], 1) - b + 1  # Fortran tile-local
    frac = orc.decode(xp, geo.izipx)
    v64 = dt_mid * vp[:n].astype(np.float64)
    delta = v64 / 4.0
    g = np.ceil((ijk - 1) + frac[:n] + delta).astype(np.int64)
    assert g.min() >= 1 - 2 * b and g.max() <= nt + 2 * b
    shift = v64 * 2.0 ** B / 4
    nint = np.sign(shift) * np.floor(np.abs(shift) + 0.5)
    xnew = ((xp[:n].astype(np.int64) + nint.astype(np.int64) + 2 ** (B - 1)) % 2 ** B - 2 ** (B - 1))
    keep = ((g >= 1) & (g <= nt)).all(1)
    pas = (ijk[:, 2] - (1 - b)) % nlayer
    keys = (l, ijk[:, 0], ijk[:, 1], ijk[:, 2], pas, g[:, 0], g[:, 1], g[:, 2], T)
    order = np.lexsort(tuple(k[keep] for k in keys))
    sel = np.nonzero(keep)[0][order]
    rho = np.zeros(geo.rhoc_shape, np.int32).reshape(-1, nte, nte, nte)
    np.add.at(rho, (T[keep], g[keep, 2] + b - 1, g[keep, 1] + b - 1, g[keep, 0] + b - 1), 1)
    return xnew[sel].astype(xp.dtype), vp[sel], pid[sel], rho.reshape(rhoc.shape), nlayer


@pytest.mark.parametrize("izipx,dt", [(2, 0.6), (1, 1.5)])
def test_drift_matches_spec_and_invariants(izipx, dt):
    geo = Geometry(64, izipx=izipx)
    xp, vp, pid, rhoc = buffered_set(geo, seed=21, mean=4.0, sigma_v=2.0)
    vz = float(np.abs(vp[:, 2]).max())
    st, xo, vo, po, ro, oo = orc.drift(oparams(geo), xp, vp, pid, rhoc, 0.4, dt, vz)
    assert st == 0
    xs, vs, ps, rs, nlayer = _drift_spec(geo, xp, vp, pid, rhoc, 0.4, dt, vz)
    assert oo["nlayer"] == nlayer
    assert np.array_equal(ro, rs)
    assert np.array_equal(po, ps) and np.array_equal(xo, xs) and np.array_equal(vo, vs)
    # invariants: every particle kept exactly once, counts match
    nphys = int(rhoc.reshape(-1, geo.nte, geo.nte, geo.nte)[:, geo.ncb:-geo.ncb, geo.ncb:-geo.ncb,
                                                           geo.ncb:-geo.ncb].sum())
    assert oo["nplocal"] == nphys == int(ro.sum())
    assert np.array_equal(np.sort(po), np.arange(1, nphys + 1))


def test_drift_moves_particles_by_dt_v():
    """Re-decoded global positions equal old position + dt_mid v/4 to within
    one coordinate bin (periodic)."""
    geo = Geometry(64)
    rng = np.random.default_rng(5)
    n = 3000
    pos = rng.uniform(0, geo.nc, (n, 3))
    vel = rng.normal(0, 3.0, (n, 3)).astype(np.float32)
    xp, vp, pid, rhoc = buffered_explicit(geo, pos, vel)
    st, xo, vo, po, ro, oo = orc.drift(oparams(geo), xp, vp, pid, rhoc, 0.5, 0.5, float(np.abs(vel[:, 2]).max()))
    assert st == 0 and oo["nplocal"] == n
    # global cell of each output particle from the physical layout
    nt, nnt, b, nte = geo.nt, geo.nnt, geo.ncb, geo.nte
    ph = ro.reshape(nnt, nnt, nnt, nte, nte, nte)[..., b:b + nt, b:b + nt, b:b + nt]
    cells = []
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                kk, jj, ii = np.meshgrid(np.arange(nt) + tz * nt, np.arange(nt) + ty * nt, np.arange(nt) + tx * nt,
                                         indexing="ij")
                cells.append(np.repeat(np.stack([ii, jj, kk], -1).reshape(-1, 3), ph[tz, ty, tx].reshape(-1), 0))
    gc = np.concatenate(cells)
    newpos = gc + orc.decode(xo, geo.izipx)
    old = pos[po - 1]
    exp = old + 0.5 * vel[po - 1].astype(np.float64) / 4
    d = (newpos - exp + geo.nc / 2) % geo.nc - geo.nc / 2
    assert np.abs(d).max() <= 1.5 / 2 ** 16


# ------------------------------------------------------------------------- PP
def _pp_brute(geo, xp, vp, rhoc, it, iapm, a_mid, dt):
    """O(N^2) PP kick of one sub-tile in FP64 from the FP32 fine positions."""
    sf = [0, 1, 2, 4, 6, 8, 12, 16][iapm]
    rng_ = np.float32(3.5) / np.float32(sf)
    nte, b, ntt, ns = geo.nte, geo.ncb, geo.ntt, geo.nns
    T = it // ns ** 3
    ifs = np.array([it % ns, (it // ns) % ns, (it // ns ** 2) % ns]) + 1
    r = rhoc.reshape(-1)
    start = np.cumsum(r) - r
    idx, cells = [], []
    lo, hi = (ifs - 1) * ntt + 1 - 1, ifs * ntt + 1
    for k in range(lo[2], hi[2] + 1):
        for j in range(lo[1], hi[1] + 1):
            for i in range(lo[0], hi[0] + 1):
                c = T * nte ** 3 + ((k + b - 1) * nte + (j + b - 1)) * nte + (i + b - 1)
                idx.append(np.arange(start[c], start[c] + r[c]))
                cells.append(np.repeat([[i, j, k]], r[c], 0))
    idx = np.concatenate(idx)
    ijk = np.concatenate(cells)
    xvec = ((ijk - 1) + orc.decode(xp[idx], geo.izipx)).astype(np.float32)
    xf = (xvec * np.float32(geo.ratio_cs)).astype(np.float64)
    pg = np.floor(np.float32(sf) * (xvec - ((ifs - 1) * ntt).astype(np.float32))).astype(np.int64) + 1
    npg = ntt * sf
    ext = ((pg >= 0) & (pg <= npg + 1)).all(1)
    phy = ((pg >= 1) & (pg <= npg)).all(1)
    xs = xf[ext]
    dv = {}
    app = np.float64(np.float32(0.06))
    from golden.make_golden import f_ra_ref

    for p in np.nonzero(phy)[0]:
        rv = xs - xf[p]
        rr = np.sqrt((rv ** 2).sum(1))
        m = (rr > 0) & (rr < rng_)
        f = (rv[m] / rr[m, None]) * (f_ra_ref(rr[m], app) - f_ra_ref(rr[m], np.float64(rng_)))[:, None]
        dv[idx[p]] = f.sum(0) * (a_mid * dt / 6 / np.pi)
    return dv


@pytest.mark.parametrize("it,iapm", [(0, 2), (1234, 2), (4095, 3)])
def test_pp_box_matches_brute_force(it, iapm):
    geo = Geometry(64)
    xp, vp, pid, rhoc = buffered_set(geo, seed=31 + it, mean=48.0)
    v = vp.copy()
    f2, vm, npairs = orc.pp_box(oparams(geo), xp, v, rhoc, it, iapm, 0.5, 0.3)
    ref = _pp_brute(geo, xp, vp, rhoc, it, iapm, 0.5, 0.3)
    ids = np.array(sorted(ref))
    assert len(ids) > 20 and npairs > 20
    got = (v[ids].astype(np.float64) - vp[ids])
    exp = np.array([ref[i] for i in ids])
    assert rel_rms(got, exp) <= 2e-5
    others = np.setdiff1d(np.arange(vp.shape[0]), ids)
    assert np.array_equal(v[others], vp[others])


# ------------------------------------------------------------- PM end to end
def test_pm1_conserves_momentum():
    geo = Geometry(64)
    xp, vp, pid, rhoc = buffered_set(geo, seed=2, mean=3.0)
    tfs = orc.tf_default(oparams(geo))
    v = vp.copy()
    orc.kick(oparams(geo), xp, v, rhoc, tfs, 0.5, 0.3, phase_mask=1)
    dv = (v.astype(np.float64) - vp)
    assert np.abs(dv).sum() > 0
    assert np.abs(dv.sum(0)).max() <= 1e-5 * np.abs(dv).sum(0).max()


def test_pm1_point_mass_force_law():
    """Two particles 6 coarse cells (24 fine units) apart attract with the
    Newtonian force m/r^2 (fine units) beyond the 3.5-cell softening:
    dv = m a dt / (6 pi r^2)."""
    geo = Geometry(256)
    c = 20.5
    for sep in (6, 9):
        pos = np.array([[c, c, c], [c + sep, c, c]])
        xp, vp, pid, rhoc = buffered_explicit(geo, pos)
        v = vp.copy()
        a, dt = 0.5, 0.3
        tfs = [orc.green(geo.nc, 3.5, 0.0, 4.0)] + [np.zeros(1, np.float32)] * 5  # PM1 reads tf1 only
        orc.kick(oparams(geo), xp, v, rhoc, tfs, a, dt, phase_mask=1)
        dv = {int(p): v[i].astype(np.float64) for i, p in enumerate(pid) if np.any(v[i] != 0)}
        r = sep * geo.ratio_cs
        exp = a * dt / (6 * np.pi * r * r)
        assert dv[1][0] > 0 and dv[2][0] < 0  # attraction along x
        print(f"separation {sep}: dv/expected = {dv[1][0] / exp:.4f}, {dv[2][0] / -exp:.4f}")
        assert abs(dv[1][0] / exp - 1) < 0.03 and abs(dv[2][0] / -exp - 1) < 0.03
        assert abs(dv[1][1]) < 1e-3 * exp and abs(dv[1][2]) < 1e-3 * exp


def test_kick_step_criteria_closed_forms():
    geo = Geometry(64)
    xp, vp, pid, rhoc = buffered_set(geo, seed=6, mean=6.0)
    v = vp.copy()
    a = np.float32(0.5)
    out, _, _ = orc.kick(oparams(geo), xp, v, rhoc, orc.tf_default(oparams(geo)), float(a), 0.3)
    G = np.float32(np.float64(np.float32(1.0) / np.float32(6.0)) / np.pi)
    f32 = np.float32
    for key, fac in (("pm1", 0.5), ("pm2", 0.5), ("pm3", 0.5), ("pp", 0.05)):
        exp = f32(fac) * np.sqrt(f32(1.0) / (np.sqrt(f32(out["f2max_" + key])) * a * G))
        assert abs(out["dt_" + key] - exp) <= 2e-7 * exp, key
    phys = physical_mask(geo, rhoc)
    assert out["vz_max"] == np.abs(v[phys, 2]).max()  # post-PP max|v_z| (kick.f90:315)
    assert abs(out["dt_vmax"] - 0.9 * 14 / np.abs(v[phys]).max()) <= 1e-6 * out["dt_vmax"]


# -------------------------------------------------------------- CIC power
def _cic_numpy(geo, xp, rhoc, nw):
    """CIC deposit of cicpower.f90:67-156 restated with numpy (FP32 positions
    and weights, periodic wrap), mesh summed in FP64."""
    nt, nnt, b, nte = geo.nt, geo.nnt, geo.ncb, geo.nte
    ph = rhoc.reshape(nnt, nnt, nnt, nte, nte, nte)[..., b:b + nt, b:b + nt, b:b + nt]
    cells = []
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                kk, jj, ii = np.meshgrid(np.arange(nt) + tz * nt, np.arange(nt) + ty * nt, np.arange(nt) + tx * nt,
                                         indexing="ij")
                cells.append(np.repeat(np.stack([ii, jj, kk], -1).reshape(-1, 3), ph[tz, ty, tx].reshape(-1), 0))
    gc = np.concatenate(cells)
    pos = (gc + orc.decode(xp, geo.izipx)).astype(np.float32)
    pos = ((pos * np.float32(nw)) / np.float32(geo.nc) - np.float32(0.5)).astype(np.float32)
    f = np.floor(pos).astype(np.int64) + 1
    dx1 = (f.astype(np.float32) - pos).astype(np.float32)
    dx2 = (np.float32(1) - dx1).astype(np.float32)
    grid = np.zeros((nw, nw, nw), np.float64)
    for sx in (0, 1):
        for sy in (0, 1):
            for sz in (0, 1):
                w = ((dx2[:, 0] if sx else dx1[:, 0]) * (dx2[:, 1] if sy else dx1[:, 1])).astype(np.float32)
                w = (w * (dx2[:, 2] if sz else dx1[:, 2])).astype(np.float32)
                ix = (f[:, 0] + sx - 1) % nw
                iy = (f[:, 1] + sy - 1) % nw
                iz = (f[:, 2] + sz - 1) % nw
                np.add.at(grid, (iz, iy, ix), w.astype(np.float64))
    return grid, gc.shape[0]


def test_cic_density_matches_numpy_and_conserves_mass():
    geo = Geometry(64)
    xp, vp, pid, rhoc = physical_set(geo, seed=12, mean=10.0)
    nw = 32
    xi, mean, delta = orc.cicpower(oparams(geo), xp, rhoc, nw, 75.0, want_delta=True)
    grid, n = _cic_numpy(geo, xp, rhoc, nw)
    assert abs(mean * nw ** 3 - n) <= 1e-5 * n  # every particle deposits unit mass
    ref = grid / (grid.sum() / nw ** 3) - 1
    assert np.abs(delta - ref).max() <= 2e-5 * np.abs(ref).max()


def test_cic_auto_power_matches_numpy_binning():
    """Columns 1-4 of auto_power restated with numpy.fft and numpy binning
    from the oracle's own delta field (powerspectrum.f90:158-199)."""
    geo = Geometry(64)
    xp, vp, pid, rhoc = physical_set(geo, seed=13, mean=20.0)
    nw, box = 32, 75.0
    xi, mean, delta = orc.cicpower(oparams(geo), xp, rhoc, nw, box, want_delta=True)
    n = int(rhoc.sum())
    ck = np.fft.rfftn(delta.astype(np.float64)) / nw ** 3  # [z][y][kx]
    nyq = nw // 2
    kf = lambda m: ((m + nyq) % nw) - nyq
    kz, ky, kx = np.meshgrid(kf(np.arange(nw)), kf(np.arange(nw)), np.arange(nyq + 1), indexing="ij")
    ig, jg, kg = kx + 1, np.arange(nw)[None, :, None] + 1 + 0 * kx, np.arange(nw)[:, None, None] + 1 + 0 * kx
    keep = ~((ig == 1) & (jg == 1) & (kg == 1))
    keep &= ~(((ig == 1) | (ig == nyq + 1)) & (jg > nyq + 1))
    keep &= ~(((ig == 1) | (ig == nyq + 1)) & ((jg == 1) | (jg == nyq + 1)) & (kg > nyq + 1))
    kr = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    ib = np.floor(kr + 0.5).astype(np.int64)
    amp = np.abs(ck) ** 2
    c1 = lambda k: 1 - (2.0 / 3.0) * np.sin(np.pi * k / nw) ** 2
    dk = c1(kx) * c1(ky) * c1(kz) / n
    nb = xi.shape[0]
    cnt = np.bincount(ib[keep], minlength=nb)[:nb]
    assert np.array_equal(xi[:, 0], cnt.astype(np.float32))
    sel = cnt > 0
    k_ref = np.bincount(ib[keep], kr[keep], nb)[:nb][sel] / cnt[sel] * 2 * np.pi / box
    p3 = np.bincount(ib[keep], amp[keep], nb)[:nb][sel] / cnt[sel] * box ** 3
    p4 = np.bincount(ib[keep], (amp - dk)[keep], nb)[:nb][sel] / cnt[sel] * box ** 3
    assert np.allclose(xi[sel, 1], k_ref, rtol=1e-6)
    assert np.allclose(xi[sel, 2], p3, rtol=2e-5)
    assert np.abs(xi[sel, 3] - p4).max() <= 2e-5 * np.abs(p3).max()


def test_cic_shot_noise_of_a_poisson_sample():
    """Unclustered particles: <|delta_k|^2> = C_CIC(k)/N, the Dk the reference
    subtracts, so the raw power equals box^3 C/N on average and the
    shot-noise-subtracted power is zero within the sampling error."""
    geo = Geometry(64)
    rng = np.random.default_rng(8)
    n = 120000
    pos = rng.uniform(0, geo.nc, (n, 3))
    from inputs import explicit_set

    xp, vp, pid, rhoc = explicit_set(geo, pos)
    nw, box = 32, 75.0
    xi, mean, _ = orc.cicpower(oparams(geo), xp, rhoc, nw, box)
    nyq = nw // 2
    b = slice(1, nyq)  # up to Nyquist
    w = xi[b, 0]
    shot = box ** 3 / n
    ratio = np.sum(w * xi[b, 2]) / np.sum(w) / shot
    c_mean = np.sum(w * (xi[b, 2] - xi[b, 3])) / np.sum(w) / shot  # the subtracted Dk, averaged
    assert abs(ratio / c_mean - 1) < 0.02
    assert abs(np.sum(w * xi[b, 3]) / np.sum(w)) < 0.02 * shot
