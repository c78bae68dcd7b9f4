"""Pins the oracle's composed pieces that the per-function tests leave open:

* the PM2/PM3 per-box pipeline (deposit windows, utile_shift, the nex crop,
  the nwork^3 scaling, the interpolation range) and its composition with PM1
  and PP in the full kick, by the force between two particles: the sum of the
  four levels is Newton's m/r^2 from 0.3 to 24 fine cells, and each level's
  share vanishes outside the range its Green function gives it
  (kick.f90:54-58, 212-219, 593-862; Green.f90:22-71);
* the alias correction of P(k) (column 5, powerspectrum.f90:212-259) by an
  independent numpy image sum over the oracle's own columns 2 and 4;

No GPU is involved.
"""
import numpy as np
import pytest

from cube_nu_amd.p3m import Geometry
from inputs import buffered_explicit, explicit_set, oparams, physical_mask
from oracle import pyoracle as orc


@pytest.fixture(scope="module")
def geo128():
    return Geometry(128)


@pytest.fixture(scope="module")
def tfs128(geo128):
    return orc.tf_default(oparams(geo128))


# separation (fine cells) -> which levels may carry force there.  Ranges: PP
# below r_pp = 3.5/ratio_sf = 1.75 fine cells (ires 2), PM3 between its
# softening (3.5 fine-fine cells) and its S2 range apm2 = 3.5 fine cells, PM2
# up to apm1 * ratio_cs = 14 fine cells, PM1 beyond.
SEPARATIONS = (0.3, 1.0, 2.0, 4.0, 8.0, 16.0, 24.0)


@pytest.mark.parametrize("r", SEPARATIONS)
def test_two_body_force_through_all_levels(geo128, tfs128, r):
    geo = geo128
    P = oparams(geo)
    a, dt = 0.5, 0.3
    u = np.array([1.0, 0.55, 0.3])
    u /= np.linalg.norm(u)  # off-axis: a transposed component or axis mix-up shows as a transverse force
    c = np.array([9.3, 17.7, 12.4])  # coarse units, inside a tile away from the mesh symmetry points
    pos = np.array([c, c + (r / geo.ratio_cs) * u])
    xp, vp, pid, rhoc = buffered_explicit(geo, pos)
    ph = physical_mask(geo, rhoc)
    i1 = int(np.where((pid == 1) & ph)[0][0])
    i2 = int(np.where((pid == 2) & ph)[0][0])
    share = {}
    for name, mask in (("pm1", 1), ("pm2", 2), ("pm3", 4), ("pp", 8)):
        v = vp.copy()
        orc.kick(P, xp, v, rhoc, tfs128, a, dt, phase_mask=mask)
        share[name] = (v[i1].astype(np.float64), v[i2].astype(np.float64))
    newton = a * dt / (6 * np.pi * r * r)  # dv = m a dt / (6 pi r^2), m = 1, r in fine cells
    tot1 = sum(s[0] for s in share.values())
    tot2 = sum(s[1] for s in share.values())
    along = {k: s[0] @ u / newton for k, s in share.items()}
    print(f"r={r}: total/Newton {tot1 @ u / newton:.4f}; " + " ".join(f"{k} {v:+.4f}" for k, v in along.items()))
    # attraction along the separation, equal and opposite, Newtonian to 5%
    assert abs(tot1 @ u / newton - 1) < 0.05
    assert abs(-(tot2 @ u) / newton - 1) < 0.05
    assert np.linalg.norm(tot1 - (tot1 @ u) * u) < 0.04 * newton
    # each level's share vanishes outside its range
    if r > 1.75:
        assert np.all(share["pp"][0] == 0) and np.all(share["pp"][1] == 0)  # PP cut at r_pp exactly
    else:
        assert along["pp"] > 0.1
    if r >= 4.0:
        assert abs(along["pm3"]) < 2e-3
    if r <= 2.0:
        assert along["pm3"] > 0.04
    if r >= 16.0:
        assert abs(along["pm2"]) < 2e-3
        assert along["pm1"] > 0.95
    if 2.0 <= r <= 8.0:
        assert along["pm2"] > 0.15
    if r <= 1.0:
        assert along["pm1"] < 0.01


# ------------------------------------------------------------ pk_correction
def _interp1(x, y, nbin, xq):
    """powerspectrum.f90:261-275 on xdata(1:nbin) = bins 0..nbin-1 (bisection)."""
    i1, i2 = 1, nbin
    while i2 - i1 > 1:
        im = (i1 + i2) // 2
        if xq > x[im - 1]:
            i1 = im
        else:
            i2 = im
    return y[i1 - 1] + (xq - x[i1 - 1]) / (x[i2 - 1] - x[i1 - 1]) * (y[i2 - 1] - y[i1 - 1])


def _alias_factor(nw, nyq, nbin, alpha, n_img=3, p=2):
    """<sum_n W^2(k + n nw) |k + n nw|^alpha / |k|^alpha> per bin over the
    octant i >= j >= k (powerspectrum.f90:225-252), vectorised."""
    i, j, k = np.meshgrid(np.arange(nyq + 1), np.arange(nyq + 1), np.arange(nyq + 1), indexing="ij")
    sel = (j <= i) & (k <= j)
    kv = np.stack([i[sel], j[sel], k[sel]], -1).astype(np.float64)
    kmag = np.linalg.norm(kv, axis=1)
    ib = np.floor(kmag + 0.5).astype(np.int64)
    keep = (ib >= 1) & (ib <= nbin)
    kv, kmag, ib = kv[keep], kmag[keep], ib[keep]
    img = np.arange(-n_img, n_img + 1) * float(nw)
    w = np.zeros(kmag.size)
    for a in img:
        for b in img:
            for c in img:
                kn = kv + np.array([a, b, c])
                ks = np.pi * kn / nw
                sinc = np.where(ks == 0, 1.0, np.sin(ks) / np.where(ks == 0, 1.0, ks))
                w += np.prod(sinc, axis=1) ** (2 * p) * np.linalg.norm(kn, axis=1) ** alpha
    cnt = np.bincount(ib, minlength=nbin + 1)
    fac = np.bincount(ib, w / kmag ** alpha, nbin + 1)
    out = np.full(nbin + 1, np.nan)
    out[cnt > 0] = fac[cnt > 0] / cnt[cnt > 0]
    return out


def test_pk_correction_matches_numpy_image_sum():
    """Column 5 = column 4 divided by the CIC alias factor of a power law
    whose slope alpha is re-measured from column 5 between k_N/2 and k_N, the
    correction applied four times as auto_power calls it
    (powerspectrum.f90:197-200), restated with numpy from columns 2 and 4."""
    geo = Geometry(64)
    # a clustered set with a falling spectrum (Soneira-Peebles hierarchy:
    # 40 centres, five levels of six children at radii shrinking by 1.9), so
    # that alpha is a number and the correction is applied
    rng = np.random.default_rng(31)
    pts, R = rng.uniform(0, geo.nc, (40, 3)), geo.nc / 3.0
    for _ in range(5):
        R /= 1.9
        pts = np.repeat(pts, 6, 0) + rng.uniform(-R, R, (pts.shape[0] * 6, 3))
    xp, vp, pid, rhoc = explicit_set(geo, pts % geo.nc)
    nw, box = 32, 75.0
    xi, mean, _ = orc.cicpower(oparams(geo), xp, rhoc, nw, box)
    nyq = nw // 2
    nbin = xi.shape[0] - 1
    kg = xi[:, 1].astype(np.float64) / (2 * np.pi / box)  # k in grid units
    p4 = xi[:, 3].astype(np.float64) / box ** 3
    p5 = p4.copy()
    for _ in range(4):
        alpha = (np.log(_interp1(kg, p5, nbin, float(nyq))) - np.log(_interp1(kg, p5, nbin, nyq / 2.0))) / np.log(2.0)
        if np.isnan(alpha):
            break
        fac = _alias_factor(nw, nyq, nbin, alpha)
        p5[1:] = p4[1:] / fac[1:]
    got = xi[1:, 4].astype(np.float64) / box ** 3
    ref = p5[1:]
    ok = np.isfinite(ref)
    print(f"alpha {alpha:.4f}; max rel diff {np.max(np.abs(got[ok] - ref[ok]) / np.abs(ref[ok])):.2e}")
    assert np.all(np.abs(got[ok] - ref[ok]) <= 1e-4 * np.abs(ref[ok]))
    # the alias factor is ~1 at low k and below 1 towards Nyquist (sinc^4 < 1
    # outweighs the images of a falling spectrum), so the correction raises P
    fac = _alias_factor(nw, nyq, nbin, alpha)
    assert abs(fac[1] - 1) < 0.01 and fac[nyq] < 0.9
