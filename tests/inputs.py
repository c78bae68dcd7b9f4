"""Seeded synthetic particle sets in CUBE's layouts, for parity tests.

Counts per coarse cell are Poisson (optionally with dense cells that push a
sub-tile to a finer PM3 level), positions are uniformly random compressed
coordinates, velocities Gaussian.  The physical layout follows drift.f90's
output order (tiles with itx fastest, cells (k,j,i)); the buffered layout is
produced by the oracle's nn=1 buffer step.
"""
from __future__ import annotations

import numpy as np

from cube_nu_amd.p3m import Geometry


def physical_set(geo: Geometry, seed: int, mean: float = 8.0, sigma_v: float = 1.5, dense=()):
    """Returns (xp, vp, pid, rhoc_phys) in the physical layout.

    ``dense`` is a list of ((gx, gy, gz), count) global coarse cells to fill.
    """
    rng = np.random.default_rng(seed)
    nc, nt, nnt = geo.nc, geo.nt, geo.nnt
    counts = rng.poisson(mean, size=(nc, nc, nc)).astype(np.int32)  # [z][y][x]
    for (gx, gy, gz), c in dense:
        counts[gz, gy, gx] = c
    rhoc = np.zeros(geo.rhoc_shape, np.int32)
    order = []
    for itz in range(nnt):
        for ity in range(nnt):
            for itx in range(nnt):
                blk = counts[itz * nt:(itz + 1) * nt, ity * nt:(ity + 1) * nt, itx * nt:(itx + 1) * nt]
                rhoc[itz, ity, itx, geo.ncb:geo.ncb + nt, geo.ncb:geo.ncb + nt, geo.ncb:geo.ncb + nt] = blk
                order.append(blk.reshape(-1))
    n = int(counts.sum())
    info = np.iinfo(geo.xp_dtype)
    xp = rng.integers(info.min, info.max + 1, size=(n, 3), dtype=np.int64).astype(geo.xp_dtype)
    vp = (rng.standard_normal((n, 3)) * sigma_v).astype(np.float32)
    pid = (rng.permutation(n) + 1).astype(np.int64)
    return xp, vp, pid, rhoc


def buffered_set(geo: Geometry, seed: int, **kw):
    """Physical set pushed through the oracle's nn=1 buffer step."""
    from oracle import pyoracle as orc

    xp, vp, pid, rhoc = physical_set(geo, seed, **kw)
    P = oparams(geo)
    cap = int(xp.shape[0] * ((geo.nte / geo.nt) ** 3) * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(P, xp, vp, pid, rhoc, cap)
    return np.ascontiguousarray(xe), np.ascontiguousarray(ve), np.ascontiguousarray(pe), re


def oparams(geo: Geometry, **kw):
    from oracle import pyoracle as orc

    return orc.params(geo.ng, izipx=geo.izipx, nnt=geo.nnt, nns=geo.nns, ratio_cs=geo.ratio_cs, ncb=geo.ncb,
                      ngb=geo.ngb, tile_buffer=geo.tile_buffer, image_buffer=geo.image_buffer, **kw)


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.mean(b * b))
    return float(np.sqrt(np.mean((a - b) ** 2)) / (den if den > 0 else 1.0))
