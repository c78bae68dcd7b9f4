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


def encode(frac, izipx):
    """Compressed coordinate of cell fractions in [0,1): u = floor(frac*2^B) as
    the two's-complement izipx-byte integer (ic.f90:708-718)."""
    B = 8 * izipx
    u = np.floor(np.asarray(frac, np.float64) * 2.0 ** B).astype(np.int64) % (2 ** B)
    return (((u + 2 ** (B - 1)) % 2 ** B) - 2 ** (B - 1)).astype(np.int8 if izipx == 1 else np.int16)


def explicit_set(geo: Geometry, pos, vel=None):
    """Physical layout of particles at global coarse-unit positions pos (n, 3)
    in [0, nc): tiles with itx fastest, cells with i fastest, input order
    within a cell.  Returns (xp, vp, pid, rhoc_phys); pid = input index + 1."""
    pos = np.asarray(pos, np.float64) % geo.nc
    n = pos.shape[0]
    vel = np.zeros((n, 3), np.float32) if vel is None else np.asarray(vel, np.float32)
    cell = np.floor(pos).astype(np.int64)
    tile = cell // geo.nt
    loc = cell % geo.nt
    T = (tile[:, 2] * geo.nnt + tile[:, 1]) * geo.nnt + tile[:, 0]
    key = ((T * geo.nt + loc[:, 2]) * geo.nt + loc[:, 1]) * geo.nt + loc[:, 0]
    order = np.argsort(key, kind="stable")
    rhoc = np.zeros(geo.rhoc_shape, np.int32)
    b = geo.ncb
    np.add.at(rhoc, (tile[:, 2], tile[:, 1], tile[:, 0], loc[:, 2] + b, loc[:, 1] + b, loc[:, 0] + b), 1)
    xp = encode(pos - cell, geo.izipx)[order]
    return np.ascontiguousarray(xp), np.ascontiguousarray(vel[order]), (order + 1).astype(np.int64), rhoc


def buffered_explicit(geo: Geometry, pos, vel=None):
    """explicit_set pushed through the oracle's nn=1 buffer step."""
    from oracle import pyoracle as orc

    xp, vp, pid, rhoc = explicit_set(geo, pos, vel)
    cap = int(xp.shape[0] * ((geo.nte / geo.nt) ** 3) * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(oparams(geo), xp, vp, pid, rhoc, cap)
    return np.ascontiguousarray(xe), np.ascontiguousarray(ve), np.ascontiguousarray(pe), re


def buffered_set(geo: Geometry, seed: int, **kw):
    """Physical set pushed through the oracle's nn=1 buffer step."""
    from oracle import pyoracle as orc

    xp, vp, pid, rhoc = physical_set(geo, seed, **kw)
    P = oparams(geo)
    cap = int(xp.shape[0] * ((geo.nte / geo.nt) ** 3) * 1.2) + 1024
    xe, ve, pe, re = orc.buffer(P, xp, vp, pid, rhoc, cap)
    return np.ascontiguousarray(xe), np.ascontiguousarray(ve), np.ascontiguousarray(pe), re


def physical_mask(geo: Geometry, rhoc):
    """True for the particles of a buffered layout that sit in physical cells."""
    b, nt, nte = geo.ncb, geo.nt, geo.nte
    e = np.arange(nte)
    inside = (e >= b) & (e < b + nt)
    cell_phys = inside[:, None, None] & inside[None, :, None] & inside[None, None, :]
    per_tile = np.broadcast_to(cell_phys, (geo.nnt ** 3, nte, nte, nte)).reshape(-1)
    return np.repeat(per_tile, rhoc.reshape(-1))


def oparams(geo: Geometry, **kw):
    from oracle import pyoracle as orc

    return orc.params(geo.ng, izipx=geo.izipx, nnt=geo.nnt, nns=geo.nns, ratio_cs=geo.ratio_cs, ncb=geo.ncb,
                      ngb=geo.ngb, tile_buffer=geo.tile_buffer, image_buffer=geo.image_buffer, **kw)


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.mean(b * b))
    return float(np.sqrt(np.mean((a - b) ** 2)) / (den if den > 0 else 1.0))
