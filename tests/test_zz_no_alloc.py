"""cube_pm_cic_noalloc (include/cube_pm.h): the allocation-free CIC density.

It must give cube_pm_cic's result bit for bit (both sum integer fixed-point
weights), on two calls in a row on one GPU, the second with fewer particles
and a smaller mesh (the fixed scratch keeps the first call's contents and
must be cleared), and reject a mesh beyond its fixed scratch.
"""
import ctypes as C

import pytest
import torch

from cube_nu_amd import pmdiff
from cube_nu_amd.p3m import CUBE_ERR_ARG

pytestmark = pytest.mark.gpu


def _positions(seed, n_total, box):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(n_total, 3, generator=g, dtype=torch.float32) * box).contiguous().cuda()


def _noalloc(pm, x, nw, scale):
    rho = torch.full((pm.batch, nw, nw, nw), float("nan"), dtype=torch.float32, device=x.device)
    rc = pm.lib.cube_pm_cic_noalloc(pm.ctx, pmdiff._ptr(x), pm._np(x), nw, float(scale), pmdiff._ptr(rho),
                                    pmdiff._stream())
    return rc, rho


def test_noalloc_cic_matches_cic_on_two_calls():
    pm = pmdiff.PMForce(32, batch=2, apm_soft=1.0)
    try:
        cases = [(1, 2 * 40000, 64, 64 / 32.0), (2, 2 * 5000, 24, 24 / 32.0)]
        for seed, ntot, nw, scale in cases:
            x = _positions(seed, ntot, 32.0)
            ref = pm.cic(x, nw, scale)
            rc, rho = _noalloc(pm, x, nw, scale)
            torch.cuda.synchronize()
            assert rc == 0, pm.lib.cube_pm_last_error(pm.ctx).decode()
            print(f"seed {seed}: {ntot} particles, nw {nw}: mass {float(rho.sum()):.3f}")
            assert torch.equal(rho.view(-1), ref.view(-1))
            assert abs(float(rho.double().sum()) - ntot) <= 1e-3 * ntot
    finally:
        pm.close()


def test_noalloc_cic_rejects_a_mesh_beyond_its_scratch():
    pm = pmdiff.PMForce(16, batch=2, apm_soft=1.0)
    try:
        x = _positions(3, 2 * 100, 16.0)
        out = torch.empty(1, dtype=torch.float32, device=x.device)
        # 2 * 256^3 = 2^25 cells > 2^23: refused before any launch
        rc = pm.lib.cube_pm_cic_noalloc(pm.ctx, pmdiff._ptr(x), 100, 256, 1.0, pmdiff._ptr(out),
                                        pmdiff._stream())
        assert rc == CUBE_ERR_ARG
        assert b"exceeds the fixed scratch" in pm.lib.cube_pm_last_error(pm.ctx)
    finally:
        pm.close()
