"""Differentiable PM-only force (include/cube_pm.h, cube_nu_amd.pmdiff).

CPU: the FP64 plain-PyTorch model (oracle/pm_torch.py) is pinned by centred
finite differences of its own autograd gradient, by exact momentum
conservation (sum_p acc_p = rho^T D G rho = 0 for antisymmetric D and
symmetric G) and by the Green kernel of the oracle.  GPU: the HIP kernels'
values and hand-written adjoints against that model: force rms relative
<= 1e-5, VJPs <= 1e-4, and the gradient of a multi-step loss with respect
to the white-noise modes <= 1e-4 of the FP64 autograd gradient.
"""
import numpy as np
import pytest
import torch

from cube_nu_amd import pmdiff
from oracle import pm_torch as ref
from oracle import pyoracle as orc

SOFT = 1.0


def _G(n):
    return torch.from_numpy(orc.green(n, SOFT, 0.0, 1.0).astype(np.float64))


def _positions(seed, batch, np_, n, dtype=torch.float64):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(batch * np_, 3, generator=g, dtype=dtype) * n).contiguous()


def test_model_force_gradient_matches_finite_differences():
    n, B = 16, 2
    G = _G(n)
    x = _positions(1, B, 300, n)
    u = torch.randn(x.shape, generator=torch.Generator().manual_seed(2), dtype=torch.float64)
    h = torch.randn(x.shape, generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    xr = x.clone().requires_grad_(True)
    L = (ref.force(xr, G, B, n) * u).sum()
    (gx,) = torch.autograd.grad(L, xr)
    eps = 1e-5
    fd = ((ref.force(x + eps * h, G, B, n) * u).sum() - (ref.force(x - eps * h, G, B, n) * u).sum()) / (2 * eps)
    assert abs(float((gx * h).sum()) - float(fd)) <= 1e-6 * abs(float(fd))


def test_model_conserves_momentum():
    n, B = 16, 1
    acc = ref.force(_positions(4, B, 500, n), _G(n), B, n)
    assert float(acc.sum(0).abs().max()) <= 1e-10 * float(acc.abs().sum())


@pytest.mark.parametrize("r", [4.0, 6.0, 8.0, 12.0])
def test_model_point_mass_attraction_is_newtonian(r):
    """Two unit masses r mesh cells apart (beyond the softening) in a periodic
    96^3 box: each is pulled toward the other with |acc| = 1/r^2 (G from
    Green_3D at apm_range = 0 is the 4 pi / k^2 kernel, the FD stencil its
    gradient), to the mesh's anisotropy of a few per cent (at most 7% over
    placements at these separations) and the small pull of the periodic
    images and of the zeroed k = 0 mode.  A flipped sign, a transposed axis or a wrong stencil
    weight fails this."""
    n = 96
    G = _G(n)
    u = torch.tensor([1.0, 0.55, 0.3], dtype=torch.float64)
    u = u / torch.linalg.vector_norm(u)
    c = torch.tensor([40.3, 47.6, 45.2], dtype=torch.float64)
    x = torch.stack([c, c + r * u])
    acc = ref.force(x, G, 1, n)
    along = acc @ u
    print(f"r={r}: r^2 acc.u = {float(along[0]) * r * r:.4f}, {float(along[1]) * r * r:.4f}")
    assert abs(float(along[0]) * r * r - 1) < 0.08  # particle 0 pulled toward +u
    assert abs(-float(along[1]) * r * r - 1) < 0.08
    assert float(torch.linalg.vector_norm(acc[0] - along[0] * u)) < 0.05 / (r * r)


def test_model_multistep_gradient_matches_finite_differences():
    n, ng, B, nw = 16, 8, 1, 8
    G = _G(n)
    k1, d, k2 = pmdiff.kdk_coefficients(0.5, 1.0, 3)
    k1, k2 = [c * 0.05 for c in k1], [c * 0.05 for c in k2]
    noise = torch.randn(B, ng, ng, ng, generator=torch.Generator().manual_seed(5), dtype=torch.float64)
    target = torch.rand(B, nw, nw, nw, generator=torch.Generator().manual_seed(6), dtype=torch.float64)

    def loss(z):
        x, v = pmdiff.zeldovich(z, box=64.0, n_mesh=n, amp=20.0)
        x, v = ref.kdk(x, v, G, B, n, k1, d, k2)
        return ((ref.cic(x, B, nw, nw / n) - target) ** 2).mean()

    zr = noise.clone().requires_grad_(True)
    (gz,) = torch.autograd.grad(loss(zr), zr)
    h = torch.randn(noise.shape, generator=torch.Generator().manual_seed(7), dtype=torch.float64)
    eps = 1e-6
    fd = (loss(noise + eps * h) - loss(noise - eps * h)) / (2 * eps)
    assert abs(float((gz * h).sum()) - float(fd)) <= 1e-5 * abs(float(fd))


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.gpu
def test_gpu_green_matches_oracle():
    pm = pmdiff.PMForce(32, 1, SOFT)
    g = pm.green()
    o = orc.green(32, SOFT, 0.0, 1.0)
    assert np.max(np.abs(g - o)) <= 2e-6 * np.abs(o).max()
    pm.close()


@pytest.mark.gpu
def test_gpu_force_and_vjp_match_fp64_model():
    n, B, np_ = 32, 2, 5000
    pm = pmdiff.PMForce(n, B, SOFT)
    G = torch.from_numpy(pm.green().astype(np.float64))
    x64 = _positions(8, B, np_, n)
    g64 = torch.randn(x64.shape, generator=torch.Generator().manual_seed(9), dtype=torch.float64)
    xr = x64.clone().requires_grad_(True)
    acc_ref = ref.force(xr, G, B, n)
    (gx_ref,) = torch.autograd.grad((acc_ref * g64).sum(), xr)
    x = x64.float().cuda()
    acc, F = pm.force(x)
    gx = pm.force_vjp(x, F, g64.float().cuda())
    e_f, e_g = _rel(acc.double().cpu(), acc_ref.detach()), _rel(gx.double().cpu(), gx_ref)
    print(f"force rel {e_f:.2e} vjp rel {e_g:.2e}")
    assert e_f <= 1e-5 and e_g <= 1e-4
    # the same call again gives the same bits (integer mesh sums)
    acc2, F2 = pm.force(x)
    assert torch.equal(acc, acc2) and torch.equal(pm.force_vjp(x, F2, g64.float().cuda()), gx)
    pm.close()


@pytest.mark.gpu
def test_gpu_cic_and_vjp_match_fp64_model():
    n, B, np_, nw = 32, 2, 4000, 16
    pm = pmdiff.PMForce(n, B, SOFT)
    x64 = _positions(10, B, np_, n)
    gr64 = torch.randn(B, nw, nw, nw, generator=torch.Generator().manual_seed(11), dtype=torch.float64)
    xr = x64.clone().requires_grad_(True)
    rho_ref = ref.cic(xr, B, nw, nw / n)
    (gx_ref,) = torch.autograd.grad((rho_ref * gr64).sum(), xr)
    x = x64.float().cuda()
    rho = pm.cic(x, nw, nw / n)
    gx = pm.cic_vjp(x, nw, nw / n, gr64.float().cuda())
    assert _rel(rho.double().cpu(), rho_ref.detach()) <= 1e-6
    assert _rel(gx.double().cpu(), gx_ref) <= 1e-5
    pm.close()


@pytest.mark.gpu
def test_gpu_multistep_gradient_matches_fp64_autograd():
    """Zel'dovich from noise -> 3 KDK steps -> CIC -> MSE: the gradient with
    respect to the noise through the HIP forward and adjoint kernels."""
    n, ng, B, nw = 32, 16, 2, 16
    k1, d, k2 = pmdiff.kdk_coefficients(0.5, 1.0, 3)
    k1, k2 = [c * 0.02 for c in k1], [c * 0.02 for c in k2]
    noise = torch.randn(B, ng, ng, ng, generator=torch.Generator().manual_seed(12), dtype=torch.float64)
    target = torch.rand(B, nw, nw, nw, generator=torch.Generator().manual_seed(13), dtype=torch.float64)
    pm = pmdiff.PMForce(n, B, SOFT)
    G = torch.from_numpy(pm.green().astype(np.float64))

    zr = noise.clone().requires_grad_(True)
    x, v = pmdiff.zeldovich(zr, box=128.0, n_mesh=n, amp=20.0)
    x, v = ref.kdk(x, v, G, B, n, k1, d, k2)
    L_ref = ((ref.cic(x, B, nw, nw / n) - target) ** 2).mean()
    (g_ref,) = torch.autograd.grad(L_ref, zr)

    zg = noise.float().cuda().requires_grad_(True)
    x, v = pmdiff.zeldovich(zg, box=128.0, n_mesh=n, amp=20.0)
    x, v = pmdiff.kdk(x, v, pm, k1, d, k2)
    L = ((pmdiff.cic_density(x, pm, nw, nw / n) - target.float().cuda()) ** 2).mean()
    L.backward()
    e = _rel(zg.grad.double().cpu(), g_ref)
    print(f"loss {float(L):.6e} vs {float(L_ref):.6e}; gradient rel {e:.2e}")
    assert abs(float(L) - float(L_ref)) <= 1e-5 * float(L_ref)
    assert e <= 1e-4
    pm.close()


@pytest.mark.gpu
def test_gpu_lattice_displacements_match_absolute_positions():
    """A context with lattice = ng takes displacements from the ng^3 sites:
    force, VJP and CIC density of the displacements s equal those of the
    absolute positions q + s (FP32 rounding of q + s aside)."""
    n, ng, B, nw = 32, 16, 2, 16
    pa = pmdiff.PMForce(n, B, SOFT)
    pl = pmdiff.PMForce(n, B, SOFT, lattice=ng)
    try:
        gen = torch.Generator().manual_seed(41)
        s64 = torch.randn(B * ng ** 3, 3, generator=gen, dtype=torch.float64) * 1.5
        q = pl.sites().repeat(B, 1)
        x = (q + s64).float().cuda()
        s = s64.float().cuda()
        g = torch.randn(B * ng ** 3, 3, generator=gen).cuda()
        acc_a, Fa = pa.force(x)
        acc_l, Fl = pl.force(s)
        assert _rel(acc_l.double(), acc_a.double()) <= 1e-5
        assert _rel(pl.force_vjp(s, None, g).double(), pa.force_vjp(x, Fa, g).double()) <= 1e-4
        assert _rel(pl.cic(s, nw, nw / n).double(), pa.cic(x, nw, nw / n).double()) <= 1e-5
    finally:
        pa.close()
        pl.close()


@pytest.mark.gpu
def test_gpu_gradient_at_the_cfg5_size_matches_fp64_model():
    """The cfg5 shape with one seed: 128^3 particles from white noise on a
    256^3 mesh (box 256 Mpc/h), 10 KDK steps from a = 0.1 to 1, loss = MSE of
    the final 128^3 CIC density against a second seed's; the gradient with
    respect to the 128^3 noise modes through the HIP forward and hand-written
    adjoint kernels (FP32, displacements from the particle lattice,
    checkpointed positions with the force meshes recomputed in the backward)
    against the FP64 model's autograd gradient (run with torch's own float64
    GPU kernels), <= 1e-4 relative."""
    B, ng, n, nw, nsteps, box, amp = 1, 128, 256, 128, 10, 256.0, 0.5
    k1, d, k2 = pmdiff.kdk_coefficients(0.1, 1.0, nsteps)
    s = 4 * np.pi * ng ** 3 / n ** 3
    k1, k2 = [c / s for c in k1], [c / s for c in k2]
    pm = pmdiff.PMForce(n, B, SOFT, lattice=ng)
    try:
        G = torch.from_numpy(pm.green().astype(np.float64)).cuda()
        gen = torch.Generator().manual_seed(31)
        noise = torch.randn(B, ng, ng, ng, generator=gen, dtype=torch.float64)
        z5 = torch.randn(1, ng, ng, ng, generator=gen, dtype=torch.float64)
        with torch.no_grad():
            x, v = pmdiff.zeldovich(z5.cuda(), box, n, amp)
            x, v = ref.kdk(x, v, G, 1, n, k1, d, k2)
            target = ref.cic(x, 1, nw, nw / n)
        zr = noise.cuda().requires_grad_(True)
        x, v = pmdiff.zeldovich(zr, box, n, amp)
        x, v = ref.kdk(x, v, G, B, n, k1, d, k2)
        L_ref = ((ref.cic(x, B, nw, nw / n) - target) ** 2).mean()
        (g_ref,) = torch.autograd.grad(L_ref, zr)
        del x, v
        zg = noise.float().cuda().requires_grad_(True)
        x, v = pmdiff.zeldovich(zg, box, n, amp, displacement=True)
        x, v = pmdiff.kdk(x, v, pm, k1, d, k2)
        L = ((pmdiff.cic_density(x, pm, nw, nw / n).double() - target) ** 2).mean()
        L.backward()
        e = _rel(zg.grad.double(), g_ref)
        el = abs(float(L.detach()) - float(L_ref)) / float(L_ref)
        print(f"cfg5 size: loss {float(L.detach()):.6e} vs {float(L_ref):.6e} (rel {el:.1e}); gradient rel {e:.2e}")
        assert el <= 1e-5
        assert e <= 1e-4
    finally:
        pm.close()
