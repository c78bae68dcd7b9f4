"""TEST INFRASTRUCTURE ONLY -- plain-PyTorch FP64 restatement of the
differentiable PM operator of include/cube_pm.h, for the checker.

Only tests/ and bench.py's checks may import it.  It evaluates in float64 on
the device of its inputs (the CPU for the small cases; torch's own GPU
kernels for the 256^3-mesh comparison, which would take many minutes on the
CPU), with torch's autograd providing the exact gradient of the same
discrete model, against which the HIP kernels' values and hand-written
adjoints are compared.  The reference has no PM-only differentiable path;
the pieces restated are PM1's (TSC deposit and interpolation,
particle_mesh.f90:36-62 and 182-205; the 4-point FD gradient,
particle_mesh.f90:237-242; Green_3D with apm_range = 0, Green.f90:22-71),
with the kernel G taken from the library (its own values are pinned in
tests/test_oracle.py against an independent numpy Green_3D).  This model is
pinned by finite differences of itself and by its adjoint identities
(tests/test_pmdiff.py); it is not pinned against any reference output.
"""
from __future__ import annotations

import torch

FD_C = (1.0 / 12.0, -2.0 / 3.0, 2.0 / 3.0, -1.0 / 12.0)
FD_S = (-2, -1, 1, 2)


def tsc_weights(x: torch.Tensor):
    """Nearest point i = floor(x + 1/2), d = x - i; weights on i-1, i, i+1:
    ((1/2 - d)^2/2, 3/4 - d^2, (1/2 + d)^2/2).  x (N, 3) -> i (N, 3) long,
    w (N, 3 dims, 3 taps)."""
    fi = torch.floor(x + 0.5).detach()
    d = x - fi
    a, b = 0.5 - d, 0.5 + d
    return fi.long(), torch.stack([0.5 * a * a, 0.75 - d * d, 0.5 * b * b], dim=-1)


def _tap_index(i, b, n, k1, k2, k3):
    z = (i[:, 2] - 1 + k3) % n
    y = (i[:, 1] - 1 + k2) % n
    x = (i[:, 0] - 1 + k1) % n
    return ((b * n + z) * n + y) * n + x


def deposit(x: torch.Tensor, batch: int, n: int) -> torch.Tensor:
    """rho (batch, n, n, n) [z][y][x] = sum_p W(x_p - m)."""
    np_ = x.shape[0] // batch
    i, w = tsc_weights(x)
    b = torch.arange(batch, device=x.device).repeat_interleave(np_)
    rho = torch.zeros(batch * n ** 3, dtype=x.dtype, device=x.device)
    for k3 in range(3):
        for k2 in range(3):
            for k1 in range(3):
                rho = rho.index_add(0, _tap_index(i, b, n, k1, k2, k3), w[:, 0, k1] * w[:, 1, k2] * w[:, 2, k3])
    return rho.view(batch, n, n, n)


def potential(rho: torch.Tensor, G: torch.Tensor) -> torch.Tensor:
    """phi = c2r(G r2c(rho)) / n^3 (torch's irfftn carries the 1/n^3)."""
    n = rho.shape[-1]
    return torch.fft.irfftn(torch.fft.rfftn(rho, dim=(1, 2, 3)) * G, s=(n, n, n), dim=(1, 2, 3))


def fd_force(phi: torch.Tensor) -> torch.Tensor:
    """F (batch, 3, n, n, n): F_d(m) = sum_s c_s phi(m + s e_d)."""
    out = []
    for axis in (3, 2, 1):  # x, y, z
        f = torch.zeros_like(phi)
        for c, s in zip(FD_C, FD_S):
            f = f + c * torch.roll(phi, shifts=-s, dims=axis)
        out.append(f)
    return torch.stack(out, dim=1)


def interp(x: torch.Tensor, F: torch.Tensor) -> torch.Tensor:
    batch, n = F.shape[0], F.shape[-1]
    np_ = x.shape[0] // batch
    i, w = tsc_weights(x)
    b = torch.arange(batch, device=x.device).repeat_interleave(np_)
    Ff = F.permute(0, 2, 3, 4, 1).reshape(batch * n ** 3, 3)
    acc = torch.zeros_like(x)
    for k3 in range(3):
        for k2 in range(3):
            for k1 in range(3):
                wt = (w[:, 0, k1] * w[:, 1, k2] * w[:, 2, k3]).unsqueeze(1)
                acc = acc + wt * Ff[_tap_index(i, b, n, k1, k2, k3)]
    return acc


def force(x: torch.Tensor, G: torch.Tensor, batch: int, n: int) -> torch.Tensor:
    return interp(x, fd_force(potential(deposit(x, batch, n), G)))


def cic(x: torch.Tensor, batch: int, nw: int, scale: float) -> torch.Tensor:
    np_ = x.shape[0] // batch
    y = x * scale
    fi = torch.floor(y).detach()
    d = y - fi
    i = fi.long()
    w = torch.stack([1 - d, d], dim=-1)
    b = torch.arange(batch, device=x.device).repeat_interleave(np_)
    rho = torch.zeros(batch * nw ** 3, dtype=x.dtype, device=x.device)
    for k3 in range(2):
        for k2 in range(2):
            for k1 in range(2):
                idx = ((b * nw + (i[:, 2] + k3) % nw) * nw + (i[:, 1] + k2) % nw) * nw + (i[:, 0] + k1) % nw
                rho = rho.index_add(0, idx, w[:, 0, k1] * w[:, 1, k2] * w[:, 2, k3])
    return rho.view(batch, nw, nw, nw)


def kdk(x, v, G, batch, n, kick1, drift, kick2):
    a = force(x, G, batch, n)
    for k1, d, k2 in zip(kick1, drift, kick2):
        v = v + k1 * a
        x = x + d * v
        a = force(x, G, batch, n)
        v = v + k2 * a
    return x, v
