"""Differentiable PM-only gravity over the C-ABI of include/cube_pm.h.

The reference has no differentiable path; this module provides the PM-only
forward+adjoint that the P3M operator's mathematics admits (PM1's TSC
deposit, Green_3D convolution, 4-point FD gradient and TSC interpolation,
particle_mesh.f90:36-205, Green.f90:22-71 with apm_range = 0):

* ``PMForce`` -- one context per (mesh size, batch): force, its
  vector-Jacobian product, CIC density and its VJP, all as HIP kernels that
  work on torch CUDA tensors in place (torch supplies device memory and the
  stream; the arithmetic is the library's).
* ``pm_force`` / ``cic_density`` -- ``torch.autograd.Function`` wrappers whose
  backward calls the hand-written adjoint kernels.
* ``kdk`` -- a multi-step kick-drift-kick PM forward as one
  ``torch.autograd.Function``: the forward keeps only each step's particle
  positions (per-step checkpoints of the particle state); the backward walks
  the steps in reverse, recomputing each force mesh from the checkpointed
  positions before its adjoint, so no mesh outlives its step.
* ``zeldovich`` -- a differentiable (torch) Zel'dovich initial condition from
  white-noise modes, so gradients reach the noise.  With ``displacement=True``
  it returns the displacements from the particle lattice, the form a context
  created with ``lattice=ng`` takes (its floor indices and fractions then
  carry the rounding of the displacement rather than of the absolute
  coordinate).

No CPU fallback: every function raises if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .p3m import CubeError, load_library

PM_ABI_FUNCTIONS = (
    "cube_pm_create", "cube_pm_destroy", "cube_pm_last_error", "cube_pm_green", "cube_pm_force",
    "cube_pm_force_vjp", "cube_pm_cic", "cube_pm_cic_vjp", "cube_pm_cic_noalloc",
)


class PMConfig(C.Structure):
    _fields_ = [("n", C.c_int32), ("batch", C.c_int32), ("apm_soft", C.c_float), ("lattice", C.c_int32)]


def _lib():
    lib = load_library()
    if getattr(lib, "_pm_bound", False):
        return lib
    P, vp, i64 = C.c_void_p, C.c_void_p, C.c_int64
    lib.cube_pm_create.argtypes = [C.POINTER(PMConfig), C.POINTER(P)]
    lib.cube_pm_destroy.argtypes = [P]
    lib.cube_pm_last_error.argtypes = [P]
    lib.cube_pm_last_error.restype = C.c_char_p
    lib.cube_pm_green.argtypes = [P, vp]
    lib.cube_pm_force.argtypes = [P, vp, i64, vp, vp, vp]
    lib.cube_pm_force_vjp.argtypes = [P, vp, i64, vp, vp, vp, vp]
    lib.cube_pm_cic.argtypes = [P, vp, i64, C.c_int32, C.c_float, vp, vp]
    lib.cube_pm_cic_vjp.argtypes = [P, vp, i64, C.c_int32, C.c_float, vp, vp, vp]
    lib.cube_pm_cic_noalloc.argtypes = [P, vp, i64, C.c_int32, C.c_float, vp, vp]
    lib._pm_bound = True
    return lib


def _ptr(t: torch.Tensor) -> C.c_void_p:
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 CUDA tensor")
    return C.c_void_p(t.data_ptr())


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class PMForce:
    """PM force on ``batch`` periodic meshes of n^3 points (mesh units)."""

    def __init__(self, n: int, batch: int = 1, apm_soft: float = 1.0, lattice: int = 0):
        """lattice = L > 0: positions are displacements from an L^3 lattice
        of sites ((i + 1/2) n / L per axis), one per particle in x-fastest
        order (include/cube_pm.h)."""
        if not torch.cuda.is_available():
            raise RuntimeError("PMForce needs a GPU (the HIP library is the only implementation)")
        self.lib = _lib()
        self.n, self.batch, self.lattice = int(n), int(batch), int(lattice)
        self.ctx = C.c_void_p()
        rc = self.lib.cube_pm_create(C.byref(PMConfig(self.n, self.batch, float(apm_soft), self.lattice)),
                                     C.byref(self.ctx))
        if rc != 0:
            raise CubeError(rc, f"cube_pm_create failed ({rc})")

    def _check(self, rc):
        if rc != 0:
            raise CubeError(rc, self.lib.cube_pm_last_error(self.ctx).decode())

    def close(self):
        if self.ctx:
            torch.cuda.synchronize()
            self.lib.cube_pm_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def green(self) -> np.ndarray:
        """G on the half-complex grid, shape (n, n, n//2+1) as [kz][ky][kx]."""
        n = self.n
        g = np.empty((n, n, n // 2 + 1), np.float32)
        self._check(self.lib.cube_pm_green(self.ctx, g.ctypes.data_as(C.c_void_p)))
        return g

    def _np(self, x: torch.Tensor) -> int:
        if x.dim() != 2 or x.shape[1] != 3 or x.shape[0] % self.batch:
            raise ValueError("x must be (batch*np, 3)")
        return x.shape[0] // self.batch

    def mesh_tensor(self, device) -> torch.Tensor:
        return torch.empty((self.batch, 3, self.n, self.n, self.n), dtype=torch.float32, device=device)

    def force(self, x: torch.Tensor, F: torch.Tensor | None = None):
        """acc (same shape as x) and the force meshes F it came from."""
        np_ = self._np(x)
        F = self.mesh_tensor(x.device) if F is None else F
        acc = torch.empty_like(x)
        self._check(self.lib.cube_pm_force(self.ctx, _ptr(x), np_, _ptr(F), _ptr(acc), _stream()))
        return acc, F

    def force_vjp(self, x: torch.Tensor, F: torch.Tensor | None, g: torch.Tensor) -> torch.Tensor:
        """(d acc / d x)^T g; F = None recomputes the force meshes of x."""
        gx = torch.empty_like(x)
        self._check(self.lib.cube_pm_force_vjp(self.ctx, _ptr(x), self._np(x), None if F is None else _ptr(F),
                                               _ptr(g.contiguous()), _ptr(gx), _stream()))
        return gx

    def sites(self, dtype=torch.float64, device="cpu") -> torch.Tensor:
        """Lattice site positions (np, 3) in mesh units, x fastest (lattice mode)."""
        L = self.lattice
        if L <= 0:
            raise ValueError("not a lattice-displacement context")
        g = (torch.arange(L, dtype=dtype, device=device) + 0.5) * (self.n / L)
        QZ, QY, QX = torch.meshgrid(g, g, g, indexing="ij")
        return torch.stack([QX, QY, QZ], dim=-1).reshape(L ** 3, 3)

    def cic(self, x: torch.Tensor, nw: int, scale: float) -> torch.Tensor:
        rho = torch.empty((self.batch, nw, nw, nw), dtype=torch.float32, device=x.device)
        self._check(self.lib.cube_pm_cic(self.ctx, _ptr(x), self._np(x), int(nw), float(scale), _ptr(rho), _stream()))
        return rho

    def cic_vjp(self, x: torch.Tensor, nw: int, scale: float, grho: torch.Tensor) -> torch.Tensor:
        gx = torch.empty_like(x)
        self._check(self.lib.cube_pm_cic_vjp(self.ctx, _ptr(x), self._np(x), int(nw), float(scale),
                                             _ptr(grho.contiguous()), _ptr(gx), _stream()))
        return gx


class _ForceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, pm):
        x = x.contiguous()
        acc, F = pm.force(x)
        ctx.pm = pm
        ctx.save_for_backward(x, F)
        return acc

    @staticmethod
    def backward(ctx, g):
        x, F = ctx.saved_tensors
        return ctx.pm.force_vjp(x, F, g), None


class _CICFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, pm, nw, scale):
        x = x.contiguous()
        ctx.pm, ctx.nw, ctx.scale = pm, nw, scale
        ctx.save_for_backward(x)
        return pm.cic(x, nw, scale)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        return ctx.pm.cic_vjp(x, ctx.nw, ctx.scale, g), None, None, None


def pm_force(x: torch.Tensor, pm: PMForce) -> torch.Tensor:
    """Differentiable PM acceleration at positions x (batch*np, 3)."""
    return _ForceFn.apply(x, pm)


def cic_density(x: torch.Tensor, pm: PMForce, nw: int, scale: float) -> torch.Tensor:
    """Differentiable CIC density (batch, nw, nw, nw) of x*scale."""
    return _CICFn.apply(x, pm, nw, scale)


class _KDKFn(torch.autograd.Function):
    """Kick-drift-kick steps with per-step checkpoints of the positions.

    Forward: a_0 = F(x_0); for each step s: v' = v_s + k1_s a_s,
    x_{s+1} = x_s + d_s v', a_{s+1} = F(x_{s+1}), v_{s+1} = v' + k2_s a_{s+1}.
    Only x_0 .. x_S are kept.  Backward, with the cotangents (X, V) of
    (x_{s+1}, v_{s+1}) and A the part of a_{s+1}'s cotangent that step s+1
    left (k1_{s+1} V' of that step): A += k2_s V; X += J_F(x_{s+1})^T A
    (the force mesh of x_{s+1} recomputed first); V' = V + d_s X; then
    (X, V) become the cotangents of (x_s, v_s) and A = k1_s V'.  Finally
    X += J_F(x_0)^T A.
    """

    @staticmethod
    def forward(ctx, x, v, pm, kick1, drift, kick2):
        x = x.contiguous()
        keep = [x]
        a, _ = pm.force(x)
        for k1, d, k2 in zip(kick1, drift, kick2):
            v = v + k1 * a
            x = (x + d * v).contiguous()
            keep.append(x)
            a, _ = pm.force(x)
            v = v + k2 * a
        ctx.pm, ctx.coef = pm, (list(kick1), list(drift), list(kick2))
        ctx.save_for_backward(*keep)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        xs = ctx.saved_tensors
        kick1, drift, kick2 = ctx.coef
        pm = ctx.pm
        X = torch.zeros_like(xs[-1]) if gx is None else gx.clone()
        V = torch.zeros_like(xs[-1]) if gv is None else gv.clone()
        A = None
        for s in range(len(drift) - 1, -1, -1):
            A = kick2[s] * V if A is None else A + kick2[s] * V
            X = X + pm.force_vjp(xs[s + 1], None, A)
            V = V + drift[s] * X
            A = kick1[s] * V
        if A is not None:
            X = X + pm.force_vjp(xs[0], None, A)
        return X, V, None, None, None, None


def kdk(x: torch.Tensor, v: torch.Tensor, pm: PMForce, kick1, drift, kick2):
    """Kick-drift-kick steps v += k1 a(x); x += d v; v += k2 a(x), with the
    force at the end of a step reused by the next step's first kick.
    Differentiable with respect to x and v; the backward recomputes each
    step's force mesh from checkpointed positions (memory: one position array
    per step, no meshes)."""
    return _KDKFn.apply(x, v, pm, list(kick1), list(drift), list(kick2))


def kdk_coefficients(a0: float, a1: float, nsteps: int, omega_m: float = 0.301):
    """Steps uniform in a for the comoving equations dx/da = v/(a^3 E),
    dv/da = (3 Omega_m / 2) g / (a^2 E), E(a) = sqrt(Om a^-3 + 1 - Om)
    (g the mesh acceleration per unit density contrast); kick and drift
    factors integrated with the midpoint rule."""
    E = lambda a: math.sqrt(omega_m / a ** 3 + 1 - omega_m)
    edges = np.linspace(a0, a1, nsteps + 1)
    k1, d, k2 = [], [], []
    for i in range(nsteps):
        lo, hi = edges[i], edges[i + 1]
        mid = 0.5 * (lo + hi)
        q1, q3 = 0.5 * (lo + mid), 0.5 * (mid + hi)
        k1.append(1.5 * omega_m * (mid - lo) / (q1 * q1 * E(q1)))
        d.append((hi - lo) / (mid ** 3 * E(mid)))
        k2.append(1.5 * omega_m * (hi - mid) / (q3 * q3 * E(q3)))
    return k1, d, k2


def bbks_amplitude(k, h=0.6817, omega_m=0.301, n_s=0.9693):
    """sqrt(P(k)) shape of BBKS (unnormalised; k in h/Mpc)."""
    q = k / (omega_m * h)
    T = torch.log(1 + 2.34 * q) / (2.34 * q) * (1 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4) ** -0.25
    return torch.sqrt(k ** n_s) * T


def zeldovich(noise: torch.Tensor, box: float, n_mesh: int, amp: float, displacement: bool = False):
    """Positions (mesh units) and velocities of ng^3 particles per batch
    from white noise (batch, ng, ng, ng): delta_k = amp * sqrt(P(k)) *
    FFT(noise), displacement psi_k = i k/k^2 delta_k, x = q + psi, v = psi
    (growing mode, in units where the drift factor carries the time).
    displacement=True returns psi (mesh units) instead of x, for a context
    with lattice=ng (the sites q are (i + 1/2) n_mesh / ng, x fastest)."""
    B, ng = noise.shape[0], noise.shape[1]
    dk = torch.fft.rfftn(noise, dim=(1, 2, 3))
    kf = 2 * math.pi / box
    kx = torch.fft.rfftfreq(ng, d=1.0 / ng, device=noise.device) * kf
    ky = torch.fft.fftfreq(ng, d=1.0 / ng, device=noise.device) * kf
    KZ, KY, KX = torch.meshgrid(ky, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k = torch.sqrt(k2)
    k2 = torch.where(k2 > 0, k2, torch.ones_like(k2))
    A = torch.where(k > 0, bbks_amplitude(torch.where(k > 0, k, torch.ones_like(k))), torch.zeros_like(k)) * amp
    cell = box / ng
    disp = []
    for K in (KX, KY, KZ):
        psi_k = dk * (1j * K / k2) * A
        disp.append(torch.fft.irfftn(psi_k, s=(ng, ng, ng), dim=(1, 2, 3)) / cell)  # in particle-grid cells
    psi = torch.stack(disp, dim=-1)  # (B, z, y, x, 3) with components (x, y, z)
    g = torch.arange(ng, device=noise.device, dtype=noise.dtype) + 0.5
    QZ, QY, QX = torch.meshgrid(g, g, g, indexing="ij")
    q = torch.stack([QX, QY, QZ], dim=-1).expand(B, ng, ng, ng, 3)
    s = n_mesh / ng
    v = (psi * s).reshape(B * ng ** 3, 3).contiguous()
    if displacement:
        return v.clone(), v
    x = ((q + psi) * s).reshape(B * ng ** 3, 3).contiguous()
    return x, v
