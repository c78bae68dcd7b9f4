"""Host-side mirror of CUBE's P3M call surface over the cube_p3m C-ABI.

The reference exposes the hot path as argument-less Fortran subroutines
acting on module arrays (kick.f90:1, drift.f90:1, particle_mesh.f90:1).  This
module keeps the same names and argument meanings, taking the module arrays
explicitly as numpy arrays in the reference's column-major layouts:

* ``xp``   int16/int8 array of shape (np, 3)   == Fortran xp(3, np)
* ``vp``   float32 array of shape (np, 3)      == Fortran vp(3, np)
* ``pid``  int64 array of shape (np,)
* ``rhoc`` int32 array of shape (nnt, nnt, nnt, nte, nte, nte)
  == Fortran rhoc(1-ncb:nt+ncb, ..., nnt, nnt, nnt)

Errors map the reference's ``error stop`` sites (drift.f90:64-73, 87-93,
139-147; buffer_grid.f90:161-171) to :class:`CubeError` with the C status code.
The library is required: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcube_p3m.so")

# status codes (include/cube_p3m.h)
CUBE_OK = 0
CUBE_ERR_ARG = 1
CUBE_ERR_HIP = 2
CUBE_ERR_TILE_OVERFLOW = 3
CUBE_ERR_IMAGE_OVERFLOW = 4
CUBE_ERR_OUT_OF_RANGE = 5
CUBE_ERR_COUNT_MISMATCH = 6
CUBE_ERR_PID = 7
CUBE_ERR_SPINE = 8
CUBE_ERR_NO_DEVICE = 9
CUBE_ERR_STATE = 10

RATIO_SF = (1, 2, 4, 6, 8, 12, 16)  # parameters.f90:26


class CubeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cube_p3m error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("ng", C.c_int32), ("nnt", C.c_int32), ("nns", C.c_int32), ("ratio_cs", C.c_int32),
        ("ncb", C.c_int32), ("ngb", C.c_int32), ("izipx", C.c_int32), ("precision", C.c_int32),
        ("box", C.c_float), ("apm1c", C.c_float), ("apm2", C.c_float), ("apm3f", C.c_float),
        ("app", C.c_float), ("pm3", C.c_int32), ("pp", C.c_int32), ("pp_corr", C.c_int32),
        ("f_nu", C.c_double), ("mass_nu", C.c_double), ("image_buffer", C.c_float),
        ("tile_buffer", C.c_float), ("workspace_bytes", C.c_int64), ("ires_fixed", C.c_int32),
    ]


class KickIn(C.Structure):
    _fields_ = [("a_mid", C.c_double), ("dt", C.c_double), ("mass_p_cdm", C.c_float), ("phase_mask", C.c_int32)]


class KickOut(C.Structure):
    _fields_ = [
        ("dt_pm1", C.c_float), ("dt_pm2", C.c_float), ("dt_pm3", C.c_float), ("dt_pp", C.c_float),
        ("dt_vmax", C.c_float), ("vz_max", C.c_float), ("f2max_pm1", C.c_float), ("f2max_pm2", C.c_float),
        ("f2max_pm3", C.c_float), ("f2max_pp", C.c_float), ("vmax", C.c_float * 3), ("napm", C.c_int32 * 7),
        ("testrho", C.c_double),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("vmax", "napm")}
        d["vmax"] = list(self.vmax)
        d["napm"] = list(self.napm)
        return d


class DriftIn(C.Structure):
    _fields_ = [("dt_old", C.c_double), ("dt", C.c_double), ("vz_max", C.c_float)]


class DriftOut(C.Structure):
    _fields_ = [("nplocal", C.c_int64), ("overhead_tile", C.c_float), ("nlayer", C.c_int32)]


class ICParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("z", C.c_double), ("h", C.c_double), ("omega_m", C.c_double),
        ("omega_b", C.c_double), ("n_s", C.c_double), ("sigma8", C.c_double), ("nic", C.c_int32),
        ("box", C.c_float),
    ]


# Names the C-ABI exports (include/cube_p3m.h).
ABI_FUNCTIONS = (
    "cube_p3m_create", "cube_p3m_destroy", "cube_p3m_last_error", "cube_p3m_set_tf", "cube_p3m_get_tf",
    "cube_p3m_particle_mesh", "cube_p3m_kick", "cube_p3m_drift", "cube_p3m_buffer", "cube_p3m_buffer_v",
    "cube_p3m_buffer_v_resident", "cube_p3m_state_upload",
    "cube_p3m_state_download", "cube_p3m_kick_resident", "cube_p3m_drift_resident", "cube_p3m_buffer_resident",
    "cube_p3m_step_resident", "cube_p3m_make_ic", "cube_p3m_cicpower", "cube_p3m_synchronize",
    "cube_p3m_kick_resident_nu", "cube_p3m_add_thermal_velocities",
    "cube_p3m_graph_enable", "cube_p3m_timing_enable", "cube_p3m_timing_reset", "cube_p3m_timing_report",
    "cube_p3m_version",
)

_lib = None


def torch_runtime_first() -> None:
    """The PyTorch-ROCm wheel carries its own HIP runtime; when both runtimes
    live in one process, torch's must initialise the device first or
    torch.cuda stays unavailable.  Called before the library's first device
    call whenever torch has been imported (no effect otherwise)."""
    t = sys.modules.get("torch")
    if t is not None:
        try:
            t.cuda.init()
        except Exception:
            pass


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"cube_p3m library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    P = C.c_void_p
    vp_ = C.c_void_p
    i64 = C.c_int64
    lib.cube_p3m_create.argtypes = [C.POINTER(Config), C.POINTER(P)]
    lib.cube_p3m_destroy.argtypes = [P]
    lib.cube_p3m_last_error.argtypes = [P]
    lib.cube_p3m_last_error.restype = C.c_char_p
    lib.cube_p3m_version.restype = C.c_char_p
    lib.cube_p3m_set_tf.argtypes = [P, C.c_int32, vp_, i64]
    lib.cube_p3m_get_tf.argtypes = [P, C.c_int32, vp_, i64]
    lib.cube_p3m_particle_mesh.argtypes = [P, vp_, vp_, vp_, vp_, i64, C.POINTER(KickIn), vp_, C.POINTER(KickOut)]
    lib.cube_p3m_kick.argtypes = [P, vp_, vp_, vp_, vp_, i64, C.POINTER(KickIn), C.POINTER(KickOut)]
    lib.cube_p3m_drift.argtypes = [P, vp_, vp_, vp_, vp_, i64] + [vp_] * 8 + [C.POINTER(DriftIn), C.POINTER(DriftOut)]
    lib.cube_p3m_buffer.argtypes = [P, vp_, vp_, vp_, vp_, i64, i64, vp_, vp_, C.POINTER(C.c_int64)]
    lib.cube_p3m_buffer_v.argtypes = [P, vp_, vp_, vp_, i64]
    lib.cube_p3m_buffer_v_resident.argtypes = [P]
    lib.cube_p3m_state_upload.argtypes = [P, vp_, vp_, vp_, vp_, i64]
    lib.cube_p3m_state_download.argtypes = [P, vp_, vp_, vp_, vp_, i64, C.POINTER(C.c_int64)]
    lib.cube_p3m_kick_resident.argtypes = [P, C.POINTER(KickIn), C.POINTER(KickOut)]
    lib.cube_p3m_drift_resident.argtypes = [P, C.POINTER(DriftIn), C.POINTER(DriftOut)]
    lib.cube_p3m_buffer_resident.argtypes = [P, C.POINTER(C.c_int64)]
    lib.cube_p3m_step_resident.argtypes = [P, C.POINTER(DriftIn), C.POINTER(KickIn), C.POINTER(DriftOut),
                                           C.POINTER(KickOut)]
    lib.cube_p3m_make_ic.argtypes = [P, C.POINTER(ICParams), C.POINTER(C.c_float), C.POINTER(C.c_int64)]
    lib.cube_p3m_cicpower.argtypes = [P, C.c_int32, vp_, C.c_int32, vp_, C.POINTER(C.c_double)]
    lib.cube_p3m_synchronize.argtypes = [P]
    lib.cube_p3m_kick_resident_nu.argtypes = [P, P, C.POINTER(KickIn), C.c_float, C.POINTER(KickOut),
                                              C.POINTER(KickOut)]
    lib.cube_p3m_add_thermal_velocities.argtypes = [P, C.c_double, C.c_uint64, C.POINTER(C.c_float)]
    lib.cube_p3m_graph_enable.argtypes = [P, C.c_int32]
    lib.cube_p3m_timing_enable.argtypes = [P, C.c_int32]
    lib.cube_p3m_timing_reset.argtypes = [P]
    lib.cube_p3m_timing_report.argtypes = [P, C.c_char_p, i64]
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Geometry:
    """Derived sizes of one image (parameters.f90:57-82, variables.f90:24-26)."""

    ng: int
    nnt: int = 4
    nns: int = 4
    ratio_cs: int = 4
    ncb: int = 4
    ngb: int = 16
    izipx: int = 2
    image_buffer: float = 2.0
    tile_buffer: float = 6.0

    @property
    def nc(self):
        return self.ng // self.ratio_cs

    @property
    def nt(self):
        return self.nc // self.nnt

    @property
    def ntt(self):
        return self.nt // self.nns

    @property
    def ngp(self):
        return self.ng // self.nnt

    @property
    def ngt(self):
        return self.ngp + 2 * self.ngb

    @property
    def nte(self):
        return self.nt + 2 * self.ncb

    def nfp(self, iapm):
        return (self.ngp // self.nns) * RATIO_SF[iapm - 1]

    def nfb(self, iapm):
        return self.ratio_cs * RATIO_SF[iapm - 1]

    def nft(self, iapm):
        return self.nfp(iapm) + 2 * self.nfb(iapm)

    @property
    def rhoc_shape(self):
        return (self.nnt, self.nnt, self.nnt, self.nte, self.nte, self.nte)

    @property
    def xp_dtype(self):
        return np.int16 if self.izipx == 2 else np.int8

    @property
    def np_image(self):
        return self.ng ** 3

    @property
    def np_image_max(self):
        r = np.float32(self.nte) / np.float32(self.nt)
        return int(np.float32(self.np_image) * (r * r * r) * np.float32(self.image_buffer))

    def tf_size(self, level):
        n = {1: self.nc, 2: self.ngt, 3: self.nft(2), 4: self.nft(3), 5: self.nft(4), 6: self.nft(5)}[level]
        return n, (n // 2 + 1) * n * n


class P3M:
    """One image (nn=1) of the P3M step on the current HIP device."""

    def __init__(self, ng: int, izipx: int = 2, precision: int = 32, *, nnt: int = 4, nns: int = 4,
                 ratio_cs: int = 4, ncb: int = 4, ngb: int = 16, box: float = 600.0, apm1c: float = 3.5,
                 apm2: float = 3.5, apm3f: float = 3.5, app: float = 0.06, pm3: bool = True, pp: bool = True,
                 pp_corr: bool = True, f_nu: float = 0.0, mass_nu: float = 0.0, image_buffer: float = 2.0,
                 tile_buffer: float = 6.0, workspace_bytes: int = 0, ires_fixed: int = 0):
        self.lib = load_library()
        self.geo = Geometry(ng, nnt, nns, ratio_cs, ncb, ngb, izipx, image_buffer, tile_buffer)
        cfg = Config(ng, nnt, nns, ratio_cs, ncb, ngb, izipx, precision, box, apm1c, apm2, apm3f, app, int(pm3),
                     int(pp), int(pp_corr), f_nu, mass_nu, image_buffer, tile_buffer, workspace_bytes,
                     int(ires_fixed))
        self.cfg = cfg
        ctx = C.c_void_p()
        torch_runtime_first()
        rc = self.lib.cube_p3m_create(C.byref(cfg), C.byref(ctx))
        if rc != CUBE_OK:
            raise CubeError(rc, "cube_p3m_create failed")
        self.ctx = ctx

    # -- helpers ----------------------------------------------------------
    def _check(self, rc):
        if rc != CUBE_OK:
            raise CubeError(rc, self.lib.cube_p3m_last_error(self.ctx).decode())

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.cube_p3m_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk_arrays(self, xp, vp, rhoc, pid=None):
        g = self.geo
        assert xp.dtype == g.xp_dtype and xp.flags.c_contiguous and xp.shape[1] == 3
        assert vp.dtype == np.float32 and vp.flags.c_contiguous and vp.shape == xp.shape
        assert rhoc.dtype == np.int32 and rhoc.shape == g.rhoc_shape and rhoc.flags.c_contiguous
        if pid is not None:
            assert pid.dtype == np.int64 and pid.shape == (xp.shape[0],)

    # -- transfer functions (variables.f90:21-22) -------------------------
    def get_tf(self, level: int) -> np.ndarray:
        n, m = self.geo.tf_size(level)
        out = np.empty(m, np.float32)
        self._check(self.lib.cube_p3m_get_tf(self.ctx, level, _ptr(out), m))
        return out.reshape(n, n, n // 2 + 1)

    def set_tf(self, level: int, tf: np.ndarray):
        n, m = self.geo.tf_size(level)
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1)
        self._check(self.lib.cube_p3m_set_tf(self.ctx, level, _ptr(tf), tf.size))

    # -- the reference call surface --------------------------------------
    def particle_mesh(self, xp, vp, rhoc, *, a_mid, dt, mass_p_cdm=1.0, idx_b_r=None):
        """PM1 (particle_mesh.f90:1-222): returns (rho1, outputs); vp updated in place."""
        self._chk_arrays(xp, vp, rhoc)
        nc = self.geo.nc
        dens = np.empty((nc, nc, nc), np.float32)
        kin = KickIn(a_mid, dt, mass_p_cdm, 1)
        out = KickOut()
        self._check(self.lib.cube_p3m_particle_mesh(self.ctx, _ptr(xp), _ptr(vp), _ptr(rhoc), _ptr(idx_b_r),
                                                     xp.shape[0], C.byref(kin), _ptr(dens), C.byref(out)))
        return dens, out.as_dict()

    def kick(self, xp, vp, rhoc, *, a_mid, dt, mass_p_cdm=1.0, phase_mask=0, idx_b_r=None):
        """kick (kick.f90:1-353): vp updated in place; returns the sim outputs."""
        self._chk_arrays(xp, vp, rhoc)
        kin = KickIn(a_mid, dt, mass_p_cdm, phase_mask)
        out = KickOut()
        self._check(self.lib.cube_p3m_kick(self.ctx, _ptr(xp), _ptr(vp), _ptr(rhoc), _ptr(idx_b_r), xp.shape[0],
                                           C.byref(kin), C.byref(out)))
        return out.as_dict()

    def drift(self, xp, vp, pid, rhoc, *, dt_old, dt, vz_max, spine=False):
        """drift (drift.f90:1-233): arrays rewritten in place to the physical
        layout; returns nplocal etc. (and the spine arrays if spine=True)."""
        self._chk_arrays(xp, vp, rhoc, pid)
        g = self.geo
        sp = None
        ptrs = [None] * 8
        if spine:
            shp_b = (g.nnt, g.nnt, g.nnt, g.nte, g.nte)
            shp_p = (g.nnt, g.nnt, g.nnt, g.nt, g.nt)
            sp = {k: np.zeros(shp_b if k.startswith("idx") else shp_p, np.int64)
                  for k in ("idx_b_l", "idx_b_r", "ppl0", "pplr", "pprl", "ppr0", "ppl", "ppr")}
            ptrs = [_ptr(sp[k]) for k in ("idx_b_l", "idx_b_r", "ppl0", "pplr", "pprl", "ppr0", "ppl", "ppr")]
        din = DriftIn(dt_old, dt, vz_max)
        out = DriftOut()
        self._check(self.lib.cube_p3m_drift(self.ctx, _ptr(xp), _ptr(vp), _ptr(pid), _ptr(rhoc), xp.shape[0],
                                            *ptrs, C.byref(din), C.byref(out)))
        res = {"nplocal": out.nplocal, "overhead_tile": out.overhead_tile, "nlayer": out.nlayer}
        if sp is not None:
            res["spine"] = sp
        return res

    def buffer(self, xp, vp, pid, rhoc, nplocal):
        """buffer_grid + buffer_x + buffer_v at nn=1; arrays rewritten in place
        to the buffered layout; returns the extended particle count."""
        self._chk_arrays(xp, vp, rhoc, pid)
        n = C.c_int64()
        self._check(self.lib.cube_p3m_buffer(self.ctx, _ptr(xp), _ptr(vp), _ptr(pid), _ptr(rhoc), nplocal,
                                             xp.shape[0], None, None, C.byref(n)))
        return n.value

    # -- resident mode ----------------------------------------------------
    def upload(self, xp, vp, pid, rhoc):
        self._chk_arrays(xp, vp, rhoc, pid)
        self._check(self.lib.cube_p3m_state_upload(self.ctx, _ptr(xp), _ptr(vp), _ptr(pid), _ptr(rhoc),
                                                   xp.shape[0]))

    def download(self, cap=None):
        g = self.geo
        n = C.c_int64()
        cap = cap or g.np_image_max
        xp = np.zeros((cap, 3), g.xp_dtype)
        vp = np.zeros((cap, 3), np.float32)
        pid = np.zeros(cap, np.int64)
        rhoc = np.zeros(g.rhoc_shape, np.int32)
        self._check(self.lib.cube_p3m_state_download(self.ctx, _ptr(xp), _ptr(vp), _ptr(pid), _ptr(rhoc), cap,
                                                     C.byref(n)))
        k = n.value
        return xp[:k].copy(), vp[:k].copy(), pid[:k].copy(), rhoc

    def kick_resident(self, *, a_mid, dt, mass_p_cdm=1.0, phase_mask=0):
        kin = KickIn(a_mid, dt, mass_p_cdm, phase_mask)
        out = KickOut()
        self._check(self.lib.cube_p3m_kick_resident(self.ctx, C.byref(kin), C.byref(out)))
        return out.as_dict()

    def kick_resident_nu(self, nu: "P3M", *, a_mid, dt, nu_mass, mass_p_cdm=1.0, phase_mask=0):
        """Kick with the PM-only second species resident in `nu`
        (cube_p3m_kick_resident_nu); returns (outputs, second species' outputs)."""
        kin = KickIn(a_mid, dt, mass_p_cdm, phase_mask)
        out, out_nu = KickOut(), KickOut()
        self._check(self.lib.cube_p3m_kick_resident_nu(self.ctx, nu.ctx, C.byref(kin), float(nu_mass),
                                                       C.byref(out), C.byref(out_nu)))
        return out.as_dict(), out_nu.as_dict()

    def add_thermal_velocities(self, v_scale: float, seed: int) -> float:
        """Fermi-Dirac thermal velocities (cube_p3m_add_thermal_velocities);
        returns the largest |v_z| afterwards."""
        vz = C.c_float()
        self._check(self.lib.cube_p3m_add_thermal_velocities(self.ctx, float(v_scale), int(seed) & (2 ** 64 - 1),
                                                             C.byref(vz)))
        return vz.value

    def drift_resident(self, *, dt_old, dt, vz_max):
        din = DriftIn(dt_old, dt, vz_max)
        out = DriftOut()
        self._check(self.lib.cube_p3m_drift_resident(self.ctx, C.byref(din), C.byref(out)))
        return {"nplocal": out.nplocal, "overhead_tile": out.overhead_tile, "nlayer": out.nlayer}

    def buffer_v(self, vp, rhoc, pid=None):
        """buffer_v (buffer_particle.f90:11-27): buffer velocities (and IDs)
        of the extended layout from their physical cells, in place."""
        assert vp.dtype == np.float32 and vp.flags.c_contiguous and vp.shape[1] == 3
        assert rhoc.dtype == np.int32 and rhoc.shape == self.geo.rhoc_shape and rhoc.flags.c_contiguous
        if pid is not None:
            assert pid.dtype == np.int64 and pid.shape == (vp.shape[0],) and pid.flags.c_contiguous
        self._check(self.lib.cube_p3m_buffer_v(self.ctx, _ptr(vp), _ptr(pid), _ptr(rhoc), vp.shape[0]))

    def buffer_v_resident(self):
        self._check(self.lib.cube_p3m_buffer_v_resident(self.ctx))

    def buffer_resident(self):
        n = C.c_int64()
        self._check(self.lib.cube_p3m_buffer_resident(self.ctx, C.byref(n)))
        return n.value

    def step_resident(self, *, dt_old, dt, vz_max, a_mid, mass_p_cdm=1.0):
        din = DriftIn(dt_old, dt, vz_max)
        kin = KickIn(a_mid, dt, mass_p_cdm, 0)
        dout, kout = DriftOut(), KickOut()
        self._check(self.lib.cube_p3m_step_resident(self.ctx, C.byref(din), C.byref(kin), C.byref(dout),
                                                    C.byref(kout)))
        return {"nplocal": dout.nplocal, "nlayer": dout.nlayer}, kout.as_dict()

    def make_ic(self, *, seed=1, z=2.0, h=0.6817, omega_m=0.301, omega_b=0.048, n_s=0.9693, sigma8=0.8, nic=2,
                box=None):
        p = ICParams(seed, z, h, omega_m, omega_b, n_s, sigma8, nic, box if box is not None else self.cfg.box)
        vz = C.c_float()
        n = C.c_int64()
        self._check(self.lib.cube_p3m_make_ic(self.ctx, C.byref(p), C.byref(vz), C.byref(n)))
        return {"vz_max": vz.value, "nplocal": n.value}

    def cicpower(self, nw: int, want_delta: bool = False):
        """CIC P(k) of the resident physical particles on an nw^3 mesh
        (utilities/cicpower.f90 + powerspectrum.f90 auto_power).  Returns
        (xi (nbin+1, 10) float32, mean, delta (nw,nw,nw) or None)."""
        nbin = int(round((nw // 2) * float(np.sqrt(np.float32(3.0)))))
        xi = np.zeros((nbin + 1, 10), np.float32)
        delta = np.zeros((nw, nw, nw), np.float32) if want_delta else None
        mean = C.c_double()
        self._check(self.lib.cube_p3m_cicpower(self.ctx, nw, xi.ctypes.data_as(C.c_void_p), nbin,
                                               None if delta is None else delta.ctypes.data_as(C.c_void_p),
                                               C.byref(mean)))
        return xi, mean.value, delta

    def synchronize(self):
        self._check(self.lib.cube_p3m_synchronize(self.ctx))

    def graph(self, enable: bool = True):
        """Graph mode for full resident kicks (cube_p3m_graph_enable)."""
        self._check(self.lib.cube_p3m_graph_enable(self.ctx, int(enable)))

    def timing(self, enable: bool = True):
        self._check(self.lib.cube_p3m_timing_enable(self.ctx, int(enable)))

    def timing_reset(self):
        self._check(self.lib.cube_p3m_timing_reset(self.ctx))

    def timing_report(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.cube_p3m_timing_report(self.ctx, buf, len(buf)))
        rep = {}
        for line in buf.value.decode().splitlines():
            name, calls, ms, by, fl = line.split()
            rep[name] = {"calls": int(calls), "ms": float(ms), "bytes": float(by), "flops": float(fl)}
        return rep
