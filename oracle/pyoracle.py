"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  Functions mirror
oracle/cube_oracle.cpp, which restates the reference hot path (see that
file's header for the per-function reference citations).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "lib", "liboracle.so")


class OParams(C.Structure):
    _fields_ = [
        ("ng", C.c_int32), ("nnt", C.c_int32), ("nns", C.c_int32), ("ratio_cs", C.c_int32), ("ncb", C.c_int32),
        ("ngb", C.c_int32), ("izipx", C.c_int32), ("app", C.c_float), ("apm1c", C.c_float), ("apm2", C.c_float),
        ("apm3f", C.c_float), ("pm3", C.c_int32), ("pp", C.c_int32), ("pp_corr", C.c_int32), ("f_nu", C.c_double),
        ("mass_nu", C.c_double), ("tile_buffer", C.c_float), ("image_buffer", C.c_float),
        ("ires_fixed", C.c_int32),
    ]


class OKickIn(C.Structure):
    _fields_ = [("a_mid", C.c_double), ("dt", C.c_double), ("mass_p_cdm", C.c_float), ("phase_mask", C.c_int32)]


class OKickOut(C.Structure):
    _fields_ = [
        ("dt_pm1", C.c_float), ("dt_pm2", C.c_float), ("dt_pm3", C.c_float), ("dt_pp", C.c_float),
        ("dt_vmax", C.c_float), ("vz_max", C.c_float), ("f2max_pm1", C.c_float), ("f2max_pm2", C.c_float),
        ("f2max_pm3", C.c_float), ("f2max_pp", C.c_float), ("vmax", C.c_float * 3), ("napm", C.c_int32 * 7),
        ("testrho", C.c_double),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("vmax", "napm")}
        d["vmax"] = list(self.vmax)
        d["napm"] = list(self.napm)
        return d


class ODriftIn(C.Structure):
    _fields_ = [("dt_old", C.c_double), ("dt", C.c_double), ("vz_max", C.c_float)]


class ODriftOut(C.Structure):
    _fields_ = [("nplocal", C.c_int64), ("overhead_tile", C.c_float), ("nlayer", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB} (make -C oracle)")
        _lib = C.CDLL(LIB)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params(ng, izipx=2, nnt=4, nns=4, ratio_cs=4, ncb=4, ngb=16, app=0.06, apm1c=3.5, apm2=3.5, apm3f=3.5, pm3=1,
           pp=1, pp_corr=1, f_nu=0.0, mass_nu=0.0, tile_buffer=6.0, image_buffer=2.0, ires_fixed=0):
    return OParams(ng, nnt, nns, ratio_cs, ncb, ngb, izipx, app, apm1c, apm2, apm3f, pm3, pp, pp_corr, f_nu,
                   mass_nu, tile_buffer, image_buffer, ires_fixed)


def green(n, apm_soft, apm_range, gridsize):
    out = np.zeros((n, n, n // 2 + 1), np.float32)
    lib().orc_green(C.c_int64(n), C.c_float(apm_soft), C.c_float(apm_range), C.c_float(gridsize), _p(out))
    return out


def r2c3d(x):
    n = x.shape[0]
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((n, n, n // 2 + 1), np.complex128)
    lib().orc_r2c3d(C.c_int(n), _p(x), _p(out))
    return out


def c2r3d(c, n):
    c = np.ascontiguousarray(c, np.complex128)
    out = np.zeros((n, n, n), np.float64)
    lib().orc_c2r3d(C.c_int(n), _p(c), _p(out))
    return out


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, np.complex128).copy()
    lib().orc_fft1d(C.c_int(x.size), _p(x), C.c_int(sign))
    return x


def decode(xp, izipx):
    """Cell fractions of compressed coordinates xp (..., 3) int8/int16."""
    xp = np.ascontiguousarray(xp)
    out = np.zeros(xp.shape, np.float64)
    lib().orc_decode(C.c_int(izipx), C.c_int64(xp.size // 3), _p(xp), _p(out))
    return out


def tsc(xpos):
    """(floor index int64, weights float32 (n, 3)) of FP64 positions."""
    xpos = np.ascontiguousarray(xpos, np.float64)
    idx = np.zeros(xpos.size, np.int64)
    w = np.zeros((xpos.size, 3), np.float32)
    lib().orc_tsc(C.c_int64(xpos.size), _p(xpos), _p(idx), _p(w))
    return idx, w


def fd_gradient(phi):
    """Force [k][j][i][d] and f2max of a periodic potential phi [k][j][i]."""
    n = phi.shape[0]
    phi = np.ascontiguousarray(phi, np.float32)
    force = np.zeros(phi.shape + (3,), np.float32)
    f2 = C.c_float(0)
    lib().orc_fd_gradient(C.c_int64(n), _p(phi), _p(force), C.byref(f2))
    return force, f2.value


def poisson(rho, tf):
    """Unnormalised periodic solve: c2r(r2c(rho) * tf) in the PM levels' rounding order."""
    n = rho.shape[0]
    rho = np.ascontiguousarray(rho, np.float32)
    tf = np.ascontiguousarray(tf, np.float32)
    phi = np.zeros_like(rho)
    lib().orc_poisson(C.c_int(n), _p(rho), _p(tf), _p(phi))
    return phi


def f_ra(r, apm):
    r = np.ascontiguousarray(r, np.float32)
    out = np.zeros_like(r)
    lib().orc_f_ra(C.c_int64(r.size), _p(r), C.c_float(apm), _p(out))
    return out


def classify(P, rhoc):
    ns3 = P.nnt ** 3 * P.nns ** 3
    ires = np.zeros(ns3, np.int32)
    napm = np.zeros(7, np.int32)
    lib().orc_classify(C.byref(P), _p(rhoc), _p(ires), _p(napm))
    return ires, napm


def tf_default(P, levels=(1, 2, 3, 4, 5, 6)):
    """tf1, tf2, tf3_2..tf3_8 as initialize.f90:248-334 builds them.  Levels
    not in `levels` (1 = tf1, 2 = tf2, 3..6 = tf3_2..tf3_8) are returned as
    zeros of the right shape, for kicks whose phase mask does not read them."""
    nc = P.ng // P.ratio_cs
    ngp = P.ng // P.nnt
    ngt = ngp + 2 * P.ngb
    sfs = (2, 4, 6, 8)
    nft = [(ngp // P.nns) * s + 2 * P.ratio_cs * s for s in sfs]

    def g(level, n, *a):
        return green(n, *a) if level in levels else np.zeros((n, n, n // 2 + 1), np.float32)

    tfs = [g(1, nc, P.apm1c, 0.0, float(P.ratio_cs)),
           g(2, ngt, P.apm2, np.float32(P.apm1c) * np.float32(P.ratio_cs), 1.0)]
    for lv, (s, n) in enumerate(zip(sfs, nft), start=3):
        tfs.append(g(lv, n, P.apm3f, np.float32(P.apm2) * np.float32(s), np.float32(1.0) / np.float32(s)))
    if P.mass_nu > 0:
        tfs = [(np.float64(1 - P.f_nu) * t.astype(np.float64)).astype(np.float32) for t in tfs]
    return tfs


def kick(P, xp, vp, rhoc, tfs, a_mid, dt, mass_p_cdm=1.0, phase_mask=0, want_rho1=False):
    """Full kick; vp updated in place.  Returns (outputs, rho1 or None, pp pairs)."""
    arr = (C.c_void_p * 6)(*[t.ctypes.data for t in tfs])
    kin = OKickIn(a_mid, dt, mass_p_cdm, phase_mask)
    out = OKickOut()
    nc = P.ng // P.ratio_cs
    rho1 = np.zeros((nc, nc, nc), np.float32) if want_rho1 else None
    npairs = C.c_int64()
    lib().orc_kick(C.byref(P), _p(xp), _p(vp), _p(rhoc), arr, C.byref(kin), C.byref(out), _p(rho1), C.byref(npairs))
    return out.as_dict(), rho1, npairs.value


def pp_box(P, xp, vp, rhoc, it, iapm, a_mid, dt, mass_p_cdm=1.0):
    kin = OKickIn(a_mid, dt, mass_p_cdm, 8)
    f2 = C.c_float()
    vm = (C.c_float * 3)()
    npairs = C.c_int64()
    lib().orc_pp_box(C.byref(P), _p(xp), _p(vp), _p(rhoc), C.c_int64(it), C.c_int(iapm), C.byref(kin), C.byref(f2),
                     vm, C.byref(npairs))
    return f2.value, list(vm), npairs.value


def spine_image(P, rhoc):
    nt = P.ng // P.ratio_cs // P.nnt
    nte = nt + 2 * P.ncb
    nnt = P.nnt
    b = lambda: np.zeros((nnt, nnt, nnt, nte, nte), np.int64)
    p = lambda: np.zeros((nnt, nnt, nnt, nt, nt), np.int64)
    res = {"idx_b_l": b(), "idx_b_r": b(), "ppl0": p(), "pplr": p(), "pprl": p(), "ppr0": p(), "ppl": p(), "ppr": p()}
    lib().orc_spine_image(C.byref(P), _p(rhoc), _p(res["idx_b_l"]), _p(res["idx_b_r"]), _p(res["ppl0"]),
                          _p(res["pplr"]), _p(res["pprl"]), _p(res["ppr0"]), _p(res["ppl"]), _p(res["ppr"]))
    return res


def drift(P, xp, vp, pid, rhoc, dt_old, dt, vz_max):
    """Returns (status, xp, vp, pid, rhoc, outputs) with physical-layout arrays."""
    n = xp.shape[0]
    xo = np.zeros_like(xp)
    vo = np.zeros_like(vp)
    po = np.zeros_like(pid)
    ro = np.zeros_like(rhoc)
    din = ODriftIn(dt_old, dt, vz_max)
    out = ODriftOut()
    st = lib().orc_drift(C.byref(P), _p(xp), _p(vp), _p(pid), _p(rhoc), C.byref(din), _p(xo), _p(vo), _p(po), _p(ro),
                         C.c_int64(n), C.byref(out))
    k = out.nplocal
    return st, xo[:k], vo[:k], po[:k], ro, {"nplocal": k, "overhead_tile": out.overhead_tile, "nlayer": out.nlayer}


def buffer(P, xp, vp, pid, rhoc_phys, cap):
    """Physical layout -> buffered layout at nn=1."""
    xo = np.zeros((cap, 3), xp.dtype)
    vo = np.zeros((cap, 3), np.float32)
    po = np.zeros(cap, np.int64)
    ro = np.zeros_like(rhoc_phys)
    n = C.c_int64()
    st = lib().orc_buffer(C.byref(P), _p(xp), _p(vp), _p(pid), _p(rhoc_phys), _p(xo), _p(vo), _p(po), _p(ro),
                          C.c_int64(cap), C.byref(n))
    if st != 0:
        raise RuntimeError(f"orc_buffer status {st}")
    k = n.value
    return xo[:k], vo[:k], po[:k], ro


def buffer_v(P, vp, rhoc_ext, pid=None):
    """buffer_v of an extended layout, in place on vp (and pid)."""
    st = lib().orc_buffer_v(C.byref(P), _p(vp), _p(pid), _p(rhoc_ext), C.c_int64(vp.shape[0]))
    if st != 0:
        raise RuntimeError(f"orc_buffer_v status {st}")


def cicpower(P, xp, rhoc_phys, nw, box, want_delta=False):
    """CIC P(k) of a physical layout: (xi (nbin+1, 10), mean, delta or None)."""
    nyq = nw // 2
    nbin = int(round(nyq * float(np.sqrt(np.float32(3.0)))))
    xi = np.zeros((nbin + 1, 10), np.float32)
    delta = np.zeros((nw, nw, nw), np.float32) if want_delta else None
    mean = C.c_double()
    lib().orc_cicpower(C.byref(P), _p(np.ascontiguousarray(xp)), _p(rhoc_phys), C.c_int32(nw), C.c_float(box),
                       _p(delta), C.byref(mean), _p(xi), C.c_int32(nbin))
    return xi, mean.value, delta


def num_threads():
    return lib().orc_num_threads()


def set_sample_stride(stride):
    """Time-only sample: kick PM2/PM3/PP and drift process tiles T with
    T % stride == 0 (1 = every tile, the default)."""
    lib().orc_set_sample_stride(C.c_int64(int(stride)))


def phase_seconds():
    """Seconds of the last kick's PM1, PM2, PM3, PP phases and the last drift
    (the reference's tcat slots 11, 12, 13, 14 and 2)."""
    t = np.zeros(5, np.float64)
    lib().orc_phase_seconds(_p(t))
    return t
