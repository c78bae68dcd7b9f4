"""Checkpoint files of a one-image (nn=1) CUBE run, read and written in the
reference's raw stream format so GPU runs restart from, and are compared
against, the reference's own files.

Files (checkpoint.f90:11-33, particle_initialization.f90:28-82; names from
output_name, basic_functions.f08:95-135)::

    <opath>image<i>/<z>_info_<i>.bin   sim_header (parameters.f90:183-210)
    <opath>image<i>/<z>_xp_<i>.bin     integer(izipx) xp(3, nplocal)
    <opath>image<i>/<z>_vp_<i>.bin     real(4) vp(3, nplocal)
    <opath>image<i>/<z>_np_<i>.bin     integer(4) rhoc(1:nt,1:nt,1:nt, nnt,nnt,nnt)
    <opath>image<i>/<z>_id_<i>.bin     integer(8) pid(nplocal)  (izipi = 8)

with <z> written as Fortran's f7.3, left-adjusted.  Arrays here are in the
layouts of cube_nu_amd.p3m: xp (nplocal, 3), vp (nplocal, 3) and rhoc
(nnt, nnt, nnt, nte, nte, nte) with zero buffer cells (the physical layout
the drift writes).
"""
from __future__ import annotations

import os

import numpy as np

from .p3m import Geometry

# sim_header, parameters.f90:183-210: 16 integer(8) then 24 real(4), no padding
SIM_HEADER = np.dtype([
    ("nplocal", "<i8"), ("npglobal", "<i8"), ("izipx", "<i8"), ("izipv", "<i8"), ("image", "<i8"), ("nn", "<i8"),
    ("nnt", "<i8"), ("nt", "<i8"), ("ncell", "<i8"), ("ncb", "<i8"), ("timestep", "<i8"), ("cur_checkpoint", "<i8"),
    ("cur_halofind", "<i8"), ("cur_powerpoint", "<i8"), ("calculate_PK", "<i8"), ("cic_iapm", "<i8"),
    ("a", "<f4"), ("t", "<f4"), ("tau", "<f4"), ("dt_pp", "<f4"), ("dt_pm3", "<f4"), ("dt_pm2", "<f4"),
    ("dt_pm1", "<f4"), ("dt_vmax", "<f4"), ("mass_p_cdm", "<f4"), ("m_nu", "<f4", (3,)), ("Mass_nu", "<f4"),
    ("box", "<f4"), ("h0", "<f4"), ("omega_m", "<f4"), ("omega_nu", "<f4"), ("omega_l", "<f4"), ("s8", "<f4"),
    ("vsim2phys", "<f4"), ("sigma_vres", "<f4"), ("sigma_vi", "<f4"), ("z_i", "<f4"), ("vz_max", "<f4"),
])
assert SIM_HEADER.itemsize == 16 * 8 + 24 * 4

IZIPV = 2  # parameters.f90:14


def output_name(opath: str, z: float, zipname: str, image: int = 1) -> str:
    """output_prefix()//zipname//output_suffix() (basic_functions.f08:95-135)."""
    return os.path.join(opath, f"image{image}", f"{z:7.3f}".strip() + f"_{zipname}_{image}.bin")


def new_header(geo: Geometry, nplocal: int, **values) -> np.ndarray:
    """A sim_header for this geometry; other fields from keyword values."""
    h = np.zeros((), SIM_HEADER)
    h["nplocal"] = h["npglobal"] = nplocal
    h["izipx"], h["izipv"], h["image"], h["nn"] = geo.izipx, IZIPV, 1, 1
    h["nnt"], h["nt"], h["ncell"], h["ncb"] = geo.nnt, geo.nt, geo.nc, geo.ncb
    for k, v in values.items():
        h[k] = v
    return h


def write_checkpoint(opath: str, z: float, header: np.ndarray, xp, vp, rhoc, pid, geo: Geometry,
                     image: int = 1) -> None:
    """checkpoint.f90:11-33: writes the physical particles and counts."""
    n = int(header["nplocal"])
    b, nt = geo.ncb, geo.nt
    os.makedirs(os.path.join(opath, f"image{image}"), exist_ok=True)
    np.asarray(header, SIM_HEADER).tofile(output_name(opath, z, "info", image))
    np.ascontiguousarray(xp[:n], geo.xp_dtype).tofile(output_name(opath, z, "xp", image))
    np.ascontiguousarray(vp[:n], np.float32).tofile(output_name(opath, z, "vp", image))
    np.ascontiguousarray(rhoc[..., b:b + nt, b:b + nt, b:b + nt], np.int32).tofile(output_name(opath, z, "np", image))
    np.ascontiguousarray(pid[:n], np.int64).tofile(output_name(opath, z, "id", image))


def read_checkpoint(opath: str, z: float, geo: Geometry, image: int = 1):
    """particle_initialization.f90:28-82.  Returns (header, xp, vp, rhoc,
    pid); raises ValueError for the reference's 'zip format incompatable'
    and for counts that do not add up to nplocal."""
    header = np.fromfile(output_name(opath, z, "info", image), SIM_HEADER, count=1)[0]
    if int(header["izipx"]) != geo.izipx or int(header["izipv"]) != IZIPV:
        raise ValueError("zip format incompatable")  # the reference's message, particle_initialization.f90:38-43
    n = int(header["nplocal"])
    xp = np.fromfile(output_name(opath, z, "xp", image), geo.xp_dtype, count=3 * n).reshape(n, 3)
    vp = np.fromfile(output_name(opath, z, "vp", image), np.float32, count=3 * n).reshape(n, 3)
    b, nt = geo.ncb, geo.nt
    phys = np.fromfile(output_name(opath, z, "np", image), np.int32).reshape(geo.nnt, geo.nnt, geo.nnt, nt, nt, nt)
    rhoc = np.zeros(geo.rhoc_shape, np.int32)
    rhoc[..., b:b + nt, b:b + nt, b:b + nt] = phys
    pid = np.fromfile(output_name(opath, z, "id", image), np.int64, count=n)
    if int(phys.sum()) != n:
        raise ValueError(f"checkpoint counts sum to {int(phys.sum())}, header nplocal is {n}")
    return header, xp, vp, rhoc, pid
