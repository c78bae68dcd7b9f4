"""Checkpoint stream files (checkpoint.f90:11-33, particle_initialization.f90:
28-82): the sim_header layout is checked against a header written by Fortran
(flang, component-wise stream I/O as the reference writes it), and the
particle files round-trip."""
import os
import subprocess

import numpy as np
import pytest

from cube_nu_amd.checkpoint import SIM_HEADER, new_header, output_name, read_checkpoint, write_checkpoint
from cube_nu_amd.p3m import Geometry
from inputs import physical_set

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cube_nu_amd", "lib", "cube_p3m_driver")


def test_sim_header_layout_matches_fortran_stream_write(tmp_path):
    f = tmp_path / "hdr.bin"
    subprocess.run([DRIVER, "--write-header", str(f)], check=True, capture_output=True)
    assert f.stat().st_size == SIM_HEADER.itemsize == 224
    h = np.fromfile(f, SIM_HEADER)[0]
    assert h["nplocal"] == 1234567 and h["nt"] == 8 and h["ncell"] == 32 and h["timestep"] == 17
    assert h["cur_checkpoint"] == 3 and h["cic_iapm"] == 0
    assert np.float32(h["a"]) == np.float32(0.25) and np.float32(h["box"]) == np.float32(600.0)
    assert np.allclose(h["m_nu"], 0.02) and np.float32(h["Mass_nu"]) == np.float32(0.06)
    assert np.float32(h["z_i"]) == np.float32(49.0) and np.float32(h["vz_max"]) == np.float32(4.3)


def test_output_name_is_the_fortran_f7_3_name():
    assert output_name("/o/", 2.0, "xp") == "/o/image1/2.000_xp_1.bin"
    assert output_name("/o/", 49.0, "info", 3) == "/o/image3/49.000_info_3.bin"


def test_checkpoint_round_trip(tmp_path):
    geo = Geometry(64)
    xp, vp, pid, rhoc = physical_set(geo, seed=2, mean=3.0)
    h = new_header(geo, xp.shape[0], a=np.float32(1 / 3.0), box=600.0, vz_max=2.5)
    write_checkpoint(str(tmp_path) + "/", 2.0, h, xp, vp, rhoc, pid, geo)
    h2, x2, v2, r2, p2 = read_checkpoint(str(tmp_path) + "/", 2.0, geo)
    assert h2.tobytes() == np.asarray(h, SIM_HEADER).tobytes()
    assert np.array_equal(x2, xp) and np.array_equal(v2, vp) and np.array_equal(r2, rhoc) and np.array_equal(p2, pid)
    # the np file holds only the physical counts, nt^3 per tile
    assert os.path.getsize(output_name(str(tmp_path) + "/", 2.0, "np")) == 4 * geo.nt ** 3 * geo.nnt ** 3


def test_checkpoint_rejects_other_zip_format(tmp_path):
    geo = Geometry(64)
    xp, vp, pid, rhoc = physical_set(geo, seed=2, mean=1.0)
    h = new_header(geo, xp.shape[0])
    write_checkpoint(str(tmp_path) + "/", 1.0, h, xp, vp, rhoc, pid, geo)
    with pytest.raises(ValueError, match="zip format"):
        read_checkpoint(str(tmp_path) + "/", 1.0, Geometry(64, izipx=1))


@pytest.mark.gpu
def test_gpu_state_checkpoint_restart(tmp_path):
    """A drifted GPU state written as a checkpoint and read back restarts
    bit-identically (the checkpoint branch of main.f90:39-48)."""
    from cube_nu_amd import P3M

    p = P3M(64)
    ic = p.make_ic(seed=5, z=9.0, nic=2)
    d = p.drift_resident(dt_old=0.0, dt=0.1, vz_max=ic["vz_max"])
    xp, vp, pid, rhoc = p.download()
    n = d["nplocal"]
    geo = p.geo
    write_checkpoint(str(tmp_path) + "/", 9.0, new_header(geo, n), xp, vp, rhoc, pid, geo)
    h, x2, v2, r2, p2 = read_checkpoint(str(tmp_path) + "/", 9.0, geo)
    p.close()
    assert np.array_equal(x2, xp[:n]) and np.array_equal(v2, vp[:n]) and np.array_equal(p2, pid[:n])
    assert np.array_equal(r2, rhoc)
