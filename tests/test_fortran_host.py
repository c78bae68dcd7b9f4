"""A Fortran host (cube_nu_amd/fortran/cube_p3m_driver.f90, flang-built) calls
kick and drift through the ISO_C_BINDING interface on raw stream files; the
results are checked against the oracle: kick to rms relative 1e-5, drift and
the rebuilt spine arrays bit-exact."""
import os
import subprocess

import numpy as np
import pytest

from cube_nu_amd.p3m import Geometry
from inputs import buffered_set, oparams, rel_rms
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cube_nu_amd", "lib", "cube_p3m_driver")


def test_fortran_host_kick_and_drift(tmp_path):
    geo = Geometry(64)
    P = oparams(geo)
    xp, vp, pid, rhoc = buffered_set(geo, seed=17, mean=5.0)
    n = xp.shape[0]
    spine = orc.spine_image(P, rhoc)
    for name, arr in (("xp", xp), ("vp", vp), ("pid", pid), ("rhoc", rhoc), ("idx_b_r", spine["idx_b_r"])):
        np.ascontiguousarray(arr).tofile(tmp_path / f"{name}.bin")
    a_mid, dt, dt_old, dt_drift = 0.5, 0.3, 0.4, 0.6
    print(f"fortran driver: ng=64 np={n}", flush=True)
    subprocess.run([DRIVER, str(tmp_path), "64", str(n), str(a_mid), str(dt), str(dt_old), str(dt_drift)],
                   check=True, timeout=120)
    # kick
    vk = np.fromfile(tmp_path / "vp_kick.bin", np.float32).reshape(n, 3)
    vo = vp.copy()
    out_o, _, _ = orc.kick(P, xp, vo, rhoc, orc.tf_default(P), a_mid, dt)
    assert rel_rms(vk - vp, vo - vp) <= 1e-5
    txt = open(tmp_path / "kick.txt").read().split()
    dts = [float(v) for v in txt[:6]]
    for k, v in zip(("dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "dt_vmax", "vz_max"), dts):
        assert abs(v - out_o[k]) <= 1e-5 * abs(out_o[k]), k
    assert [int(v) for v in txt[6:13]] == out_o["napm"]
    # buffer_v, then the drift of the GPU-kicked velocities, bit-exact against
    # the oracle
    vz = np.float32(dts[5])
    vkb = vk.copy()
    orc.buffer_v(P, vkb, rhoc)
    # the driver holds the step values in real(c_float)
    st, xe, ve, pe, re, oo = orc.drift(P, xp, vkb, pid, rhoc, float(np.float32(dt_old)), float(np.float32(dt_drift)),
                                       float(vz))
    assert st == 0
    m = oo["nplocal"]
    nplocal, _, nlayer = open(tmp_path / "drift.txt").read().split()
    assert int(nplocal) == m and int(nlayer) == oo["nlayer"]
    assert np.array_equal(np.fromfile(tmp_path / "xp_drift.bin", xp.dtype).reshape(m, 3), xe)
    assert np.array_equal(np.fromfile(tmp_path / "vp_drift.bin", np.float32).reshape(m, 3), ve)
    assert np.array_equal(np.fromfile(tmp_path / "pid_drift.bin", np.int64), pe)
    assert np.array_equal(np.fromfile(tmp_path / "rhoc_drift.bin", np.int32).reshape(re.shape), re)
    sp = np.fromfile(tmp_path / "spine.bin", np.int64)
    ref = orc.spine_image(P, re)
    nb, npb = ref["idx_b_l"].size, ref["ppl"].size
    parts = [sp[:nb], sp[nb:2 * nb]] + [sp[2 * nb + q * npb:2 * nb + (q + 1) * npb] for q in range(6)]
    for got, key in zip(parts, ("idx_b_l", "idx_b_r", "ppl0", "pplr", "pprl", "ppr0", "ppl", "ppr")):
        assert np.array_equal(got, ref[key].reshape(-1)), key


REFHOST = os.path.join(ROOT, "cube_nu_amd", "lib", "cube_ref_host")


def test_reference_side_shims_kick_buffer_v_drift(tmp_path):
    """The drop-in kick / buffer_v / drift subroutines a CUBE maintainer adds
    (cube_nu_amd/fortran/cube_ref_shims.f90), compiled against stand-in
    parameters/variables modules with the reference's names and shapes and
    called argument-less as main.f90:34-38 calls them: the kick's outputs
    against the oracle (velocity change rms relative <= 1e-5, the Delta-t
    criteria to 1e-5, napm exact), buffer_v and then the drift of the
    kicked state bit-exact."""
    geo = Geometry(64)
    P = oparams(geo)
    xp, vp, pid, rhoc = buffered_set(geo, seed=19, mean=5.0)
    n = xp.shape[0]
    spine = orc.spine_image(P, rhoc)
    for name, arr in (("xp", xp), ("vp", vp), ("pid", pid), ("rhoc", rhoc), ("idx_b_r", spine["idx_b_r"])):
        np.ascontiguousarray(arr).tofile(tmp_path / f"{name}.bin")
    a_mid, dt, dt_old, dt_drift = 0.5, 0.3, 0.4, 0.6
    print(f"reference-side shims: ng=64 np={n}", flush=True)
    subprocess.run([REFHOST, str(tmp_path), str(n), str(a_mid), str(dt), str(dt_old), str(dt_drift)],
                   check=True, timeout=120)
    raw = open(tmp_path / "ref_kick.bin", "rb").read()
    dts = np.frombuffer(raw[:24], np.float32)
    testrho = np.frombuffer(raw[24:32], np.float64)[0]
    napm = np.frombuffer(raw[32:60], np.int32)
    vk = np.frombuffer(raw[60:60 + 12 * n], np.float32).reshape(n, 3)
    # the host keeps a_mid, dt and dt_old in real(4), as variables.f90:35 does
    a32, dt32, dtd32 = (float(np.float32(v)) for v in (a_mid, dt, dt_drift))
    vo = vp.copy()
    out_o, _, _ = orc.kick(P, xp, vo, rhoc, orc.tf_default(P), a32, dt32)
    orc.buffer_v(P, vo, rhoc)
    assert rel_rms(vk - vp, vo - vp) <= 1e-5
    for k, v in zip(("dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "dt_vmax", "vz_max"), dts):
        assert abs(v - out_o[k]) <= 1e-5 * abs(out_o[k]), k
    assert abs(testrho - out_o["testrho"]) <= 1e-6 * out_o["testrho"]
    assert list(napm) == out_o["napm"]
    # the drift of the kicked state, dt_old = dt of the kick
    st, xo, vdo, po, ro, oo = orc.drift(P, xp, np.ascontiguousarray(vk), pid, rhoc, dt32, dtd32, float(dts[5]))
    assert st == 0
    raw = open(tmp_path / "ref_drift.bin", "rb").read()
    m = int(np.frombuffer(raw[:8], np.int64)[0])
    off = 8
    xg = np.frombuffer(raw[off:off + 6 * m], np.int16).reshape(m, 3); off += 6 * m
    vg = np.frombuffer(raw[off:off + 12 * m], np.float32).reshape(m, 3); off += 12 * m
    pg = np.frombuffer(raw[off:off + 8 * m], np.int64); off += 8 * m
    rg = np.frombuffer(raw[off:off + 4 * rhoc.size], np.int32).reshape(rhoc.shape); off += 4 * rhoc.size
    bl = np.frombuffer(raw[off:off + 8 * spine["idx_b_l"].size], np.int64).reshape(spine["idx_b_l"].shape)
    off += 8 * spine["idx_b_l"].size
    br = np.frombuffer(raw[off:off + 8 * spine["idx_b_r"].size], np.int64).reshape(spine["idx_b_r"].shape)
    assert m == oo["nplocal"]
    assert np.array_equal(rg, ro) and np.array_equal(xg, xo) and np.array_equal(vg, vdo) and np.array_equal(pg, po)
    sp = orc.spine_image(P, ro)
    assert np.array_equal(bl, sp["idx_b_l"]) and np.array_equal(br, sp["idx_b_r"])
