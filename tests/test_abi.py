"""The C-ABI library loads and exports exactly the entry points include/cube_p3m.h declares."""
import ctypes
import os
import re

import pytest

from cube_nu_amd import ABI_FUNCTIONS, load_library
from cube_nu_amd.p3m import CUBE_ERR_NO_DEVICE, CUBE_OK, Config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "cube_p3m.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cube_p3m_\w+)\(", src, re.M)))


def test_header_lists_all_bound_functions():
    assert header_functions() == sorted(ABI_FUNCTIONS)


def test_library_exports_every_declared_symbol():
    lib = load_library()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_version_string():
    assert b"gfx950" in load_library().cube_p3m_version()


def test_create_without_device_or_bad_config():
    lib = load_library()
    cfg = Config(64, 4, 4, 4, 4, 16, 3, 32, 600.0, 3.5, 3.5, 3.5, 0.06, 1, 1, 1, 0.0, 0.0, 2.0, 6.0, 0)
    ctx = ctypes.c_void_p()
    # izipx=3 is rejected before any device call
    assert lib.cube_p3m_create(ctypes.byref(cfg), ctypes.byref(ctx)) == 1
    cfg.izipx = 2
    rc = lib.cube_p3m_create(ctypes.byref(cfg), ctypes.byref(ctx))
    if rc == CUBE_OK:
        lib.cube_p3m_destroy(ctx)
        pytest.skip("a HIP device is present")
    assert rc == CUBE_ERR_NO_DEVICE


def test_create_rejects_ratio_cs_not_a_power_of_two():
    """The drift's division by ratio_cs is a product with its reciprocal and
    the PP binning takes xvec = xf / ratio_cs: both exact only for a power of
    two, so another ratio_cs is refused with CUBE_ERR_ARG before any device
    call."""
    lib = load_library()
    ctx = ctypes.c_void_p()
    cfg = Config(96, 4, 4, 3, 4, 16, 2, 32, 600.0, 3.5, 3.5, 3.5, 0.06, 1, 1, 1, 0.0, 0.0, 2.0, 6.0, 0)
    assert lib.cube_p3m_create(ctypes.byref(cfg), ctypes.byref(ctx)) == 1


DRIVER = os.path.join(ROOT, "cube_nu_amd", "lib", "cube_p3m_driver")


def test_fortran_bindc_types_match_c_structs():
    """The ISO_C_BINDING derived types have the C structs' sizes."""
    import subprocess

    from cube_nu_amd.p3m import DriftIn, DriftOut, ICParams, KickIn, KickOut

    out = subprocess.run([DRIVER, "--sizes"], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out[:6]]
    assert sizes == [ctypes.sizeof(c) for c in (Config, KickIn, KickOut, DriftIn, DriftOut, ICParams)]


def test_fortran_driver_links_the_library():
    import subprocess

    r = subprocess.run([DRIVER, "--version"], capture_output=True, text=True, check=True)
    assert "gfx950" in r.stdout


def test_pm_header_lists_all_bound_functions_and_library_exports_them():
    from cube_nu_amd.pmdiff import PM_ABI_FUNCTIONS

    src = open(os.path.join(ROOT, "include", "cube_pm.h")).read()
    names = sorted(set(re.findall(r"^(?:int|const char\*)\s+(cube_pm_\w+)\(", src, re.M)))
    assert names == sorted(PM_ABI_FUNCTIONS)
    lib = load_library()
    for name in names:
        assert hasattr(lib, name), name
