"""The C ABI: libifl.so loads, exports every entry point include/ifl.h
declares, and its host-side logic answers without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ifl.h")
LIB = os.path.join(ROOT, "incremental-fluids_amd", "libifl.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint32_t|const char\*)\s+(ifl_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("ifl_create", "ifl_destroy", "ifl_update", "ifl_add_inflow", "ifl_set_bodies",
              "ifl_upload", "ifl_download", "ifl_stage", "ifl_profile_read", "ifl_init_particles",
              "ifl_frand_state", "ifl_frand_set_state"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing


def test_config_defaults_are_the_reference_literals():
    import ifl_fluids as F
    c = F.default_config(7, 128, 96)
    assert (c.level, c.w, c.h, c.pcg_limit) == (7, 128, 96, 2000)
    assert (c.tau, c.sigma, c.tol, c.t_amb, c.gravity, c.flip_alpha) == (0.97, 0.25, 1e-5, 294.0, 9.81, 0.001)
    assert F.default_config(3, 8, 8).pcg_limit == 600


def test_bad_arguments_are_rejected_without_touching_a_device():
    import ifl_fluids as F
    h = ctypes.c_void_p()
    for lv, w, hh in [(2, 64, 64), (9, 64, 64), (7, 2, 64), (7, 64, 0)]:
        cfg = F.default_config(lv, w, hh)
        assert F.lib.ifl_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
        assert b"level" in F.lib.ifl_last_error(None)
    assert F.lib.ifl_destroy(None) == 0
    assert F.lib.ifl_update(None, 0.005, None) == 1


def test_no_cpu_fallback_without_a_device():
    """On a host without a GPU, creation fails loudly instead of computing."""
    import ifl_fluids as F
    try:
        s = F.FluidSolver(7, 64, 64, (0.1, 1.0, 0.01))
    except F.IflError as e:
        assert "no HIP device" in str(e)
        return
    s.close()
    pytest.skip("a GPU is present")


def test_frand_state_is_process_wide_host_state():
    """frand()'s seed (8:39) starts at 0xBA5EBA11 and is read and set on the
    host, without a device."""
    import ifl_fluids as F
    old = F.frand_state()
    try:
        F.frand_set_state(0x12345678)
        assert F.frand_state() == 0x12345678
    finally:
        F.frand_set_state(old)
    assert F.frand_state() == old
