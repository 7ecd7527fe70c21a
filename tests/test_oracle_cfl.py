"""maxTimestep (1:310-328), solver 1's CFL bound: the oracle's restatement
against the reference's solver 1 compiled from its own sources
(oracle/_ref/libref1.so), on the velocity fields of the reference's own
scene (1:340-370) after several steps, plus the closed forms."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as ol

REF1 = os.path.join(ol.REF_DIR, "libref1.so")


def ref1():
    lib = ctypes.CDLL(REF1)
    lib.ref1_create.restype = ctypes.c_void_p
    lib.ref1_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.ref1_add_inflow.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 7
    lib.ref1_update.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.ref1_max_timestep.restype = ctypes.c_double
    lib.ref1_max_timestep.argtypes = [ctypes.c_void_p]
    lib.ref1_velocity.restype = ctypes.POINTER(ctypes.c_double)
    lib.ref1_velocity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
    return lib


def velocity(lib, h, q):
    n = ctypes.c_long()
    p = lib.ref1_velocity(h, q, ctypes.byref(n))
    return np.ctypeslib.as_array(p, shape=(n.value,)).copy()


def oracle_dt(w, h, u, v):
    orc = ol.OracleSolver(3, w, h, (0.1,), [])
    orc.set_field(ol.qfield(ol.Q_U, ol.QF_SRC), u)
    orc.set_field(ol.qfield(ol.Q_V, ol.QF_SRC), v)
    lib = ol.oracle_lib()
    lib.orc_max_timestep.restype = ctypes.c_double
    lib.orc_max_timestep.argtypes = [ctypes.c_void_p]
    r = lib.orc_max_timestep(orc.h)
    orc.close()
    return r


@pytest.mark.skipif(not os.path.exists(REF1), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("w,h", [(64, 64), (48, 80), (80, 48)])
def test_max_timestep_matches_solver1(w, h):
    lib = ref1()
    s = lib.ref1_create(w, h, 0.1)
    checked = 0
    for k in range(12):
        lib.ref1_add_inflow(s, 0.45, 0.2, 0.1, 0.01, 1.0, 0.0, 3.0)  # 1:363
        lib.ref1_update(s, 0.005)
        if k % 3 == 2:
            want = lib.ref1_max_timestep(s)
            got = oracle_dt(w, h, velocity(lib, s, 2), velocity(lib, s, 3))
            assert np.float64(got).tobytes() == np.float64(want).tobytes(), (k, got, want)
            assert 0 < want < 1
            checked += 1
    assert checked == 4


def test_max_timestep_closed_forms():
    w, h = 32, 24
    zero_u, zero_v = np.zeros((w + 1) * h), np.zeros(w * (h + 1))
    assert oracle_dt(w, h, zero_u, zero_v) == 1.0          # no flow: the clamp
    # a uniform flow (3, 4) at every face: speed 5 at every centre
    u = np.full((w + 1) * h, 3.0)
    v = np.full(w * (h + 1), 4.0)
    assert oracle_dt(w, h, u, v) == 2.0 * (1.0 / 24) / 5.0
