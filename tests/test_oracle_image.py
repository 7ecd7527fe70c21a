"""toImage (3:455, 4:903, 5:961, 6:1021, 7:1047, 8:1426): the oracle's
restatement against the reference's own classes (oracle/_ref) after a few
steps of each solver's scene, both panels for solvers 6-8, byte for byte."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(7), reason="reference build (oracle/_ref) not present")


def image(solver, lib_fn, w, h, heat):
    buf = (ctypes.c_ubyte * (w * h * 4 * (2 if heat else 1)))()
    lib_fn(solver.h, buf, heat)
    return np.frombuffer(buf, dtype=np.uint8).copy()


@pytest.mark.parametrize("level", [3, 4, 5, 6, 7, 8])
def test_to_image_matches_reference(level):
    w, h, steps = 64, 56, 6
    sc = ol.REF_SCENES[level]
    ref = ol.RefSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    orc = ol.OracleSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    for _ in range(steps):
        for s in (ref, orc):
            if sc["inflow"] is not None:
                s.add_inflow(*sc["inflow"])
            s.update(sc["dt"])
    ref.lib.ref_to_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    olib = ol.oracle_lib()
    olib.orc_to_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    for heat in ((0, 1) if level >= 6 else (0,)):
        a = image(ref, ref.lib.ref_to_image, w, h, heat)
        b = image(orc, olib.orc_to_image, w, h, heat)
        assert a.tobytes() == b.tobytes(), f"level {level} heat {heat}: {np.count_nonzero(a != b)} bytes differ"
        assert len(np.unique(a)) > 2  # the frame is not blank
    orc.close()
    ref.close()
