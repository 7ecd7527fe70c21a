"""ifl_to_image_static, the allocation-free toImage, against ifl_to_image
through the C ABI: the same bytes on two calls in a row, the second on a
smaller grid, with and without the heat panel. The first grid's heat frame
(4.9 MB) needs two row windows of the 4 MiB device scratch.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import make_product

pytestmark = pytest.mark.gpu


def _bind():
    import ifl_fluids as F
    f = F.lib.ifl_to_image_static
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    f.restype = ctypes.c_int
    return F, f


def _frame(F, fn, solver, heat):
    a = np.empty((solver.h, solver.w * (2 if heat else 1), 4), dtype=np.uint8)
    a.fill(0xA5)
    rc = fn(solver._h, a.ctypes.data, a.nbytes, 1 if heat else 0)
    assert rc == 0, F.lib.ifl_last_error(solver._h).decode()
    return a


@pytest.mark.parametrize("level", [7, 4])
def test_to_image_static_matches_to_image(level):
    F, static = _bind()
    sc = ol.REF_SCENES[level]
    for w, h in ((880, 700), (96, 64)):  # second call: a smaller grid
        print(f"ifl_to_image_static level {level} {w}x{h}", flush=True)
        solver = make_product(level, w, h, sc["params"], sc["bodies"])
        for _ in range(2):
            solver.add_inflow(*sc["inflow"])
            solver.update(sc["dt"])
        for heat in ((False, True) if level >= 6 else (False,)):
            want = solver.to_image(render_heat=heat)
            got = _frame(F, static, solver, heat)
            assert got.tobytes() == want.tobytes(), f"{w}x{h} heat={heat}: frames differ"
            assert np.any(got[..., 0] != got[0, 0, 0])  # not a blank frame
        solver.close()


def test_to_image_static_rejects_a_wrong_size():
    F, static = _bind()
    sc = ol.REF_SCENES[7]
    solver = make_product(7, 48, 40, sc["params"], sc["bodies"])
    a = np.empty(48 * 40 * 4 - 4, dtype=np.uint8)
    rc = static(solver._h, a.ctypes.data, a.nbytes, 0)
    assert rc == 1 and b"w*h*4" in F.lib.ifl_last_error(solver._h)
    solver.close()
