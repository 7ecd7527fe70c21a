"""Teacher-forced parity of every stage of FluidSolver::update() on the GPU.

The oracle runs a scene for a few steps and then walks through the stages of
one update() in the reference's order (3:404-441, 4:862-894, 6:987-1026,
7:987-1030). Before each stage the oracle's whole state is uploaded into the
product, both run the stage, and every state array (and the stage's scalar
result) must agree bit for bit. Dot products use the reference's summation
order (dot_order = SEQUENTIAL) so that they are bit-exact too.

project() is also run with a 3-iteration budget so that the budget-exit state
(_z holding A s at non-FLUID cells, the deferred s update) is covered.
"""
import pytest

import oracle_lib as ol
from gpu_util import compare_fields, copy_state, make_product

pytestmark = pytest.mark.gpu

MULTI = [(ol.BODY_BOX, 0.5, 0.6, 0.3, 0.1, 0.7, 0.0, 0.0, 0.0),
         (ol.BODY_SPHERE, 0.3, 0.35, 0.15, 0.0, 0.0, 0.0, 0.0, 0.0),
         (ol.BODY_BOX, 0.7, 0.3, 0.1, 0.12, -0.4, 0.0, 0.0, 0.0)]


def stage_sequence(level, dt):
    """(stage, iarg, darg, advance): the stages of one update() in order.
    advance=False marks a probe: it is compared, then the oracle's state is
    restored, so that the chain continues from the state update() would see."""
    qs = [ol.Q_D, ol.Q_U, ol.Q_V] + ([ol.Q_T] if level >= 6 else [])
    seq = []
    if level >= 4:
        seq += [("FILL_SOLID", q, 0.0, True) for q in qs]
    if level >= 6:
        seq += [("COPY_T_TO_R", 0, 0.0, True), ("BUILD_HEAT", 0, dt, True), ("BUILD_PRECON", 0, 0.0, True),
                ("PROJECT", 3, 0.0, False), ("PROJECT", 2000, 0.0, True), ("COPY_P_TO_T", 0, 0.0, True)]
    seq += [("BUILD_RHS", 0, 0.0, True)]
    if level >= 7:
        seq += [("COMPUTE_DENSITIES", 0, 0.0, True)]
    seq += [("BUILD_PRESSURE", 0, dt, True), ("BUILD_PRECON", 0, 0.0, True),
            ("APPLY_PRECON", 0, 0.0, False), ("DOT", 0, 0.0, False), ("INF_NORM", 0, 0.0, False),
            ("MATVEC", 0, 0.0, False), ("DOT", 1, 0.0, False),
            ("PROJECT", 0, 0.0, False), ("PROJECT", 1, 0.0, False), ("PROJECT", 3, 0.0, False),
            ("PROJECT", 600, 0.0, True), ("APPLY_PRESSURE", 0, dt, True)]
    if level >= 6:
        seq += [("ADD_BUOYANCY", 0, dt, True)]
    if level >= 4:
        seq += [("SET_BOUNDARY", 0, 0.0, True)]
    seq += [("ADVECT", q, dt, True) for q in qs]
    if level >= 4:
        seq += [("EXTRAPOLATE", q, 0.0, True) for q in qs]
    seq += [("FLIP", q, 0.0, True) for q in qs]
    return seq


CASES = [
    (3, 48, 40, None, 3),
    (4, 72, 40, MULTI, 3),
    (5, 72, 40, MULTI, 3),
    (6, 72, 40, MULTI, 3),
    (6, 64, 64, None, 4),
    (7, 72, 40, MULTI, 3),
    (7, 130, 70, MULTI, 2),
]


@pytest.mark.parametrize("level,w,h,bodies,warm", CASES)
def test_stages_teacher_forced(level, w, h, bodies, warm):
    sc = dict(ol.REF_SCENES[level])
    if bodies is not None:
        sc["bodies"] = bodies
    orc = ol.OracleSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    for _ in range(warm):
        if sc["inflow"] is not None:
            orc.add_inflow(*sc["inflow"])
        orc.update(sc["dt"])
    gpu = make_product(level, w, h, sc["params"], sc["bodies"], dot_order=1)
    gpu._push_bodies()
    fields = ol.state_fields(level)
    for name, iarg, darg, advance in stage_sequence(level, sc["dt"]):
        copy_state(orc, gpu, fields)
        saved = None if advance else {f: orc.field(f) for f in fields}
        want = orc.stage(name, iarg, darg)
        got = gpu.stage(name, iarg, darg)
        where = f"level {level} {w}x{h} stage {name}({iarg})"
        assert got == want or (got != got and want != want), f"{where}: result {got!r} vs {want!r}"
        bad = compare_fields(orc, gpu, fields, where=where)
        assert not bad, "\n".join(bad[:8])
        if saved is not None:
            for f, v in saved.items():
                if v is not None:
                    orc.set_field(f, v)
    gpu.close()
    orc.close()


def test_max_timestep_matches_oracle():
    """maxTimestep (1:310-328) on the device against the oracle's restatement
    (itself pinned to the reference's solver 1), on velocity fields from the
    solver-7 scene and on random fields, bit for bit."""
    import ctypes
    import numpy as np
    import ifl_fluids as F
    lib = ol.oracle_lib()
    lib.orc_max_timestep.restype = ctypes.c_double
    lib.orc_max_timestep.argtypes = [ctypes.c_void_p]
    sc = ol.REF_SCENES[7]
    for w, h, steps in ((96, 80, 3), (130, 70, 2)):
        orc = ol.OracleSolver(7, w, h, sc["params"], sc["bodies"], fresh=True)
        for _ in range(steps):
            orc.add_inflow(*sc["inflow"])
            orc.update(sc["dt"])
        gpu = make_product(7, w, h, sc["params"], sc["bodies"])
        rng = np.random.default_rng(w)
        fu, fv = ol.qfield(ol.Q_U, ol.QF_SRC), ol.qfield(ol.Q_V, ol.QF_SRC)
        for case in ("scene", "random"):
            if case == "random":
                for f in (fu, fv):
                    orc.set_field(f, rng.normal(0, 2.0, orc.field(f).shape))
            for f in (fu, fv):
                gpu.set_field(f, orc.field(f))
            want = lib.orc_max_timestep(orc.h)
            got = gpu.max_timestep()
            print(f"{w}x{h} {case}: dt {got!r} (oracle {want!r})", flush=True)
            assert np.float64(got).tobytes() == np.float64(want).tobytes()
        gpu.close()
        orc.close()


@pytest.mark.parametrize("level", [3, 4, 5, 6, 7, 8])
def test_to_image_matches_oracle(level, tmp_path):
    """toImage rendered on the device against the oracle's restatement (itself
    pinned to the reference's toImage), after a few steps, both panels for
    solvers 6-8, byte for byte; and the frame file decodes to the same pixels."""
    import ctypes
    import numpy as np
    import ifl_fluids as F
    from test_png import decode
    w, h = 64, 56
    sc = ol.REF_SCENES[level]
    if level == 8:
        F.frand_set_state(0xBA5EBA11)
    orc = ol.OracleSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    gpu = make_product(level, w, h, sc["params"], sc["bodies"], dot_order=1)
    for _ in range(4):
        for s in (orc, gpu):
            if sc["inflow"] is not None:
                s.add_inflow(*sc["inflow"])
            s.update(sc["dt"])
    lib = ol.oracle_lib()
    lib.orc_to_image.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    for heat in ((0, 1) if level >= 6 else (0,)):
        want = np.zeros((h, w * (2 if heat else 1), 4), dtype=np.uint8)
        lib.orc_to_image(orc.h, want.ctypes.data, heat)
        got = gpu.to_image(heat)
        assert got.tobytes() == want.tobytes(), f"heat {heat}: {np.count_nonzero(got != want)} bytes differ"
        p = tmp_path / f"frame{heat}.png"
        F.write_png(str(p), got)
        assert np.array_equal(decode(str(p)), want)
    gpu.close()
    orc.close()


def test_step_reports_nonfinite_fields():
    """ifl_step_stats.nonfinite: 0 on a normal step, 1 once a field holds a
    NaN (the reference carries such values on without a word)."""
    import numpy as np
    sc = ol.REF_SCENES[7]
    gpu = make_product(7, 64, 48, sc["params"], sc["bodies"])
    gpu.add_inflow(*sc["inflow"])
    st = gpu.update(sc["dt"])
    assert st.nonfinite == 0
    f = ol.qfield(ol.Q_D, ol.QF_SRC)
    d = gpu.field(f)
    d[50 + 10 * 64] = np.nan  # a fluid cell far from the box
    gpu.set_field(f, d)
    st = gpu.update(sc["dt"])
    assert st.nonfinite == 1
    gpu.close()


@pytest.mark.gpu
@pytest.mark.parametrize("level", [4, 7])
def test_step_reports_border_solid(level):
    """ifl_step_stats.border_solid: 1 when a solid body covers cells on the
    domain border (the reference's extrapolation then reads outside the
    grid), 0 otherwise; read at the end of the step, also in the batched form."""
    import math
    import ifl_fluids as F
    sc = ol.REF_SCENES[7 if level == 7 else 4]
    inside = (F.BODY_BOX, 0.5, 0.4, 0.2, 0.1, 0.0, 0.0, 0.0, 0.0)
    touching = (F.BODY_BOX, 0.02, 0.4, 0.2, 0.1, math.pi * 0.25, 0.0, 0.0, 0.0)
    got = []
    for body in (inside, touching):
        gpu = make_product(level, 64, 48, sc["params"], [body])
        gpu.add_inflow(*sc["inflow"])
        got.append(gpu.update(sc["dt"]).border_solid)
        gpu.close()
    assert got == [0, 1]
    # the same through the batched form, two solvers side by side
    objs = [[F.SolidBox(*inside[1:])], [F.SolidBox(*touching[1:])]]
    solvers = [F.FluidSolver(level, 64, 48, sc["params"], o) for o in objs]
    batch = F.Batch(solvers)
    for s in solvers:
        s.add_inflow(*sc["inflow"])
    stats = batch.update(sc["dt"])
    assert [st.border_solid for st in stats] == [0, 1]
    batch.close()
    for s in solvers:
        s.close()
