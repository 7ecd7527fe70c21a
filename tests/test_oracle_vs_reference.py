"""The C restatement (oracle/liboracle.so) against the reference's own
Fluid.cpp compiled from its sources (oracle/_ref/libref<N>.so): every array
of the solver state, bit for bit, after every step, plus the reference's
stdout lines (iteration counts and residuals), for solvers 3..8.

Scenes avoid configurations where the reference itself has undefined
behaviour (bodies touching the domain border; tall grids at levels 7-8, where
_vDensity is read past its end through the u-grid stride, 7:700).
"""
import math

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(7), reason="reference build (oracle/_ref) not present")

MULTI = [(ol.BODY_BOX, 0.5, 0.6, 0.3, 0.1, 0.7, 0.0, 0.0, 0.0),
         (ol.BODY_SPHERE, 0.3, 0.35, 0.15, 0.15, 0.0, 0.0, 0.0, 0.0),
         (ol.BODY_BOX, 0.7, 0.3, 0.1, 0.12, -0.4, 0.0, 0.0, 0.0)]
MOVING = [(ol.BODY_SPHERE, 0.55, 0.5, 0.2, 0.2, 0.3, 0.1, -0.2, 1.0),
          (ol.BODY_BOX, 0.3, 0.8, 0.2, 0.08, 0.2, 0.0, 0.1, -2.0)]

CASES = [
    (3, 48, 40, None, 16), (3, 40, 72, None, 8), (3, 72, 40, None, 8),
    (4, 48, 40, None, 12), (4, 64, 64, dict(bodies=MULTI), 12), (4, 40, 72, dict(bodies=MOVING), 8),
    (5, 48, 40, None, 12), (5, 64, 64, dict(bodies=MULTI), 12), (5, 72, 40, dict(bodies=MOVING), 8),
    (6, 48, 40, None, 12), (6, 64, 64, dict(bodies=MULTI), 10), (6, 40, 72, dict(bodies=MOVING), 8),
    (7, 48, 40, None, 12), (7, 64, 64, dict(bodies=MULTI), 10), (7, 72, 40, dict(bodies=MOVING), 8),
    (8, 48, 40, None, 12), (8, 64, 64, dict(bodies=MULTI), 8),
]


def _state_equal(a, b, level):
    bad = []
    for f in ol.state_fields(level):
        x, y = a.field(f), b.field(f)
        if x is None and y is None:
            continue
        if x is None or y is None or x.tobytes() != y.tobytes():
            bad.append(f)
    return bad


@pytest.mark.parametrize("level,w,h,scene,steps", CASES)
def test_trajectory_bit_identical(level, w, h, scene, steps):
    sc = dict(ol.REF_SCENES[level])
    if scene:
        sc.update(scene)
    ref = ol.RefSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    orc = ol.OracleSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    assert not _state_equal(ref, orc, level), "initial state differs"
    for k in range(steps):
        if sc["inflow"] is not None:
            ref.add_inflow(*sc["inflow"])
            orc.add_inflow(*sc["inflow"])
        ref.update(sc["dt"])
        orc.update(sc["dt"])
        bad = _state_equal(ref, orc, level)
        assert not bad, f"step {k}: fields {bad} differ"
        if (k + 1) % sc["substeps"] == 0:
            ref.bodies_update(sc["dt"])
            orc.bodies_update(sc["dt"])
    assert ref.log() == orc.log()
    assert ref.log().count("\n") >= steps
    orc.close()
    ref.close()


@pytest.mark.parametrize("level", [3, 5, 7])
def test_single_stages_bit_identical(level):
    """Each stage on the same mid-trajectory state, reference vs oracle."""
    sc = ol.REF_SCENES[level]
    w, h = 56, 40
    ref = ol.RefSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    orc = ol.OracleSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    for k in range(5):
        for s in (ref, orc):
            s.add_inflow(*sc["inflow"])
            s.update(sc["dt"])
    stages = [("BUILD_RHS", 0), ("BUILD_PRESSURE", 0), ("BUILD_PRECON", 0), ("APPLY_PRECON", 0),
              ("MATVEC", 0), ("APPLY_PRESSURE", 0)]
    if level >= 4:
        stages = [("FILL_SOLID", 0), ("FILL_SOLID", 2), ("FILL_SOLID", 3), ("SET_BOUNDARY", 0)] + stages + \
                 [("EXTRAPOLATE", 0), ("EXTRAPOLATE", 2)]
    if level >= 7:
        stages.insert(0, ("COMPUTE_DENSITIES", 0))
    stages += [("ADVECT", 0), ("ADVECT", 2)]
    for name, ia in stages:
        ref.stage(name, ia, sc["dt"])
        orc.stage(name, ia, sc["dt"])
        bad = _state_equal(ref, orc, level)
        assert not bad, f"after {name}({ia}): {bad}"
    for which in (0, 1):
        assert ref.stage("DOT", which) == orc.stage("DOT", which)
    assert ref.stage("INF_NORM") == orc.stage("INF_NORM")


def test_frand_matches_reference_lcg():
    """frand() 8:38-46 is an LCG; the oracle's stream equals the reference's."""
    ref = ol.RefSolver(8, 8, 8, (0.1, 0.25, 0.01), [], fresh=True)
    lib = ol.oracle_lib()
    lib.orc_frand_seed(0xBA5EBA11)
    for _ in range(2 * 4 * 8 * 8):  # initParticles draws x and y per candidate
        lib.orc_frand()
    a = [ref.lib.ref_frand() for _ in range(1000)]
    b = [lib.orc_frand() for _ in range(1000)]
    assert a == b
