"""Full FluidSolver::update() trajectories on the GPU against the oracle.

With dot_order = SEQUENTIAL the product sums dot products in the reference's
left-to-right order, so every array of the state must match the oracle bit
for bit after every step, for solvers 3..7 (the oracle is itself pinned
bit-for-bit to the reference build, tests/test_oracle_vs_reference.py).
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import compare_fields, make_product

pytestmark = pytest.mark.gpu

MULTI = [(ol.BODY_BOX, 0.5, 0.6, 0.3, 0.1, 0.7, 0.0, 0.0, 0.0),
         (ol.BODY_SPHERE, 0.3, 0.35, 0.15, 0.0, 0.0, 0.0, 0.0, 0.0),
         (ol.BODY_BOX, 0.7, 0.3, 0.1, 0.12, -0.4, 0.0, 0.0, 0.0)]
MOVING = [(ol.BODY_SPHERE, 0.55, 0.5, 0.2, 0.2, 0.3, 0.1, -0.2, 1.0)]

CASES = [
    (3, 48, 40, None, 12),
    (3, 64, 130, None, 6),
    (4, 64, 64, None, 10),
    (4, 72, 40, dict(bodies=MULTI), 8),
    (5, 64, 64, dict(bodies=MULTI), 8),
    (5, 40, 72, dict(bodies=MOVING), 8),
    (6, 64, 64, None, 8),
    (6, 72, 40, dict(bodies=MULTI), 8),
    (7, 64, 64, None, 8),
    (7, 72, 40, dict(bodies=MULTI), 8),
    # 257 strips, more than the GPU's CUs: one launch per sweep with the
    # strips dealt upstream first
    (3, 32, 16448, None, 2),
]


@pytest.mark.parametrize("level,w,h,scene,steps", CASES)
def test_update_bit_exact_sequential_dots(level, w, h, scene, steps):
    sc = dict(ol.REF_SCENES[level])
    if scene:
        sc.update(scene)
    orc = ol.OracleSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
    gpu = make_product(level, w, h, sc["params"], sc["bodies"], dot_order=1)
    fields = ol.state_fields(level)
    for k in range(steps):
        if sc["inflow"] is not None:
            orc.add_inflow(*sc["inflow"])
            gpu.add_inflow(*sc["inflow"])
        orc.update(sc["dt"])
        st = gpu.update(sc["dt"])
        bad = compare_fields(orc, gpu, fields, where=f"step {k}")
        assert not bad, "\n".join(bad[:8])
        if (k + 1) % sc["substeps"] == 0:
            orc.bodies_update(sc["dt"])
            gpu.bodies_update(sc["dt"])
    gpu.close()
    orc.close()
