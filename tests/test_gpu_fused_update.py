"""The PCG iterations of levels 6-8 (FP64, tree-ordered dots) form the
vector update s = z + beta s (and p += alpha s) inside the matrix-vector
product (7:866-882 reordered into one pass, the new s written to a second
array by turns). Every value is the separate kernels' value, so a solver
built with IFL_NO_FUSED_UPDATE=1 (the separate update kernel) must give the
same iterates bit for bit: iteration counts, statuses and the whole state,
including budget exits after an odd and an even number of iterations (the
last s then lies in either array)."""
import os

import pytest

import oracle_lib as ol
from gpu_util import compare_fields

pytestmark = pytest.mark.gpu


def _make(level, n, fused, scene, **kw):
    import ifl_fluids as F
    old = os.environ.get("IFL_NO_FUSED_UPDATE")
    os.environ["IFL_NO_FUSED_UPDATE"] = "0" if fused else "1"
    try:
        bodies = [F.SolidBox(*b[1:]) if b[0] == F.BODY_BOX else F.SolidSphere(b[1], b[2], b[3], *b[5:])
                  for b in scene["bodies"]]
        return F.FluidSolver(level, n, n, scene["params"], bodies, dot_order=F.DOT_TREE, **kw)
    finally:
        if old is None:
            del os.environ["IFL_NO_FUSED_UPDATE"]
        else:
            os.environ["IFL_NO_FUSED_UPDATE"] = old


def _run(sv, scene, steps, level):
    rows = []
    for _ in range(steps):
        if level != 8:
            for inf in scene["inflows"]:
                sv.add_inflow(*inf)
        st = sv.update(scene["dt"])
        rows.append((st.heat.status, st.heat.iterations, st.pressure.status, st.pressure.iterations,
                     st.particle_count))
    return rows


@pytest.mark.parametrize("level,n,limit", [(7, 192, None), (6, 160, None), (8, 128, None),
                                           (7, 128, 7), (7, 128, 8)])
def test_fused_update_identical(level, n, limit):
    import ifl_fluids as F
    import ifl_scenes as S
    scene = {7: S.config3_scene(), 6: S.config2_scene(static=True), 8: S.config5_scene()}[level]
    kw = {} if limit is None else {"pcg_limit": limit}
    steps = 3
    out = []
    for fused in (True, False):
        if level == 8:
            F.frand_set_state(0xBA5EBA11)
        sv = _make(level, n, fused, scene, **kw)
        out.append((sv, _run(sv, scene, steps, level)))
    (a, ra), (b, rb) = out
    print(f"level {level} {n}^2 limit {limit}: {ra}", flush=True)
    assert ra == rb
    if limit is not None:
        assert all(r[2] == F.SOLVE_BUDGET for r in ra)
    fields = ol.state_fields(level)  # with the particles at level 8
    bad = compare_fields(a, b, fields, where=f"level {level}")
    assert not bad, "\n".join(bad[:6])
    a.close()
    b.close()
