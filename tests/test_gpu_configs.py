"""The seeded workloads (incremental-fluids_amd/ifl_scenes.py) on the GPU.

* Solver 3 with the literal 0.1 x 0.01 inflow at 128x128: the rectangle
  covers one row of cell centres whose cubic pulse is 0 (3:167-184), so every
  field stays exactly zero and no solve runs, for 64 steps, as in the
  reference (tests/test_oracle_configs.py pins that on the reference build).
* Solver 6 with the seeded two boxes and sphere, static variant, at 256x256:
  8 whole update() calls bit for bit against the oracle.
* The same scene at 1024x1024: the first update() from construction bit for
  bit against digests of the reference build's own state
  (tests/golden/make_configs.py), in the static variant and in the moving
  one, where the reference's pressure solve runs out of its 2000-iteration
  budget (6:990) and the product must exit the same way.
* Solver 7, the benchmark scene at 2048x2048: the first update() with both
  project() budgets capped (the reference's update() sequence with
  project(cap), run stage by stage on the reference build) bit for bit
  against the reference's digests. The uncapped step is checked in
  tests/test_gpu_fullsize.py (iteration counts).

All bit-exact cases use the reference's dot-product order
(dot_order = SEQUENTIAL).
"""
import hashlib
import json
import os
import re

import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import compare_fields

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def _solver(F, level, n, sc, **kw):
    objs = []
    for b in sc["bodies"]:
        t, x, y, sx, sy, th, vx, vy, om = b
        if t == F.BODY_SPHERE:
            objs.append(F.SolidSphere(x, y, sx, th, vx, vy, om))
        else:
            objs.append(F.SolidBox(x, y, sx, sy, th, vx, vy, om))
    return F.FluidSolver(level, n, n, sc["params"], objs, **kw)


def _check_digests(gpu, fixture, where):
    bad = []
    for fid, d in fixture.items():
        v = gpu.field(int(fid))
        assert v is not None and v.size == d["n"], f"{where} field {fid}: size"
        h = hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
        if h != d["sha256"]:
            linf = float(np.nanmax(np.abs(v.astype(np.float64))))
            bad.append(f"{where} field {fid}: digest differs (max|v| {linf!r} vs reference {d['linf']!r})")
    return bad


def _solves(log):
    """(status, iterations) per project() line of the reference's stdout."""
    out = []
    for line in log.splitlines():
        m = re.match(r"Exiting solver after (\d+) iterations", line)
        if m:
            out.append((0, int(m.group(1))))
            continue
        m = re.match(r"Exceeded budget of (\d+) iterations", line)
        if m:
            out.append((2, int(m.group(1))))
            continue
        if line.startswith("Initial guess sufficiently small"):
            out.append((1, 0))
    return out


def test_config1_degenerate_inflow_stays_zero():
    import ifl_fluids as F
    import ifl_scenes as S
    n = 128
    sc = S.config1_scene()
    gpu = F.FluidSolver(3, n, n, sc["params"], [], dot_order=F.DOT_SEQUENTIAL)
    orc = ol.OracleSolver(3, n, n, sc["params"], [], fresh=True)
    fields = ol.state_fields(3)
    for k in range(64):
        for inf in sc["inflows"]:
            gpu.add_inflow(*inf)
            orc.add_inflow(*inf)
        st = gpu.update(sc["dt"])
        orc.update(sc["dt"])
        assert st.pressure.status == F.SOLVE_INITIAL_SMALL, f"step {k}: a solve ran"
    # the matrix and its preconditioner depend on the grid only; every
    # quantity and every PCG vector stays zero
    zero = [ol.qfield(q, k) for q in (ol.Q_D, ol.Q_U, ol.Q_V) for k in (ol.QF_SRC, ol.QF_DST)] + \
        [ol.F_R, ol.F_P, ol.F_Z, ol.F_S]
    for f in zero:
        v = gpu.field(f)
        assert not np.any(v), f"field {f} is not zero after 64 steps"
    bad = compare_fields(orc, gpu, fields, where="config 1 after 64 steps")
    assert not bad, "\n".join(bad[:8])
    gpu.close()
    orc.close()


def test_config2_static_256_bit_exact():
    import ifl_fluids as F
    import ifl_scenes as S
    n = 256  # at 128x128 the bodies enclose a fluid pocket and the reference's
    # pressure solve turns NaN, after which its advection reads out of bounds
    sc = S.config2_scene(static=True)
    gpu = _solver(F, 6, n, sc, dot_order=F.DOT_SEQUENTIAL)
    orc = ol.OracleSolver(6, n, n, sc["params"], sc["bodies"], fresh=True)
    fields = ol.state_fields(6)
    for k in range(8):
        for inf in sc["inflows"]:
            gpu.add_inflow(*inf)
            orc.add_inflow(*inf)
        st = gpu.update(sc["dt"])
        orc.log_clear()
        orc.update(sc["dt"])
        bad = compare_fields(orc, gpu, fields, where=f"config 2 256x256 step {k}")
        assert not bad, "\n".join(bad[:8])
        assert F.reference_log_line(6, st.heat) + F.reference_log_line(6, st.pressure) == orc.log()
    gpu.close()
    orc.close()


@pytest.mark.parametrize("variant", ["static", "moving"])
def test_config2_1024_first_step_matches_reference(variant):
    import ifl_fluids as F
    import ifl_scenes as S
    fx = _load("config2_1024_step1.json")
    ref = fx[variant]
    n = fx["grid"][0]
    sc = S.config2_scene(static=variant == "static")
    gpu = _solver(F, 6, n, sc, dot_order=F.DOT_SEQUENTIAL)
    for inf in sc["inflows"]:
        gpu.add_inflow(*inf)
    st = gpu.update(sc["dt"])
    got = [(st.heat.status, st.heat.iterations), (st.pressure.status, st.pressure.iterations)]
    want = _solves(ref["log"])
    print(f"config 2 {variant} 1024x1024: solves {got} (reference {want})", flush=True)
    assert got == want
    log = F.reference_log_line(6, st.heat) + F.reference_log_line(6, st.pressure)
    assert log == ref["log"]
    bad = _check_digests(gpu, ref["fields"], f"config 2 {variant}")
    assert not bad, "\n".join(bad[:8])
    gpu.close()


def test_config3_2048_capped_step_matches_reference():
    import ifl_fluids as F
    import ifl_scenes as S
    fx = _load("config3_2048_capped.json")
    n = fx["grid"][0]
    sc = S.config3_scene()
    gpu = _solver(F, 7, n, sc, dot_order=F.DOT_SEQUENTIAL, pcg_limit=fx["cap"])
    for inf in sc["inflows"]:
        gpu.add_inflow(*inf)
    st = gpu.update(sc["dt"])
    log = F.reference_log_line(7, st.heat) + F.reference_log_line(7, st.pressure)
    print(f"config 3 2048x2048, budgets capped at {fx['cap']}: {log!r}", flush=True)
    assert log == fx["log"]
    bad = _check_digests(gpu, fx["fields"], "config 3 capped")
    assert not bad, "\n".join(bad[:8])
    gpu.close()
