"""Solver 8 in the reduced-precision form: the particle state is float and
fromParticles (8:663-688) is an LDS-binned scatter accumulated in single
precision, checked against the oracle's double-precision transfer on the
same particles (teacher-forced from the oracle's state after a few steps of
the reference scene, the oracle's particles rounded to float first so that
both transfer the same values). The sums are
taken in no fixed order, so the comparison is to a tolerance: per node
weight within 2e-6 of the oracle's (relative to max(1, weight)), every
transferred field within 4e-6 of its largest magnitude, and the EMPTY
marking identical (a node's single-precision weight is zero exactly when
the double weight is)."""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import make_product
from test_gpu_flip import MULTI, PARTICLES, push_state, warm

pytestmark = pytest.mark.gpu


def _close(name, a, b, tol):
    scale = max(np.abs(a).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    print(f"  {name}: max |diff| / max |field| = {err:.2e}", flush=True)
    assert np.isfinite(b).all() and err <= tol, name


@pytest.mark.parametrize("w,h,bodies,warm_steps", [(48, 40, None, 3), (64, 64, MULTI, 2), (200, 136, MULTI, 1)])
def test_from_particles_single_precision(w, h, bodies, warm_steps):
    import ifl_fluids as F
    orc, sc = warm(w, h, bodies, warm_steps)
    mix = make_product(8, w, h, sc["params"], sc["bodies"], precision=F.PREC_MIXED)
    for f in PARTICLES:
        assert mix.field_info(f)[1] == 4, "the reduced-precision form keeps the particles in float"
        orc.set_field(f, orc.field(f).astype(np.float32))
    push_state(orc, mix)
    for f in PARTICLES:   # the float state holds the rounded values exactly
        n = orc.particle_count()
        assert (mix.field(f)[:n] == orc.field(f)[:n]).all()
    for q in range(4):
        orc.stage("FILL_SOLID", q)
        mix.stage("FILL_SOLID", q)
    print(f"{w}x{h}: {orc.particle_count()} particles", flush=True)
    for q in (ol.Q_D, ol.Q_U, ol.Q_V):   # d carries t (one grid, one pass)
        orc.stage("FROM_PARTICLES", q)
        mix.stage("FROM_PARTICLES", q)
        wo, wm = orc.field(ol.F_WEIGHT), mix.field(ol.F_WEIGHT)
        assert ((wo == 0) == (wm == 0)).all(), f"q{q}: zero weights differ"
        werr = (np.abs(wo - wm) / np.maximum(1.0, wo)).max()
        print(f"q{q}: weight max rel diff {werr:.2e}", flush=True)
        assert werr <= 2e-6
        qs = [q, ol.Q_T] if q == ol.Q_D else [q]
        for qq in qs:
            co, cm = orc.field(ol.qfield(qq, ol.QF_CELL)), mix.field(ol.qfield(qq, ol.QF_CELL))
            assert (co == cm).all(), f"q{qq}: cell flags (EMPTY marking) differ"
            _close(f"q{qq} src", orc.field(ol.qfield(qq, ol.QF_SRC)), mix.field(ol.qfield(qq, ol.QF_SRC)), 4e-6)
    mix.close()
    orc.close()


def test_flip_step_against_fp64():
    """One update() of the reference scene at 128x128 in both forms from the
    same state (teacher-forced): every grid field of the reduced-precision
    form close to the FP64 drop-in's; per field the relative difference is
    printed and held to 1e-4."""
    import ifl_fluids as F
    orc, sc = warm(128, 128, None, 3)
    runs = {}
    for prec in (F.PREC_FP64, F.PREC_MIXED):
        gpu = make_product(8, 128, 128, sc["params"], sc["bodies"], precision=prec)
        push_state(orc, gpu)
        st = gpu.update(sc["dt"])
        assert st.pressure.status == F.SOLVE_CONVERGED
        runs[prec] = ({q: gpu.field(ol.qfield(q, ol.QF_SRC)) for q in range(4)}, st.particle_count)
        gpu.close()
    (a, na), (b, nb) = runs[F.PREC_FP64], runs[F.PREC_MIXED]
    print(f"particles {na} / {nb}", flush=True)
    for q in range(4):
        _close(f"q{q}", a[q], b[q], 1e-4)
    orc.close()
