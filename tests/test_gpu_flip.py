"""Solver 8 (FLIP) on the GPU against the oracle, bit for bit.

* particle creation: initParticles (8:735-751) draws from frand(), whose
  process-wide seed both sides start from 0xBA5EBA11;
* every particle stage teacher-forced from the oracle's state after a few
  steps of the reference scene (8:1469-1513), including constructed cases
  for the order-dependent parts: a hole with no particles (EMPTY cells that
  touch each other and have no fluid neighbour, 8:503-511 / 8:558-565), a
  crowded cell (pruneParticles' swap-with-last, 8:766-786), the particle
  limit during seeding (8:795-796), fewer particles than cells (the
  rejection test of 8:802 then reads particles the pass writes) and bodies
  deep enough to trigger backProject (8:828);
* free-running update() trajectories with the reference's dot order.
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_util import compare_fields, make_product

pytestmark = pytest.mark.gpu

SEED0 = 0xBA5EBA11
PARTICLES = [ol.F_POSX, ol.F_POSY, ol.F_PROP_D, ol.F_PROP_T, ol.F_PROP_U, ol.F_PROP_V]
MULTI = [(ol.BODY_BOX, 0.5, 0.6, 0.3, 0.1, 0.7, 0.0, 0.0, 0.0),
         (ol.BODY_SPHERE, 0.3, 0.35, 0.15, 0.0, 0.0, 0.0, 0.0, 0.0),
         (ol.BODY_BOX, 0.7, 0.3, 0.1, 0.12, -0.4, 0.0, 0.0, 0.0)]
MOVING = [(ol.BODY_SPHERE, 0.55, 0.5, 0.2, 0.2, 0.3, 0.1, -0.2, 1.0)]


def grid_fields():
    out = []
    for q in range(4):
        out += [ol.qfield(q, f) for f in (ol.QF_SRC, ol.QF_DST, ol.QF_CELL, ol.QF_BODY, ol.QF_NORMALX,
                                          ol.QF_NORMALY, ol.QF_MASK, ol.QF_VOLUME, ol.QF_PHI)]
    return out


ALL = ol.state_fields(8) + [ol.F_WEIGHT] + [ol.qfield(q, ol.QF_DST) for q in range(4)]


def pair(w, h, bodies=None, dot_order=1):
    import ifl_fluids as F
    sc = dict(ol.REF_SCENES[8])
    if bodies is not None:
        sc["bodies"] = bodies
    F.frand_set_state(SEED0)
    orc = ol.OracleSolver(8, w, h, sc["params"], sc["bodies"], fresh=True)
    gpu = make_product(8, w, h, sc["params"], sc["bodies"], dot_order=dot_order)
    return orc, gpu, sc


def orc_rng():
    return ol.oracle_lib().orc_frand_state()


def check(orc, gpu, fields, where):
    import ifl_fluids as F
    bad = compare_fields(orc, gpu, fields, where=where)
    if orc.particle_count() != gpu.particle_count():
        bad.append(f"{where}: particle count {orc.particle_count()} vs {gpu.particle_count()}")
    if orc_rng() != F.frand_state():
        bad.append(f"{where}: frand state {orc_rng():#x} vs {F.frand_state():#x}")
    assert not bad, "\n".join(bad[:8])


def push_state(orc, gpu):
    """Teacher forcing: the product takes the oracle's whole state."""
    import ifl_fluids as F
    for f in ALL + grid_fields():
        v = orc.field(f)
        if v is not None:
            gpu.set_field(f, v)
    gpu.particle_count(orc.particle_count())
    F.frand_set_state(orc_rng())


def warm(w, h, bodies, steps):
    sc = dict(ol.REF_SCENES[8])
    if bodies is not None:
        sc["bodies"] = bodies
    orc = ol.OracleSolver(8, w, h, sc["params"], sc["bodies"], fresh=True)
    for _ in range(steps):
        orc.update(sc["dt"])
    return orc, sc


@pytest.mark.parametrize("w,h,bodies", [(48, 40, None), (64, 64, MULTI), (33, 17, [])])
def test_init_particles(w, h, bodies):
    """initParticles + gridToParticles(1.0) (8:735-751, 8:1314)."""
    orc, gpu, _ = pair(w, h, bodies)
    check(orc, gpu, PARTICLES + [ol.qfield(ol.Q_T, ol.QF_SRC)], "after construction")
    print(f"{w}x{h}: {gpu.particle_count()} particles", flush=True)
    gpu.close()
    orc.close()


def stage_both(orc, gpu, stage, iarg=0, darg=0.0):
    orc.stage(stage, iarg, darg)
    gpu.stage(stage, iarg, darg)


@pytest.mark.parametrize("w,h,bodies,warm_steps", [(48, 40, None, 3), (64, 64, MULTI, 2)])
def test_particle_stages_teacher_forced(w, h, bodies, warm_steps):
    orc, sc = warm(w, h, bodies, warm_steps)
    import ifl_fluids as F
    gpu = make_product(8, w, h, sc["params"], sc["bodies"], dot_order=1)
    push_state(orc, gpu)
    check(orc, gpu, ALL, "pushed state")
    for q in range(4):
        stage_both(orc, gpu, "FILL_SOLID", q)
    for q in range(4):
        stage_both(orc, gpu, "FROM_PARTICLES", q)
        check(orc, gpu, [ol.qfield(q, ol.QF_SRC), ol.qfield(q, ol.QF_CELL), ol.F_WEIGHT], f"fromParticles q{q}")
        stage_both(orc, gpu, "EXTRAPOLATE", q)
        check(orc, gpu, [ol.qfield(q, f) for f in (ol.QF_SRC, ol.QF_CELL, ol.QF_MASK)], f"extrapolate q{q}")
    stage_both(orc, gpu, "COUNT_PARTICLES")
    check(orc, gpu, [ol.F_COUNTS], "countParticles")
    stage_both(orc, gpu, "PRUNE")
    check(orc, gpu, PARTICLES + [ol.F_COUNTS], "pruneParticles")
    stage_both(orc, gpu, "SEED")
    check(orc, gpu, PARTICLES, "seedParticles")
    for q in range(4):
        stage_both(orc, gpu, "COPY_OLD", q)
    for q in range(4):
        stage_both(orc, gpu, "DIFF", q, 0.001)
    check(orc, gpu, [ol.qfield(q, f) for q in range(4) for f in (ol.QF_SRC, ol.QF_DST)], "diff")
    stage_both(orc, gpu, "G2P", 0, 0.001)
    check(orc, gpu, PARTICLES, "gridToParticles")
    for q in range(4):
        stage_both(orc, gpu, "UNDIFF", q, 0.001)
    check(orc, gpu, [ol.qfield(q, ol.QF_SRC) for q in range(4)], "undiff")
    stage_both(orc, gpu, "PARTICLE_ADVECT", 0, sc["dt"])
    check(orc, gpu, PARTICLES, "advect")
    # the whole particlesToGrid on the advected particles
    for q in range(4):
        stage_both(orc, gpu, "FILL_SOLID", q)
    stage_both(orc, gpu, "P2G")
    check(orc, gpu, ALL, "particlesToGrid")
    gpu.close()
    orc.close()
    assert F.frand_state() == orc_rng()


def _constructed(w, h, bodies, warm_steps, edit):
    """Warm the oracle, edit its particles, push to the product, then run
    particlesToGrid on both and compare the whole state."""
    orc, sc = warm(w, h, bodies, warm_steps)
    edit(orc)
    gpu = make_product(8, w, h, sc["params"], sc["bodies"], dot_order=1)
    push_state(orc, gpu)
    for q in range(4):
        stage_both(orc, gpu, "FILL_SOLID", q)
    stage_both(orc, gpu, "P2G")
    check(orc, gpu, ALL, "particlesToGrid")
    return orc, gpu, sc


def test_hole_and_crowded_cell():
    """Particles in a 12x9-cell block are moved into one cell: the block
    becomes EMPTY cells that touch each other, most without a fluid
    neighbour (the stack order matters), and the target cell holds far more
    than 12 particles (swap-with-last pruning), with some of them in the
    tail of the array."""
    def edit(orc):
        n = orc.particle_count()
        px, py = orc.view(ol.F_POSX), orc.view(ol.F_POSY)
        sel = np.nonzero((px[:n] >= 10) & (px[:n] < 22) & (py[:n] >= 20) & (py[:n] < 29))[0]
        assert len(sel) > 200
        px[sel] = 40.25 + 0.5 * (np.arange(len(sel)) % 7) / 7
        py[sel] = 8.5
        # a crowded cell fed from the end of the array as well
        tail = np.arange(n - 30, n)
        px[tail] = 5.5
        py[tail] = 30.5
    orc, gpu, sc = _constructed(64, 48, None, 2, edit)
    # and the step after it
    orc.update(sc["dt"])
    gpu.update(sc["dt"])
    check(orc, gpu, ALL, "step after the hole")
    gpu.close()
    orc.close()


def test_hole_next_to_a_body():
    """A particle-free ring around a body: EMPTY cells whose only processed
    neighbours are solid (average over no fluid cells, 8:543)."""
    def edit(orc):
        n = orc.particle_count()
        px, py = orc.view(ol.F_POSX), orc.view(ol.F_POSY)
        # box of MULTI[0] spans roughly x 0.35..0.65, y 0.55..0.65 at 64x64
        near = (px[:n] > 18) & (px[:n] < 46) & (py[:n] > 31) & (py[:n] < 46)
        idx = np.nonzero(near)[0]
        px[idx] = 60.5
        py[idx] = 60.5
    orc, gpu, sc = _constructed(64, 64, MULTI, 1, edit)
    gpu.close()
    orc.close()


def test_seed_reaches_particle_limit():
    """seedParticles stops when the count reaches 12 per cell (8:795-796)."""
    import ifl_fluids as F
    w, h = 24, 20
    orc, sc = warm(w, h, None, 1)
    gpu = make_product(8, w, h, sc["params"], sc["bodies"], dot_order=1)
    mp = w * h * 12
    rng = np.random.default_rng(5)
    px = orc.view(ol.F_POSX)
    py = orc.view(ol.F_POSY)
    px[:] = rng.uniform(0, w - 0.001, mp)
    py[:] = rng.uniform(0, h - 0.001, mp)
    orc.particle_count(mp - 7)
    counts = orc.view(ol.F_COUNTS)
    counts[:] = rng.integers(0, 5, w * h)
    push_state(orc, gpu)
    stage_both(orc, gpu, "SEED")
    check(orc, gpu, PARTICLES, "seed at the limit")
    assert gpu.particle_count() == mp
    gpu.close()
    orc.close()


def test_seed_with_fewer_particles_than_cells():
    """With fewer particles than cells the rejection test of 8:802 reads
    slots the pass itself writes; the pass is replayed in order."""
    w, h = 24, 20
    orc, sc = warm(w, h, None, 1)
    gpu = make_product(8, w, h, sc["params"], sc["bodies"], dot_order=1)
    orc.particle_count(150)
    counts = orc.view(ol.F_COUNTS)
    counts[:] = np.random.default_rng(6).integers(0, 4, w * h)
    push_state(orc, gpu)
    stage_both(orc, gpu, "SEED")
    check(orc, gpu, PARTICLES, "seed, few particles")
    gpu.close()
    orc.close()


def test_back_project_in_a_large_body():
    """backProject (8:816-839) acts only deeper than one world unit inside a
    body, so it needs a body larger than the domain."""
    w, h = 32, 32
    big = [(ol.BODY_SPHERE, 0.5, 0.5, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0)]
    orc, gpu, sc = pair(w, h, big)
    rng = np.random.default_rng(7)
    for q, f in ((ol.Q_U, ol.qfield(ol.Q_U, ol.QF_SRC)), (ol.Q_V, ol.qfield(ol.Q_V, ol.QF_SRC))):
        v = rng.normal(0, 0.5, orc.field(f).shape)
        orc.set_field(f, v)
        gpu.set_field(f, v)
    n = w * h * 4
    px = rng.uniform(0, w - 0.001, n)
    py = rng.uniform(0, h - 0.001, n)
    for o in (orc, gpu):
        o.set_field(ol.F_POSX, np.concatenate([px, np.zeros(w * h * 12 - n)]))
        o.set_field(ol.F_POSY, np.concatenate([py, np.zeros(w * h * 12 - n)]))
        o.particle_count(n)
    stage_both(orc, gpu, "PARTICLE_ADVECT", 0, 0.01)
    check(orc, gpu, [ol.F_POSX, ol.F_POSY], "advect with back-projection")
    gpu.close()
    orc.close()


@pytest.mark.parametrize("w,h,bodies,steps", [
    (48, 40, None, 12),
    (64, 64, MULTI, 8),
    (72, 40, MOVING, 8),  # wide: a tall grid makes the reference read _vDensity past its end (7:700)
])
def test_update_bit_exact_sequential_dots(w, h, bodies, steps):
    orc, gpu, sc = pair(w, h, bodies)
    for k in range(steps):
        orc.log_clear()
        orc.update(sc["dt"])
        st = gpu.update(sc["dt"])
        check(orc, gpu, ALL, f"step {k}")
        assert f"Particle count: {st.particle_count}\n" in orc.log()
        if (k + 1) % sc["substeps"] == 0:
            orc.bodies_update(sc["dt"])
            gpu.bodies_update(sc["dt"])
    gpu.close()
    orc.close()


def test_update_256():
    """Two steps at 256x256 (262k particles), bit for bit."""
    orc, gpu, sc = pair(256, 256, None)
    for k in range(2):
        orc.update(sc["dt"])
        st = gpu.update(sc["dt"])
        check(orc, gpu, ALL, f"256x256 step {k}")
        print(f"step {k}: {st.particle_count} particles, iterations heat {st.heat.iterations} "
              f"pressure {st.pressure.iterations}", flush=True)
    gpu.close()
    orc.close()


def test_variant_alpha_and_cubic_g2p():
    """The parameterised FLIP variant, not the reference: alpha 0.05 (a
    "0.95 FLIP" blend, 8:1401) and the FLIP update's gridToParticles sampling
    with the Catmull-Rom cerp of solvers 3-7 (7:404-422) instead of lerp
    (flip_g2p = 1), against the oracle's restatement of the same variant, bit
    for bit over 8 steps with bodies; and it really differs from the
    reference's form after one step."""
    import numpy as np
    import ifl_fluids as F
    sc = dict(ol.REF_SCENES[8])
    sc["bodies"] = MULTI
    F.frand_set_state(SEED0)
    orc = ol.OracleSolver(8, 64, 64, sc["params"], sc["bodies"], fresh=True)
    orc.set_flip(0.05, cubic=True)
    gpu = make_product(8, 64, 64, sc["params"], sc["bodies"], dot_order=1, flip_alpha=0.05, flip_g2p=1)
    for k in range(8):
        orc.update(sc["dt"])
        gpu.update(sc["dt"])
        check(orc, gpu, ALL, f"variant step {k}")
        if k == 0:
            state = F.frand_state()
            F.frand_set_state(SEED0)
            std = make_product(8, 64, 64, sc["params"], sc["bodies"], dot_order=1)
            std.update(sc["dt"])
            assert not np.array_equal(std.field(ol.F_PROP_U), gpu.field(ol.F_PROP_U))
            std.close()
            F.frand_set_state(state)
        if (k + 1) % sc["substeps"] == 0:
            orc.bodies_update(sc["dt"])
            gpu.bodies_update(sc["dt"])
    gpu.close()
    orc.close()
