"""The oracle's FLIP branches that the reference's own scenes never reach,
pinned to the reference build (solver 8 compiled from its sources) on the
same constructed states the GPU tests use (tests/test_gpu_flip.py):

* a particle-free hole: EMPTY cells that touch each other, most without a
  fluid neighbour, so the LIFO order of extrapolate's stack decides the
  values (8:611-651), next to a crowded cell fed from the tail of the
  particle array (pruneParticles' swap-with-last, 8:766-786);
* a particle-free ring around a body (EMPTY cells whose processed
  neighbours are solid: the 0/0 average, 8:543);
* seeding up to the particle limit (8:795-796);
* seeding with fewer particles than cells (the rejection test of 8:802 reads
  particle `idx`, a slot the pass itself may write).

Both solvers are created fresh (frand() starts from 0xBA5EBA11 in each and
draws in the same order), warmed with the same steps, edited identically,
and then every state array, the particle count and the following frand()
draws must agree bit for bit.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(8), reason="reference build (oracle/_ref) not present")

MULTI = [(ol.BODY_BOX, 0.5, 0.6, 0.3, 0.1, 0.7, 0.0, 0.0, 0.0),
         (ol.BODY_SPHERE, 0.3, 0.35, 0.15, 0.15, 0.0, 0.0, 0.0, 0.0),
         (ol.BODY_BOX, 0.7, 0.3, 0.1, 0.12, -0.4, 0.0, 0.0, 0.0)]
PARTICLES = [ol.F_POSX, ol.F_POSY, ol.F_PROP_D, ol.F_PROP_T, ol.F_PROP_U, ol.F_PROP_V]
ALL = ol.state_fields(8) + [ol.F_WEIGHT] + [ol.qfield(q, ol.QF_DST) for q in range(4)]


def warm_pair(w, h, bodies, steps):
    sc = dict(ol.REF_SCENES[8])
    if bodies is not None:
        sc["bodies"] = bodies
    ref = ol.RefSolver(8, w, h, sc["params"], sc["bodies"], fresh=True)
    orc = ol.OracleSolver(8, w, h, sc["params"], sc["bodies"], fresh=True)
    for _ in range(steps):
        ref.update(sc["dt"])
        orc.update(sc["dt"])
    return ref, orc, sc


def same(ref, orc, fields, where):
    bad = []
    for f in fields:
        a, b = ref.field(f), orc.field(f)
        if a is None and b is None:
            continue
        if a is None or b is None or a.tobytes() != b.tobytes():
            bad.append(f)
    assert not bad, f"{where}: fields {bad} differ"
    assert ref.particle_count() == orc.particle_count(), f"{where}: particle count"


def same_frand(ref, n=16):
    a = [ref.lib.ref_frand() for _ in range(n)]
    b = [ol.oracle_lib().orc_frand() for _ in range(n)]
    assert a == b, "frand() streams differ"


def both(ref, orc, stage, iarg=0, darg=0.0):
    ref.stage(stage, iarg, darg)
    orc.stage(stage, iarg, darg)


def p2g_both(ref, orc):
    for q in range(4):
        both(ref, orc, "FILL_SOLID", q)
    both(ref, orc, "P2G")


def test_hole_and_crowded_cell():
    ref, orc, sc = warm_pair(64, 48, None, 2)
    same(ref, orc, ALL, "warm")
    for s in (ref, orc):
        n = s.particle_count()
        px, py = s.view(ol.F_POSX), s.view(ol.F_POSY)
        sel = np.nonzero((px[:n] >= 10) & (px[:n] < 22) & (py[:n] >= 20) & (py[:n] < 29))[0]
        assert len(sel) > 200
        px[sel] = 40.25 + 0.5 * (np.arange(len(sel)) % 7) / 7
        py[sel] = 8.5
        tail = np.arange(n - 30, n)
        px[tail] = 5.5
        py[tail] = 30.5
    # particlesToGrid (8:916-927) stage by stage
    for q in range(4):
        both(ref, orc, "FILL_SOLID", q)
    for q in range(4):
        both(ref, orc, "FROM_PARTICLES", q)
        both(ref, orc, "EXTRAPOLATE", q)
        same(ref, orc, ALL, f"fromParticles and extrapolate of quantity {q} over the hole")
    # the hole really is a connected patch of particle-free nodes (EMPTY
    # before extrapolate turns them FLUID again, 8:630): zero weights of the
    # last grid transferred (v, 64 x 49 nodes)
    wgt = orc.field(ol.F_WEIGHT)[:64 * 49].reshape(49, 64)
    assert np.count_nonzero(wgt[21:28, 11:21] == 0) > 40
    both(ref, orc, "COUNT_PARTICLES")
    counts = orc.field(ol.F_COUNTS)
    assert counts.max() > 12  # the crowded cells
    both(ref, orc, "PRUNE")
    same(ref, orc, PARTICLES + [ol.F_COUNTS], "prune of the crowded cells")
    both(ref, orc, "SEED")
    same(ref, orc, PARTICLES, "seed into the hole")
    ref.update(sc["dt"])
    orc.update(sc["dt"])
    same(ref, orc, ALL, "step after the hole")
    same_frand(ref)


def test_hole_next_to_a_body():
    ref, orc, sc = warm_pair(64, 64, MULTI, 1)
    for s in (ref, orc):
        n = s.particle_count()
        px, py = s.view(ol.F_POSX), s.view(ol.F_POSY)
        idx = np.nonzero((px[:n] > 18) & (px[:n] < 46) & (py[:n] > 31) & (py[:n] < 46))[0]
        px[idx] = 60.5
        py[idx] = 60.5
    p2g_both(ref, orc)
    same(ref, orc, ALL, "particlesToGrid around the body")
    ref.update(sc["dt"])
    orc.update(sc["dt"])
    same(ref, orc, ALL, "step after the ring")
    same_frand(ref)


def test_seed_reaches_particle_limit():
    w, h = 24, 20
    ref, orc, sc = warm_pair(w, h, None, 1)
    mp = w * h * 12
    rng = np.random.default_rng(5)
    px = rng.uniform(0, w - 0.001, mp)
    py = rng.uniform(0, h - 0.001, mp)
    counts = rng.integers(0, 5, w * h).astype(np.int32)
    for s in (ref, orc):
        s.view(ol.F_POSX)[:] = px
        s.view(ol.F_POSY)[:] = py
        s.particle_count(mp - 7)
        s.view(ol.F_COUNTS)[:] = counts
    both(ref, orc, "SEED")
    same(ref, orc, PARTICLES, "seed at the limit")
    assert orc.particle_count() == mp
    same_frand(ref)


def test_seed_with_fewer_particles_than_cells():
    w, h = 24, 20
    ref, orc, sc = warm_pair(w, h, None, 1)
    counts = np.random.default_rng(6).integers(0, 4, w * h).astype(np.int32)
    for s in (ref, orc):
        s.particle_count(150)
        s.view(ol.F_COUNTS)[:] = counts
    both(ref, orc, "SEED")
    same(ref, orc, PARTICLES, "seed, few particles")
    assert orc.particle_count() > 150
    same_frand(ref)
