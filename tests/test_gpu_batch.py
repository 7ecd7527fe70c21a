"""Independent solver instances stepped concurrently from several host
threads (each handle has its own HIP stream; the library releases the GIL in
every call): every instance must still match its own oracle bit for bit.
Instances follow the batched workload's draws (solver 4, sides 64..192 here
so the oracle stays quick)."""
import threading

import pytest

import oracle_lib as ol
from gpu_util import compare_fields

pytestmark = pytest.mark.gpu


def test_concurrent_instances_bit_exact():
    import ifl_fluids as F
    import ifl_scenes as S
    insts = [i for i in S.config4_instances(64) if max(i["w"], i["h"]) <= 192][:6]
    assert len(insts) == 6
    steps = 4
    gpus = [F.FluidSolver(4, i["w"], i["h"], i["params"], [F.SolidBox(*b[1:]) for b in i["bodies"]], dot_order=1)
            for i in insts]
    errors = []

    def run(ks):
        try:
            for _ in range(steps):
                for k in ks:
                    for inf in insts[k]["inflows"]:
                        gpus[k].add_inflow(*inf)
                    gpus[k].update(insts[k]["dt"])
        except Exception as e:
            errors.append(repr(e))

    th = [threading.Thread(target=run, args=([k for k in range(6) if k % 3 == t],)) for t in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k, inst in enumerate(insts):
        orc = ol.OracleSolver(4, inst["w"], inst["h"], inst["params"], inst["bodies"])
        for _ in range(steps):
            for inf in inst["inflows"]:
                orc.add_inflow(*inf)
            orc.update(inst["dt"])
        bad = compare_fields(orc, gpus[k], ol.state_fields(4), where=f"instance {k} {inst['w']}x{inst['h']}")
        assert not bad, "\n".join(bad[:6])
        print(f"instance {k} {inst['w']}x{inst['h']}: bit-exact after {steps} steps", flush=True)
        orc.close()
        gpus[k].close()


@pytest.mark.parametrize("level", [3, 4, 5, 6, 7])
def test_fused_batch_equals_single_solvers(level):
    """ifl_batch_update runs every stage of update() as one launch for all
    instances and batches the PCG iterations of their solves; each instance
    must end exactly where the same instance stepped on its own ends (the
    same kernel bodies on the same arrays, the same reduction order per
    instance), while the instances converge after different iteration
    counts. Instance 1 has no body (its fill tables are empty); from level 5
    on the others' boxes move and turn, advanced between steps as the
    reference's main loop does."""
    import numpy as np
    import ifl_fluids as F
    import ifl_scenes as S
    insts = [i for i in S.config4_instances(64) if max(i["w"], i["h"]) <= 192][:6]
    params = (0.1,) if level <= 5 else (0.1, 1.0, 0.01)
    inflow = lambda i: i["inflows"][0] if level <= 5 else i["inflows"][0][:5] + (600.0, 0.0, 0.0)

    def inflows(i):
        # the second entry negates d over the same rectangle: addInflow keeps
        # the value of larger magnitude, the first on a tie, so the entries'
        # order decides the result
        e = tuple(inflow(i))
        return [e, e[:4] + (-e[4],) + e[5:]]

    def make(k, i):
        if level == 3 or k == 1:
            bodies = []
        else:
            # (x, y, sx, sy, theta, vx, vy, vtheta); moving from level 5 on
            vel = (0.08, -0.05, 0.6) if level >= 5 else (0.0, 0.0, 0.0)
            bodies = [F.SolidBox(*(tuple(b[1:6]) + vel)) for b in i["bodies"]]
        return F.FluidSolver(level, i["w"], i["h"], params, bodies)
    steps = 3
    batched = [make(k, i) for k, i in enumerate(insts)]
    batch = F.Batch(batched)
    iters_b, stat_b = [], []
    for _ in range(steps):
        batch.add_inflow([(k,) + e for k, i in enumerate(insts) for e in inflows(i)])
        st = batch.update(insts[0]["dt"])
        for s in batched:
            s.bodies_update(insts[0]["dt"])
        iters_b.append([(x.heat.iterations, x.pressure.iterations) for x in st])
        stat_b.append([(x.heat.status, x.pressure.status, x.border_solid, x.nonfinite) for x in st])
    batch.close()
    fields = ol.state_fields(level)
    for k, i in enumerate(insts):
        single = make(k, i)
        its, stat = [], []
        for _ in range(steps):
            for e in inflows(i):
                single.add_inflow(*e)
            st = single.update(i["dt"])
            single.bodies_update(i["dt"])
            its.append((st.heat.iterations, st.pressure.iterations))
            stat.append((st.heat.status, st.pressure.status, st.border_solid, st.nonfinite))
        assert its == [row[k] for row in iters_b]
        assert stat == [row[k] for row in stat_b]
        bad = compare_fields(single, batched[k], fields, where=f"instance {k} {i['w']}x{i['h']}")
        assert not bad, "\n".join(bad[:6])
        single.close()
        batched[k].close()
    counts = sorted(set(row[k][1] for row in iters_b for k in range(len(insts))))
    print(f"level {level}: (heat, pressure) iterations per step and instance {iters_b}; statuses {stat_b}",
          flush=True)
    assert len(counts) > 1  # the instances really converge at different iterations


def test_fused_batch_flip_16x256():
    """Solver 8 in the batch: 16 instances of the reference's FLIP scene at
    256x256 (about 4.2M particles in all) stepped as one batch, against the
    same 16 solvers stepped one after the other from the same frand() state
    (the batch draws in instance order within each step): identical
    statuses, iteration counts, particle counts, grid fields and particles."""
    import ifl_fluids as F
    import ifl_scenes as S
    sc = S.config5_scene()
    n, size, steps = 16, 256, 2

    def make():
        return F.FluidSolver(8, size, size, sc["params"], [F.SolidBox(*b[1:]) for b in sc["bodies"]])
    seed = 0xBA5EBA11
    F.frand_set_state(seed)
    batched = [make() for _ in range(n)]
    batch = F.Batch(batched)
    rows_b = []
    for _ in range(steps):
        st = batch.update(sc["dt"])
        rows_b.append([(x.heat.iterations, x.pressure.iterations, x.heat.status, x.pressure.status,
                        x.particle_count) for x in st])
    batch.close()
    state_b = F.frand_state()
    F.frand_set_state(seed)
    singles = [make() for _ in range(n)]
    rows_s = []
    for _ in range(steps):
        row = []
        for s in singles:
            x = s.update(sc["dt"])
            row.append((x.heat.iterations, x.pressure.iterations, x.heat.status, x.pressure.status,
                        x.particle_count))
        rows_s.append(row)
    assert rows_s == rows_b
    assert F.frand_state() == state_b
    fields = ol.state_fields(8)
    for k in range(n):
        bad = compare_fields(singles[k], batched[k], fields, where=f"instance {k}")
        assert not bad, "\n".join(bad[:6])
        singles[k].close()
        batched[k].close()
    print(f"16 x 256^2 FLIP batch: (heat, pressure, statuses, particles) per step {rows_b[-1][:3]} ...", flush=True)


def test_fused_batch_full_draw():
    """Config 4's whole draw as one batch: 256 solver-4 instances with sides
    from {64 .. 512} (14.5M cells), two steps. Every solve converges and
    every field stays finite; 8 instances spread over the sizes, stepped on
    their own, end exactly where their batched copies do; and the 4 smallest
    are within 1e-6 (relative to each field's largest magnitude) of their
    oracle, which sums dot products in the reference's order where the
    batch uses tree-ordered sums."""
    import numpy as np
    import ifl_fluids as F
    import ifl_scenes as S
    insts = S.config4_instances(256)
    steps = 2

    def make(i):
        return F.FluidSolver(4, i["w"], i["h"], i["params"], [F.SolidBox(*b[1:]) for b in i["bodies"]])
    batched = [make(i) for i in insts]
    batch = F.Batch(batched)
    rows = []
    for _ in range(steps):
        for s, i in zip(batched, insts):
            for inf in i["inflows"]:
                s.add_inflow(*inf)
        st = batch.update(insts[0]["dt"])
        rows.append([(x.pressure.status, x.pressure.iterations) for x in st])
        assert all(x.pressure.status in (F.SOLVE_CONVERGED, F.SOLVE_INITIAL_SMALL) for x in st)
        assert not any(x.nonfinite for x in st)
    batch.close()
    fields = ol.state_fields(4)
    by_size = sorted(range(len(insts)), key=lambda k: insts[k]["w"] * insts[k]["h"])
    for k in by_size[::len(insts) // 8][:8]:
        single = make(insts[k])
        for t in range(steps):
            for inf in insts[k]["inflows"]:
                single.add_inflow(*inf)
            st = single.update(insts[k]["dt"])
            assert (st.pressure.status, st.pressure.iterations) == rows[t][k]
        bad = compare_fields(single, batched[k], fields, where=f"instance {k} {insts[k]['w']}x{insts[k]['h']}")
        assert not bad, "\n".join(bad[:6])
        single.close()
    for k in by_size[:4]:
        i = insts[k]
        orc = ol.OracleSolver(4, i["w"], i["h"], i["params"], i["bodies"])
        for _ in range(steps):
            for inf in i["inflows"]:
                orc.add_inflow(*inf)
            orc.update(i["dt"])
        for f in (ol.qfield(q, ol.QF_SRC) for q in (ol.Q_D, ol.Q_U, ol.Q_V)):
            a, b = orc.field(f), batched[k].field(f)
            err = np.abs(a - b).max() / max(np.abs(a).max(), 1e-300)
            print(f"instance {k} {i['w']}x{i['h']} field {f}: batch vs oracle {err:.2e}", flush=True)
            assert err <= 1e-6
        orc.close()
    for s in batched:
        s.close()
    its = [r[1] for row in rows for r in row]
    print(f"256 instances, {sum(i['w'] * i['h'] for i in insts)} cells: pressure iterations {min(its)}..{max(its)}",
          flush=True)


def test_batch_add_inflow_arguments():
    """ifl_batch_add_inflow rejects a solver index outside the batch (and a
    missing array) with IFL_ERR_ARG before touching any solver; an empty
    call is a no-op."""
    import ctypes
    import numpy as np
    import ifl_fluids as F
    solvers = [F.FluidSolver(4, 64, 48, (0.1,), []) for _ in range(2)]
    batch = F.Batch(solvers)
    before = [s.field(ol.qfield(ol.Q_D, ol.QF_SRC)) for s in solvers]
    for bad in ([(2, 0.4, 0.2, 0.1, 0.05, 1.0, 0.0, 0.0, 0.0)], [(-1, 0.4, 0.2, 0.1, 0.05, 1.0, 0.0, 0.0, 0.0)]):
        with pytest.raises(F.IflError):
            batch.add_inflow(bad)
    rect = (ctypes.c_double * 4)(0.4, 0.2, 0.1, 0.05)
    inst = (ctypes.c_int * 1)(0)
    assert F.lib.ifl_batch_add_inflow(batch._h, 1, inst, rect, None) == 1  # IFL_ERR_ARG
    assert F.lib.ifl_batch_add_inflow(batch._h, 0, None, None, None) == 0
    batch.add_inflow([])
    for s, b in zip(solvers, before):
        assert np.array_equal(s.field(ol.qfield(ol.Q_D, ol.QF_SRC)), b)
    batch.add_inflow([(1, 0.4, 0.2, 0.1, 0.05, 1.0, 0.0, 0.0, 0.0)])
    assert np.array_equal(solvers[0].field(ol.qfield(ol.Q_D, ol.QF_SRC)), before[0])
    assert solvers[1].field(ol.qfield(ol.Q_D, ol.QF_SRC)).max() > 0
    batch.close()
    for s in solvers:
        s.close()


@pytest.mark.parametrize("level", [4, 6])
def test_fused_batch_host_bodies(level):
    """Solvers holding a body evaluated on the host (ifl_set_host_body_ops;
    here a sphere written in numpy) run their stages per solver inside the
    batch: three instances of different sizes, stepped as one batch, end
    exactly where the same solvers stepped one after the other end."""
    import ctypes
    import numpy as np
    import ifl_fluids as F
    dp = ctypes.POINTER(ctypes.c_double)
    DIST = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, dp, dp, dp, ctypes.c_size_t)
    CLOSEST = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, dp, dp, ctypes.c_size_t)
    NORMAL = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, dp, dp, dp, dp, ctypes.c_size_t)

    class Ops(ctypes.Structure):
        _fields_ = [("user", ctypes.c_void_p), ("distance", DIST), ("closest", CLOSEST), ("normal", NORMAL)]

    cx, cy, r = 0.42, 0.5, 0.12

    def view(p, n):
        return np.ctypeslib.as_array(p, shape=(n,))

    def dist(user, body, x, y, out, n):
        if n:
            view(out, n)[:] = np.hypot(view(x, n) - cx, view(y, n) - cy) - r

    def closest(user, body, x, y, n):
        if n:
            X, Y = view(x, n), view(y, n)
            d = np.maximum(np.hypot(X - cx, Y - cy), 1e-300)
            X[:], Y[:] = cx + (X - cx) * (r / d), cy + (Y - cy) * (r / d)

    def normal(user, body, x, y, nx, ny, n):
        if n:
            X, Y = view(x, n), view(y, n)
            d = np.maximum(np.hypot(X - cx, Y - cy), 1e-300)
            view(nx, n)[:], view(ny, n)[:] = (X - cx) / d, (Y - cy) / d

    ops = Ops(None, DIST(dist), CLOSEST(closest), NORMAL(normal))

    class HostSphere(F.SolidBody):
        type = 2  # IFL_BODY_HOST

    params = (0.1,) if level <= 5 else (0.1, 1.0, 0.01)
    inflow = (0.3, 0.2, 0.1, 0.05, 1.0, 600.0, 0.0, 0.5)
    sizes = [(64, 48), (80, 64), (48, 72)]

    def make(w, h):
        s = F.FluidSolver(level, w, h, params, [])
        assert F.lib.ifl_set_host_body_ops(s._h, ctypes.byref(ops)) == 0
        s.bodies = [HostSphere(cx, cy, r, r, 0.0, 0.0, 0.0, 0.0)]
        return s
    steps, dt = 3, 0.005
    batched = [make(w, h) for w, h in sizes]
    batch = F.Batch(batched)
    rows_b = []
    for _ in range(steps):
        for s in batched:
            s.add_inflow(*inflow)
        st = batch.update(dt)
        rows_b.append([(x.heat.iterations, x.pressure.iterations, x.pressure.status) for x in st])
    batch.close()
    fields = ol.state_fields(level)
    for k, (w, h) in enumerate(sizes):
        single = make(w, h)
        rows = []
        for _ in range(steps):
            single.add_inflow(*inflow)
            st = single.update(dt)
            rows.append((st.heat.iterations, st.pressure.iterations, st.pressure.status))
        assert rows == [row[k] for row in rows_b]
        bad = compare_fields(single, batched[k], fields, where=f"instance {k} {w}x{h}")
        assert not bad, "\n".join(bad[:6])
        single.close()
        batched[k].close()
    print(f"level {level}: {rows_b}", flush=True)


@pytest.mark.parametrize("level,limit", [(7, 7), (7, 8), (8, 7), (8, 8)])
def test_fused_batch_budget_exits(level, limit):
    """With a PCG budget the pressure solves end out of budget after an odd or
    an even number of iterations. At levels 6-8 the batch folds the vector
    update into the product (s and a second array by turns, 7:869 and 7:881
    in one pass), so an instance's last s lies in either array: its tail
    (the deferred s update, 7:881) must work on that array and leave the
    result in s, as a solver stepped alone does. Level 7 takes the batched
    tail, level 8 the tails on the solvers' own streams."""
    import ifl_fluids as F
    import ifl_scenes as S
    n, steps = 4, 2
    if level == 8:
        sc = S.config5_scene()
        dt = sc["dt"]

        def make(k):
            return F.FluidSolver(8, 128, 128, sc["params"], [F.SolidBox(*b[1:]) for b in sc["bodies"]],
                                 pcg_limit=limit)

        def inflows(k):
            return []  # solver 8 adds its inflow inside update()
    else:
        insts = [i for i in S.config4_instances(64) if max(i["w"], i["h"]) <= 160][:n]
        dt = insts[0]["dt"]

        def make(k):
            i = insts[k]
            bodies = [F.SolidBox(*(tuple(b[1:6]) + (0.0, 0.0, 0.0))) for b in i["bodies"]]
            return F.FluidSolver(7, i["w"], i["h"], (0.1, 1.0, 0.01), bodies, pcg_limit=limit)

        def inflows(k):
            # d = 1, t = 600, (u, v) = (0, 3): every pressure solve has work to do
            return [insts[k]["inflows"][0][:5] + (600.0, 0.0, 3.0)]
    seed = 0x5EED1234
    F.frand_set_state(seed)
    batched = [make(k) for k in range(n)]
    batch = F.Batch(batched)
    rows_b = []
    for _ in range(steps):
        entries = [(k,) + tuple(e) for k in range(n) for e in inflows(k)]
        if entries:
            batch.add_inflow(entries)
        st = batch.update(dt)
        rows_b.append([(x.heat.status, x.heat.iterations, x.pressure.status, x.pressure.iterations)
                       for x in st])
    batch.close()
    F.frand_set_state(seed)
    singles = [make(k) for k in range(n)]
    rows_s = []
    for _ in range(steps):
        row = []
        for k, s in enumerate(singles):
            for e in inflows(k):
                s.add_inflow(*e)
            x = s.update(dt)
            row.append((x.heat.status, x.heat.iterations, x.pressure.status, x.pressure.iterations))
        rows_s.append(row)
    print(f"level {level} limit {limit}: {rows_b}", flush=True)
    assert rows_s == rows_b
    # the budget exits this test is about (some drawn instances' inflows
    # leave their pressure solves nothing to do)
    budget = [r[2] == F.SOLVE_BUDGET for row in rows_b for r in row]
    assert sum(budget) >= len(budget) // 2
    fields = ol.state_fields(level)
    for k in range(n):
        bad = compare_fields(singles[k], batched[k], fields, where=f"level {level} instance {k}")
        assert not bad, "\n".join(bad[:6])
        singles[k].close()
        batched[k].close()
