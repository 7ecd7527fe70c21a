#!/usr/bin/env python3
"""Benchmark: FluidSolver::update() steps/s of solver 7 (variable density) at
2048x2048 on one MI355X, plus the PCG's achieved HBM bandwidth.

Workload: the seeded solver-7 scene (ifl_scenes.config3_scene): the
reference's box, two inflows with seeded temperatures, rho_air 0.1,
rho_soot 1.0, diffusion 0.01, dt 0.005. Inputs live in HBM before timing.
`python bench.py --steps K --warmup W` runs W untimed then K timed steps and
prints one JSON line. The CPU baseline times the reference's own solver 7
(oracle/_ref) on the host over a bounded sample (see cpu_baseline.sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "incremental-fluids_amd"))

PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per cell per PCG iteration (matvec 40, axpys + norm 48,
# forward sweep 40, backward sweep 48, s update 24) plus the 5 one-byte cell
# flag reads of the FLUID-masked solvers 6-8.
PCG_BYTES_PER_CELL_ITER = {3: 200.0, 4: 200.0, 5: 200.0, 6: 205.0, 7: 205.0, 8: 205.0}
# Algorithmic bytes per cell of one launch of the per-iteration kernels
# (solver 7): each array the kernel must read or write, once.
KERNEL_BYTES_PER_CELL = {
    # r, cx, cy, precon, q read; r, y written (the 64-lane FLUID masks add 1/8 B)
    "sweep_fwd": 56.0,
    # y, cx, cy, precon, r read; z written
    "sweep_bwd": 48.0,
    # adiag, aplusx, aplusy, s read; q written
    "matvec": 40.0,
}
# the same for solver 4 (the batched form): its backward sweep also reads
# the reference's _z at non-FLUID cells (A s, kept in q) for the unmasked dot
BATCH_BYTES_PER_CELL = dict(KERNEL_BYTES_PER_CELL, sweep_bwd=56.0)
# levels 6-8 in the FP64 form: the iterations after the first form the vector
# update inside the product (adiag, aplusx, aplusy, old s, z, p read; p, new
# s, q written; the FLUID flags): timed as the product
FUSED_MATVEC_BYTES_PER_CELL = 73.0


def kernel_bytes_per_cell(name, form):
    return FUSED_MATVEC_BYTES_PER_CELL if name == "matvec" and form == "fp64" else KERNEL_BYTES_PER_CELL[name]
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
# The order-preserving floor of a MIC(0) sweep: cell (x, y) needs (x-1, y) and
# (x, y-1), so any schedule that keeps each cell's operations takes W + H - 1
# dependent steps, each at least the recurrence's dependency chain (44 clocks,
# measured in a one-wave loop of the compute step) at the 2.4 GHz engine clock.
SWEEP_CHAIN_CLOCKS = 44
ENGINE_GHZ = 2.4


def sweep_floor_us(w, h):
    return (w + h - 1) * SWEEP_CHAIN_CLOCKS / (ENGINE_GHZ * 1e3)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=None, help="grid side (default 2048; 1024 for --workload flip)")
    ap.add_argument("--workload", choices=["solver7", "flip", "batch"], default="solver7",
                    help="solver7: the headline 2048^2 variable-density step; flip: solver 8 at 1024^2; "
                         "batch: 256 independent solver-4 instances of mixed sizes")
    ap.add_argument("--threads", type=int, default=16, help="batch: host threads stepping the instances")
    ap.add_argument("--instances", type=int, default=256, help="batch: number of instances")
    ap.add_argument("--batch-level", type=int, choices=[4, 8], default=4,
                    help="batch: solver 4 with the drawn sizes (config 4), or solver 8 (FLIP) at 256^2 (config 5)")
    ap.add_argument("--batch-mode", choices=["fused", "streams"], default="fused",
                    help="batch: one ifl_batch (PCG iterations fused over the instances) or one stream per instance")
    ap.add_argument("--flip-alpha", type=float, default=None,
                    help="flip: the FLIP blend (the reference's 0.001; 0.05 for the '0.95 FLIP' variant)")
    ap.add_argument("--g2p", choices=["bilinear", "cubic"], default="bilinear",
                    help="flip: gridToParticles interpolation (cubic: the Catmull-Rom variant, not the reference)")
    ap.add_argument("--form", choices=["fp64", "mixed", "refine"], default=None,
                    help="time only this form (default: FP64 as the value, the mixed form beside it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="solver7: skip the batched form's line")
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for rocprofv3 byte counters) and exit")
    return ap.parse_args()


def calibrate_copy():
    import torch
    n = 1 << 27  # doubles: 1 GiB
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n * 8, "bytes_written": n * 8, "launches": 3}))


class _OneCore:
    """Pins the calling thread to one host core (the first it may run on,
    as `taskset -c <core>` would) while a CPU baseline runs, and describes
    the host: the CPU model (/proc/cpuinfo, what lscpu reports), its logical
    CPU count and the core used."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.core = min(self.saved)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.saved)
        return False

    def host(self):
        model, mhz = "unknown", None
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name") and model == "unknown":
                        model = line.split(":", 1)[1].strip()
                    elif line.startswith("cpu MHz") and mhz is None:
                        mhz = float(line.split(":", 1)[1])
        except (OSError, ValueError):
            pass
        try:  # the pinned core's current clock, where cpufreq reports it
            with open(f"/sys/devices/system/cpu/cpu{self.core}/cpufreq/scaling_cur_freq") as f:
                mhz = int(f.read()) / 1000.0
        except (OSError, ValueError):
            pass
        return {"cpu_model": model, "cpu_mhz": mhz, "nproc": os.cpu_count(), "pinned_core": self.core}


def cpu_baseline(scene, n, mean_iters):
    with _OneCore() as pin:
        out = _cpu_baseline(scene, n, mean_iters)
        out["host"] = pin.host()
    return out


def _cpu_baseline(scene, n, mean_iters):
    """Times the reference's solver 7 (compiled from its own sources, or the
    C restatement when that build is absent) on one host core: one update()
    sequence with both PCG budgets capped, plus a run of PCG iterations, and
    extrapolates to the GPU run's mean iterations per step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    kind = "reference" if ol.ref_available(7) else "port"
    S = ol.RefSolver if kind == "reference" else ol.OracleSolver
    s = S(7, n, n, scene["params"], scene["bodies"])
    for inf in scene["inflows"]:
        s.add_inflow(*inf)
    dt = scene["dt"]
    cap = 2
    seq = [("FILL_SOLID", 0), ("FILL_SOLID", 1), ("FILL_SOLID", 2), ("FILL_SOLID", 3), ("COPY_T_TO_R", 0),
           ("BUILD_HEAT", 0), ("BUILD_PRECON", 0), ("PROJECT", cap), ("COPY_P_TO_T", 0), ("EXTRAPOLATE", 1),
           ("ADD_BUOYANCY", 0), ("SET_BOUNDARY", 0), ("BUILD_RHS", 0), ("COMPUTE_DENSITIES", 0),
           ("BUILD_PRESSURE", 0), ("BUILD_PRECON", 0), ("PROJECT", cap), ("APPLY_PRESSURE", 0),
           ("EXTRAPOLATE", 0), ("EXTRAPOLATE", 2), ("EXTRAPOLATE", 3), ("SET_BOUNDARY", 0),
           ("ADVECT", 0), ("ADVECT", 1), ("ADVECT", 2), ("ADVECT", 3)]
    t0 = time.perf_counter()
    for name, ia in seq:
        s.stage(name, ia, dt)
    t_seq = time.perf_counter() - t0
    # PCG iterations: rebuild the pressure system and run a capped solve
    s.stage("BUILD_RHS")
    s.stage("BUILD_PRESSURE", 0, dt)
    s.stage("BUILD_PRECON")
    t0 = time.perf_counter()
    s.stage("PROJECT", 0)
    t_p0 = time.perf_counter() - t0
    s.stage("BUILD_RHS")
    t0 = time.perf_counter()
    s.stage("PROJECT", cpu_iters_global)
    t_pk = time.perf_counter() - t0
    t_iter = max(t_pk - t_p0, 0.0) / cpu_iters_global
    t_fixed = t_seq - 2 * cap * t_iter
    step_s = t_fixed + mean_iters * t_iter
    if hasattr(s, "close"):
        s.close()
    return dict(value=1.0 / step_s, unit="steps/s", cores=1, kind=kind,
                sample=(f"solver 7 at {n}x{n} on one host core: one update() sequence with both PCG budgets "
                        f"capped at {cap} ({t_seq:.2f} s) and {cpu_iters_global} PCG iterations "
                        f"({t_iter * 1e3:.1f} ms each); step time extrapolated to the GPU run's mean "
                        f"{mean_iters:.0f} PCG iterations/step = {step_s:.1f} s"),
                seconds_per_step=step_s, ms_per_pcg_iteration=t_iter * 1e3)


cpu_iters_global = 12


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE runs recorded in profiles/traffic.json (each scaled by the same
    counter's reading for a device copy of known size), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def run_form(form, n, steps, warmup):
    """Creates the solver in `form`, runs `warmup` untimed steps, then times
    `steps` whole update() calls (inputs resident in HBM) with per-kernel
    HIP-event profiling on."""
    import ifl_fluids as F
    import ifl_scenes as S
    scene = S.config3_scene()
    prec = {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED, "refine": F.PREC_REFINE}[form]
    objs = [F.SolidBox(*b[1:]) if b[0] == S.BOX else F.SolidSphere(b[1], b[2], b[3], *b[5:]) for b in scene["bodies"]]
    solver = F.FluidSolver(7, n, n, scene["params"], objs, precision=prec)
    S.drive(solver, scene, warmup)
    solver.synchronize()
    solver.profile(True)
    solver.profile_reset()
    solver.synchronize()
    iters = []
    t0 = time.perf_counter()
    # the scene's body motion cadence continues from the warmup
    for k in range(steps):
        for inf in scene["inflows"]:
            solver.add_inflow(*inf)
        st = solver.update(scene["dt"])
        iters.append((st.heat.iterations, st.pressure.iterations, st.heat.status, st.pressure.status))
        if (warmup + k + 1) % scene["substeps"] == 0:
            solver.bodies_update(scene["dt"])
    solver.synchronize()
    wall = time.perf_counter() - t0
    prof = solver.profile_read(F.KERNEL_PCG)
    kern = {}
    for name, kid in (("sweep_fwd", F.KERNEL_SWEEP_FWD), ("sweep_bwd", F.KERNEL_SWEEP_BWD),
                      ("matvec", F.KERNEL_MATVEC)):
        p = solver.profile_read(kid)
        if p["launches"] == 0:
            continue
        avg_ms = p["ms"] / p["launches"]
        bytes_launch = kernel_bytes_per_cell(name, form) * n * n
        kern[name] = {"launches": p["launches"], "total_ms": p["ms"], "avg_us": 1e3 * avg_ms,
                      "bytes_per_launch": bytes_launch, "achieved_GBps": bytes_launch / (avg_ms * 1e-3) / 1e9}
    solver.close()
    pcg_iters = prof["iterations"]
    bpc = PCG_BYTES_PER_CELL_ITER[7]
    return {
        "scene": scene, "wall": wall, "steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
        "mean_iters": pcg_iters / max(1, steps), "solves": [list(x) for x in iters], "kernels": kern,
        "pcg": {"launches": prof["launches"], "ms_total": prof["ms"], "iterations": pcg_iters,
                "ms_per_iteration": prof["ms"] / max(1, pcg_iters), "bytes_per_cell_iteration": bpc,
                "achieved_GBps": bpc * prof["cell_iterations"] / (prof["ms"] * 1e-3) / 1e9 if prof["ms"] > 0 else 0.0},
    }


def mixed_frame_errors(n, frames=2, warm=2, form="mixed"):
    """A reduced-precision form's error per frame against the FP64 form,
    teacher-forced: from the FP64 form's state, one update() in each form, max relative
    difference per field (d, t, u, v; p less its mean over FLUID cells,
    since the pressure of these closed-wall systems is fixed only up to a
    constant, which the FP64 iteration lets drift and the mixed form
    deflates)."""
    import numpy as np
    import ifl_fluids as F
    import ifl_scenes as S
    scene = S.config3_scene()
    objs = [F.SolidBox(*b[1:]) for b in scene["bodies"]]
    ref = F.FluidSolver(7, n, n, scene["params"], objs)
    mix = F.FluidSolver(7, n, n, scene["params"], objs,
                        precision=F.PREC_MIXED if form == "mixed" else F.PREC_REFINE)
    S.drive(ref, scene, warm)
    fields = F.state_fields(7)
    names = {"d": F.Q_D * 16, "t": F.Q_T * 16, "u": F.Q_U * 16, "v": F.Q_V * 16, "p": F.F_P}
    out = []
    for _ in range(frames):
        for f in fields:
            v = ref.field(f)
            if v is not None:
                mix.set_field(f, v)
        for sv in (ref, mix):
            for inf in scene["inflows"]:
                sv.add_inflow(*inf)
        sts = [sv.update(scene["dt"]) for sv in (ref, mix)]
        fl = ref.field(F.Q_D * 16 + F.QF_CELL) == 0
        err = {}
        for k, f in names.items():
            a, b = ref.field(f), mix.field(f)
            if k == "p":
                a = np.where(fl, a - a[fl].mean(), 0.0)
                b = np.where(fl, b - b[fl].mean(), 0.0)
            err[k] = float(np.abs(a - b).max() / max(np.abs(a).max(), 1e-300))
        err["pressure_iterations"] = [st.pressure.iterations for st in sts]
        out.append(err)
    ref.close()
    mix.close()
    return out


# Algorithmic bytes of the particle kernels: per particle (P) and per grid
# node (N) of the grid the launch writes or samples, each array once.
#   P2G gather, d with t: posx, posy, prop d, prop t, sorted index (36 B/P);
#     group start, src d, src t, weight, two cell bytes (30 B/N)
#   P2G gather, u or v: 28 B/P; 21 B/N
#   G2P: posx, posy read, four properties read and written (80 B/P); the four
#     grids read (32 B/N)
#   particle advect: posx, posy read and written (32 B/P); u, v read (16 B/N)
#   P2G, the reduced-precision forms' LDS-binned scatter, counted as the
#     transfer: posx, posy and the properties read once (d with t 16 B/P,
#     u or v 12 B/P, float particles); src (and src t), weight and the cell
#     bytes written (26 B/N, 17 B/N). The binned records are the kernels'
#     own traffic.
# The reduced-precision forms keep the particle state in float: G2P 40 B/P,
# the particle advection 16 B/P.
def flip_kernel_bytes(n, p_before, p_after, binned=False):
    nd, nu, nv = n * n, (n + 1) * n, n * (n + 1)
    pb = 4 if binned else 8   # bytes per particle value
    if binned:
        p2g = (4 * pb * p_before + 26 * nd) + (3 * pb * p_before + 17 * nu) + (3 * pb * p_before + 17 * nv)
    else:
        p2g = (36 * p_before + 30 * nd) + (28 * p_before + 21 * nu) + (28 * p_before + 21 * nv)
    return {
        "p2g": p2g,
        "g2p": 10 * pb * p_after + 32 * nd,
        "particle_advect": 4 * pb * p_after + 16 * nd,
    }


def cpu_baseline_flip(scene, n):
    with _OneCore() as pin:
        out = _cpu_baseline_flip(scene, n)
        out["host"] = pin.host()
    return out


def _cpu_baseline_flip(scene, n):
    """One update() of the reference's solver 8 (or the C restatement) at the
    same size on one host core, from its own constructor."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    kind = "reference" if ol.ref_available(8) else "port"
    S = ol.RefSolver if kind == "reference" else ol.OracleSolver
    s = S(8, n, n, scene["params"], scene["bodies"], fresh=True)
    t0 = time.perf_counter()
    s.update(scene["dt"])
    step_s = time.perf_counter() - t0
    return dict(value=1.0 / step_s, unit="steps/s", cores=1, kind=kind,
                sample=f"one update() of solver 8 at {n}x{n} from a fresh solver on one host core ({step_s:.1f} s)",
                seconds_per_step=step_s)


def run_flip(args):
    """Solver 8 (FLIP) on the reference's own scene: steps/s, the PCG
    kernels and the particle kernels (HIP events on the solver's stream)."""
    import ifl_fluids as F
    import ifl_scenes as S
    n = args.size or 1024
    scene = S.config5_scene()
    F.frand_set_state(0xBA5EBA11)
    objs = [F.SolidBox(*b[1:]) for b in scene["bodies"]]
    form = args.form or "fp64"
    prec = {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED, "refine": F.PREC_REFINE}[form]
    variant = {}
    if args.flip_alpha is not None:
        variant["flip_alpha"] = args.flip_alpha
    if args.g2p == "cubic":
        variant["flip_g2p"] = 1
    solver = F.FluidSolver(8, n, n, scene["params"], objs, precision=prec, **variant)
    for _ in range(args.warmup):
        solver.update(scene["dt"])
    solver.synchronize()
    solver.profile(True)
    solver.profile_reset()
    byt = {"p2g": 0, "g2p": 0, "particle_advect": 0}
    solves, counts = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p0 = solver.particle_count()
        st = solver.update(scene["dt"])
        for k, v in flip_kernel_bytes(n, p0, st.particle_count, binned=form != "fp64").items():
            byt[k] += v
        solves.append([st.heat.iterations, st.pressure.iterations])
        counts.append(st.particle_count)
    solver.synchronize()
    wall = time.perf_counter() - t0
    kern = {}
    for name, kid in (("sweep_fwd", F.KERNEL_SWEEP_FWD), ("sweep_bwd", F.KERNEL_SWEEP_BWD),
                      ("matvec", F.KERNEL_MATVEC)):
        p = solver.profile_read(kid)
        if p["launches"]:
            b = kernel_bytes_per_cell(name, form) * n * n
            avg = p["ms"] / p["launches"]
            kern[name] = {"launches": p["launches"], "total_ms": p["ms"], "avg_us": 1e3 * avg,
                          "bytes_per_launch": b, "achieved_GBps": b / (avg * 1e-3) / 1e9}
    for name, kid in (("p2g", F.KERNEL_P2G), ("g2p", F.KERNEL_G2P),
                      ("particle_advect", F.KERNEL_PARTICLE_ADVECT)):
        p = solver.profile_read(kid)
        if p["launches"]:
            kern[name] = {"launches": p["launches"], "total_ms": p["ms"], "avg_us": 1e3 * p["ms"] / p["launches"],
                          "bytes_per_launch": byt[name] / p["launches"],
                          "achieved_GBps": byt[name] / (p["ms"] * 1e-3) / 1e9}
    prof = solver.profile_read(F.KERNEL_PCG)
    solver.close()
    dom = max(("sweep_fwd", "sweep_bwd", "matvec"), key=lambda k: kern.get(k, {"total_ms": -1})["total_ms"])
    k = kern[dom]
    out = {
        "metric": f"solver 8 (FLIP) steps/s at 1024^2, {'FP64 drop-in' if form == 'fp64' else form + ' form'}"
                  + ("" if not variant else f", variant {variant} (not the reference's)"),
        "value": args.steps / wall, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if form == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"solver8-flip-{n}x{n}-reference-scene", "grid": [n, n], "level": 8,
                   "dt": scene["dt"], "particles": counts, "solves": solves},
        "pcg": {"iterations": prof["iterations"], "ms_total": prof["ms"],
                "ms_per_iteration": prof["ms"] / max(1, prof["iterations"])},
        "kernels": kern,
        "roofline": {"bound": "hbm", "achieved": k["achieved_GBps"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": k["achieved_GBps"] / PEAK_HBM_GBS, "traffic": None,
                     "kernel": f"{dom}: {KERNEL_BYTES_PER_CELL[dom]:.0f} B/cell x {n}x{n} cells per launch"},
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline_flip(scene, n)
        except Exception as e:
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


def cpu_baseline_batch(insts, sample):
    with _OneCore() as pin:
        out = _cpu_baseline_batch(insts, sample)
        out["host"] = pin.host()
    return out


def _cpu_baseline_batch(insts, sample):
    """One update() of the first `sample` instances on one host core with the
    reference's solver 4 (or the C restatement): cell-updates/s over them."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    kind = "reference" if ol.ref_available(4) else "port"
    S = ol.RefSolver if kind == "reference" else ol.OracleSolver
    cells, secs = 0, 0.0
    for inst in insts[:sample]:
        s = S(4, inst["w"], inst["h"], inst["params"], inst["bodies"])
        for inf in inst["inflows"]:
            s.add_inflow(*inf)
        t0 = time.perf_counter()
        s.update(inst["dt"])
        secs += time.perf_counter() - t0
        cells += inst["w"] * inst["h"]
        if hasattr(s, "close"):
            s.close()
    return dict(value=cells / secs, unit="cell-updates/s", cores=1, kind=kind,
                sample=f"first update() of the first {sample} instances ({cells} cells) on one host core: {secs:.1f} s")


def flip_instances(n, size=256):
    """Config 5's batched part: n instances of the reference's FLIP scene
    (8:1469-1513) at size x size."""
    import ifl_scenes as S
    sc = S.config5_scene()
    return [dict(w=size, h=size, params=sc["params"], bodies=sc["bodies"], inflows=[], dt=sc["dt"])
            for _ in range(n)]


def measure_batch(n_inst, steps, warmup, mode="fused", threads=16, level=4):
    """Config 4: independent solver-4 instances (sides drawn from 64..512,
    one static box and one inflow each, ifl_scenes.config4_instances),
    stepped as one ifl_batch (mode "fused") or each on its own stream from
    `threads` host threads (mode "streams"; the library releases the GIL in
    every call). Returns the measurement: total cells x steps / s, and in the
    fused mode the fused PCG launches' times (HIP events on the batch's
    stream) with the cells that worked in them."""
    import threading
    import ifl_fluids as F
    import ifl_scenes as S
    insts = S.config4_instances(n_inst) if level == 4 else flip_instances(n_inst)
    if level == 8:
        F.frand_set_state(0xBA5EBA11)
    solvers = [F.FluidSolver(level, i["w"], i["h"], i["params"], [F.SolidBox(*b[1:]) for b in i["bodies"]])
               for i in insts]
    cells = sum(i["w"] * i["h"] for i in insts)
    nt = max(1, threads)
    errors = []

    def run(idx, steps, iters):
        try:
            for _ in range(steps):
                for k in idx:
                    for inf in insts[k]["inflows"]:
                        solvers[k].add_inflow(*inf)
                    st = solvers[k].update(insts[k]["dt"])
                    iters[k] += st.pressure.iterations
        except Exception as e:  # reported, the run is invalid
            errors.append(repr(e))

    def step_all(steps):
        iters = [0] * len(insts)
        # largest instances first, dealt round robin so the threads finish together
        order = sorted(range(len(insts)), key=lambda k: -insts[k]["w"] * insts[k]["h"])
        th = [threading.Thread(target=run, args=(order[t::nt], steps, iters)) for t in range(nt)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return time.perf_counter() - t0, iters

    batch = None
    if mode == "fused":
        batch = F.Batch(solvers)

        inflows = [(k,) + tuple(inf) for k in range(len(insts)) for inf in insts[k]["inflows"]]

        def step_all(steps):  # noqa: F811 - the fused form replaces the threaded one
            iters = [0] * len(insts)
            t0 = time.perf_counter()
            for _ in range(steps):
                batch.add_inflow(inflows)  # every instance's addInflow, one call
                st = batch.update(insts[0]["dt"])
                for k, x in enumerate(st):
                    iters[k] += x.pressure.iterations
            return time.perf_counter() - t0, iters

    step_all(warmup)
    wall, iters = step_all(steps)
    kern = {}
    if batch is not None:
        # a separate profiled step: events around every fused launch
        batch.profile(True)
        step_all(1)
        for name, kid in (("sweep_fwd", F.KERNEL_SWEEP_FWD), ("sweep_bwd", F.KERNEL_SWEEP_BWD),
                          ("matvec", F.KERNEL_MATVEC)):
            p = batch.profile_read(kid)
            if p["launches"] == 0 or p["ms"] <= 0:
                continue
            # levels 6-8 fold the vector update into the batch's product too
            bpc = (FUSED_MATVEC_BYTES_PER_CELL if name == "matvec" and level >= 6
                   else (BATCH_BYTES_PER_CELL if level == 4 else KERNEL_BYTES_PER_CELL)[name])
            byts = bpc * p["cell_iterations"]
            kern[name] = {"launches": p["launches"], "total_ms": p["ms"], "avg_us": 1e3 * p["ms"] / p["launches"],
                          "cell_iterations": p["cell_iterations"], "bytes": byts,
                          "achieved_GBps": byts / (p["ms"] * 1e-3) / 1e9}
        batch.profile(False)
        batch.close()
    for sv in solvers:
        sv.close()
    if errors:
        raise RuntimeError("; ".join(errors[:3]))
    out = {"value": cells * steps / wall, "unit": "cell-updates/s", "steps": steps, "warmup": warmup,
           "ms_per_step": 1e3 * wall / steps, "instances": len(insts), "cells": cells, "mode": mode,
           "threads": nt if mode == "streams" else 1, "streams": len(insts) if mode == "streams" else 1,
           "pressure_iterations_per_step": sum(iters) / steps, "insts": insts}
    if kern:
        out["kernels"] = kern
        dom = max(kern, key=lambda k: kern[k]["total_ms"])
        k = kern[dom]
        out["roofline"] = {"bound": "hbm", "achieved": k["achieved_GBps"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": k["achieved_GBps"] / PEAK_HBM_GBS, "traffic": None,
                           "kernel": f"fused {dom} over all instances: "
                                     f"{(BATCH_BYTES_PER_CELL if level == 4 else KERNEL_BYTES_PER_CELL)[dom]:.0f} B/cell x the "
                                     f"cells of the instances iterating in each launch, over the launches' summed "
                                     f"time (HIP events on the batch's stream, one profiled step)"}
    return out


def run_batch(args):
    """`--workload batch`: config 4 as the headline (cell-updates/s)."""
    try:
        m = measure_batch(args.instances, args.steps, args.warmup, args.batch_mode, args.threads, args.batch_level)
    except RuntimeError as e:
        print(json.dumps({"error": str(e)}))
        sys.exit(1)
    insts = m.pop("insts")
    out = {
        "metric": (f"cell-updates/s, {m['instances']} independent solver-4 instances of mixed sizes, FP64 drop-in"
                   if args.batch_level == 4 else
                   f"cell-updates/s, {m['instances']} independent solver-8 (FLIP) instances at 256^2, FP64 drop-in"),
        "value": m["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"solver4-batch-{m['instances']}-instances-64to512" if args.batch_level == 4
                                else f"solver8-flip-batch-{m['instances']}-instances-256x256"),
                   "instances": m["instances"],
                   "cells": m["cells"], "mode": m["mode"], "threads": m["threads"], "streams": m["streams"],
                   "pressure_iterations_per_step": m["pressure_iterations_per_step"]},
    }
    if "kernels" in m:
        out["kernels"] = m["kernels"]
        out["roofline"] = m["roofline"]
    if not args.no_cpu_baseline and args.batch_level == 4:
        try:
            out["cpu_baseline"] = cpu_baseline_batch(insts, 16)
        except Exception as e:
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


def main():
    global cpu_iters_global
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU strip decomposition is a design note only)"}))
        sys.exit(2)
    cpu_iters_global = args.cpu_iters
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "flip":
        run_flip(args)
        return
    if args.workload == "batch" and args.batch_mode == "streams":
        # one hardware queue per stepping thread (HIP's default is 4), set
        # before the runtime starts; at most 16 sweeps of <= 8 strips are then
        # in flight, well under one workgroup per CU
        want = min(16, max(4, args.threads))
        if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < want:
            os.environ["GPU_MAX_HW_QUEUES"] = str(want)
    if args.workload == "batch":
        run_batch(args)
        return
    n = args.size or 2048
    # the headline value is the form asked for (FP64 by default); without
    # --form the mixed form is run too and reported beside it
    main_form = args.form or "fp64"
    forms = [main_form] if args.form else ["fp64", "mixed", "refine"]
    # the forms beside the headline run at most 4 steps after 1 (bounded run time)
    res = {f: run_form(f, n, args.steps, args.warmup) if f == main_form
           else run_form(f, n, min(args.steps, 4), min(args.warmup, 1)) for f in forms}
    r = res[main_form]
    kern = r["kernels"]
    dom = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None
    out = {
        "metric": "solver steps/s at 2048^2 (solver 7; FP64 drop-in, mixed form beside it); PCG HBM GB/s",
        "value": r["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"solver7-variable-density-{n}x{n}-box-2inflows-seeded",
                   "grid": [n, n], "level": 7, "form": main_form, "dt": r["scene"]["dt"],
                   "pcg_iterations_per_step": r["mean_iters"], "solves": r["solves"]},
        "pcg": r["pcg"],
        "kernels": kern,
    }
    if dom:
        k = kern[dom]
        out["roofline"] = {"bound": "hbm", "achieved": k["achieved_GBps"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": k["achieved_GBps"] / PEAK_HBM_GBS,
                           "traffic": measured_traffic(dom) if main_form == "fp64" else None,
                           "kernel": f"{dom}: {KERNEL_BYTES_PER_CELL[dom]:.0f} B/cell x {n}x{n} cells per launch, "
                                     f"average over {k['launches']} launches (HIP events on the solver's stream)"}
        if dom.startswith("sweep"):
            fl = sweep_floor_us(n, n)
            # the same kernel against its dependency floor: the fraction of the
            # fastest order-preserving sweep's rate (HBM peak not reachable)
            out["roofline"]["latency_floor"] = {
                "us": fl, "avg_us": k["avg_us"], "frac": fl / k["avg_us"],
                "achieved_at_floor_GBps": k["bytes_per_launch"] / fl / 1e3,
                "basis": f"{2 * n - 1} dependent steps x {SWEEP_CHAIN_CLOCKS} clocks at {ENGINE_GHZ} GHz"}
    if len(forms) > 1:
        out["forms"] = {f: {"steps_per_s": res[f]["steps_per_s"], "ms_per_step": res[f]["ms_per_step"],
                            "steps": len(res[f]["solves"]),
                            "pcg_iterations_per_step": res[f]["mean_iters"],
                            "ms_per_pcg_iteration": res[f]["pcg"]["ms_per_iteration"],
                            "sweep_us": {k: v["avg_us"] for k, v in res[f]["kernels"].items()}} for f in forms}
        for f in forms[1:]:
            out["forms"][f"{f}_over_fp64"] = res[f]["steps_per_s"] / res["fp64"]["steps_per_s"]
            try:
                out["forms"][f"{f}_frame_error"] = mixed_frame_errors(n, form=f)
            except Exception as e:  # reported beside the measurement, never in place of it
                out["forms"][f"{f}_frame_error"] = {"error": str(e)}
    if not args.no_batch:
        # config 4 beside the headline: the batched form (one fused launch per
        # PCG phase for all instances), with its own roofline
        try:
            m = measure_batch(256, 2, 1)
            m.pop("insts")
            out["batch"] = m
        except Exception as e:  # reported beside the measurement, never in place of it
            out["batch"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(r["scene"], n, r["mean_iters"])
        except Exception as e:  # the baseline is reported, never the measured value
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
