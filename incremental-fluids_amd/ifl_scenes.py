"""Seeded synthetic workloads for the solver step.

All random draws come from SplitMix64 (Steele, Lea and Flood, "Fast
splittable pseudorandom number generators", OOPSLA 2014), seeded with
0x5EED0000 + workload number and consumed in the order documented per
workload, so the CPU checker and the GPU product see identical doubles.
Scene constants not drawn come from the reference programs' main() functions
(N-<solver>/Fluid.cpp).
"""
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def choice(self, seq):
        return seq[self.next_u64() % len(seq)]


BOX, SPHERE = 0, 1


def smoke_scene():
    """Solver 7's own scene (7:1084-1129) at test size."""
    return dict(level=7, params=(0.1, 1.0, 0.01), dt=0.005,
                bodies=[(BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0.0, 0.0, 0.0)],
                inflows=[(0.45, 0.2, 0.1, 0.05, 1.0, 294.0, 0.0, 0.0)], substeps=4)


def config5_scene():
    """Solver 8 (FLIP), the reference's own scene (8:1469-1513): the box of
    8:1484, rho_air 0.1, rho_soot 0.25, diffusion 0.01, dt 0.0025; the inflow
    is applied inside update() (8:1368). Particles start at 4 per cell from
    frand() (8:735-751)."""
    return dict(level=8, params=(0.1, 0.25, 0.01), dt=0.0025,
                bodies=[(BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0.0, 0.0, 0.0)], inflows=[], substeps=4)


def config3_scene(seed=0x5EED0003):
    """Solver 7 (variable density): the reference box (7:1099), rho_air 0.1,
    rho_soot 1.0, diffusion 0.01, dt 0.005, and two inflows (0.45, 0.2) and
    (0.25, 0.2), 0.1 x 0.05, d = 1, u = v = 0, with temperatures T1, T2 drawn
    in that order from U[300, 1200]."""
    rng = SplitMix64(seed)
    t1 = rng.uniform(300.0, 1200.0)
    t2 = rng.uniform(300.0, 1200.0)
    return dict(level=7, params=(0.1, 1.0, 0.01), dt=0.005,
                bodies=[(BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0.0, 0.0, 0.0)],
                inflows=[(0.45, 0.2, 0.1, 0.05, 1.0, t1, 0.0, 0.0),
                         (0.25, 0.2, 0.1, 0.05, 1.0, t2, 0.0, 0.0)], substeps=4)


def config2_scene(seed=0x5EED0002, static=False):
    """Solver 6 (heat): two SolidBox and one SolidSphere; per body, in order:
    pos x, pos y ~ U[0.2, 0.8]; sx, sy (sphere: s) ~ U[0.05, 0.15];
    theta ~ U[0, pi); omega ~ U[-pi, pi]; vx, vy ~ U[-0.3, 0.3]. static=True
    zeroes the velocities (the reference PCG does not converge with moving
    bodies). Inflow (0.35, 0.9, 0.1, 0.05), d = 1, T = 1000 K (6:1086)."""
    rng = SplitMix64(seed)
    bodies = []
    for kind in (BOX, BOX, SPHERE):
        x = rng.uniform(0.2, 0.8)
        y = rng.uniform(0.2, 0.8)
        sx = rng.uniform(0.05, 0.15)
        sy = sx if kind == SPHERE else rng.uniform(0.05, 0.15)
        th = rng.uniform(0.0, math.pi)
        om = rng.uniform(-math.pi, math.pi)
        vx = rng.uniform(-0.3, 0.3)
        vy = rng.uniform(-0.3, 0.3)
        if static:
            om = vx = vy = 0.0
        bodies.append((kind, x, y, sx, sy, th, vx, vy, om))
    return dict(level=6, params=(0.1, 0.1, 0.01), dt=0.005, bodies=bodies,
                inflows=[(0.35, 0.9, 0.1, 0.05, 1.0, 1000.0, 0.0, 0.0)], substeps=1)


def config1_scene(own_inflow=False):
    """Solver 3 (3:468-499): density 0.1, dt 0.005; inflow (0.45, 0.2) of
    0.1 x 0.01 (d = 1, u = 0, v = 3), or the program's own 0.15 x 0.03."""
    wh = (0.15, 0.03) if own_inflow else (0.1, 0.01)
    return dict(level=3, params=(0.1,), dt=0.005, bodies=[],
                inflows=[(0.45, 0.2, wh[0], wh[1], 1.0, 0.0, 0.0, 3.0)], substeps=4)


def config4_instances(n=256, seed=0x5EED0004):
    """Batched solver 4: per instance, in order: w, h from the side list;
    box centre ~ U[0.2, 0.8]^2 (scaled by w/min(w,h), h/min(w,h)); sx, sy ~
    U[0.05, 0.15]; theta ~ U[0, pi); inflow origin ~ U[0.2, 0.8]^2 with size
    0.15 x 0.03, d = 1, u = 0, v = 3. Bodies are static."""
    sides = (64, 96, 128, 192, 256, 384, 512)
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        w = rng.choice(sides)
        h = rng.choice(sides)
        m = min(w, h)
        cx = rng.uniform(0.2, 0.8) * w / m
        cy = rng.uniform(0.2, 0.8) * h / m
        sx = rng.uniform(0.05, 0.15)
        sy = rng.uniform(0.05, 0.15)
        th = rng.uniform(0.0, math.pi)
        ix = rng.uniform(0.2, 0.8)
        iy = rng.uniform(0.2, 0.8)
        out.append(dict(level=4, w=w, h=h, params=(0.1,), dt=0.005,
                        bodies=[(BOX, cx, cy, sx, sy, th, 0.0, 0.0, 0.0)],
                        inflows=[(ix, iy, 0.15, 0.03, 1.0, 0.0, 0.0, 3.0)], substeps=4))
    return out


def drive(solver, scene, steps, on_step=None):
    """Runs `steps` updates like the reference main(): inflows, update, and
    body motion once per frame of `substeps` updates (7:1110-1126)."""
    for k in range(steps):
        for inf in scene["inflows"]:
            solver.add_inflow(*inf)
        st = solver.update(scene["dt"])
        if on_step:
            on_step(k, st)
        if (k + 1) % scene["substeps"] == 0:
            solver.bodies_update(scene["dt"])
