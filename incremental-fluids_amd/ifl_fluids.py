"""Python front-end of libifl.so (the C ABI in include/ifl.h).

Mirrors the reference's FluidSolver / SolidBox / SolidSphere surface
(N-<solver>/Fluid.cpp): construct with the grid size and densities, add
bodies, call add_inflow() and update(dt). Solid bodies are plain Python
objects that the caller may move between steps, exactly like the reference's
borrowed `const vector<const SolidBody*>&` (7:622); their rotation is
evaluated with the host libm (math.cos / math.sin) so the device sees the
values the reference computes.

The native library is required: if libifl.so cannot be loaded, importing
this module raises. There is no CPU fallback.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libifl.so")

Q_D, Q_T, Q_U, Q_V = 0, 1, 2, 3
QF_SRC, QF_DST, QF_PHI, QF_VOLUME, QF_NORMALX, QF_NORMALY, QF_CELL, QF_BODY, QF_MASK = range(9)
F_R, F_P, F_Z, F_S, F_PRECON, F_ADIAG, F_APLUSX, F_APLUSY, F_UDENSITY, F_VDENSITY = range(64, 74)
F_POSX, F_POSY, F_PROP_D, F_PROP_T, F_PROP_U, F_PROP_V, F_WEIGHT, F_COUNTS = range(80, 88)
BODY_BOX, BODY_SPHERE = 0, 1
PREC_FP64, PREC_MIXED, PREC_REFINE = 0, 1, 2
DOT_TREE, DOT_SEQUENTIAL = 0, 1
SOLVE_CONVERGED, SOLVE_INITIAL_SMALL, SOLVE_BUDGET, SOLVE_NOT_RUN = 0, 1, 2, 3
KERNEL_PCG, KERNEL_PRECON, KERNEL_ADVECT, KERNEL_SWEEP_FWD, KERNEL_SWEEP_BWD, KERNEL_MATVEC = 0, 1, 2, 3, 4, 5
KERNEL_P2G, KERNEL_G2P, KERNEL_PARTICLE_ADVECT = 6, 7, 8

STAGE = dict(
    FILL_SOLID=1, EXTRAPOLATE=2, ADVECT=3, FLIP=4, BUILD_RHS=5, BUILD_PRESSURE=6,
    BUILD_HEAT=7, BUILD_PRECON=8, APPLY_PRECON=9, MATVEC=10, DOT=11, INF_NORM=12,
    PROJECT=13, APPLY_PRESSURE=14, ADD_BUOYANCY=15, SET_BOUNDARY=16,
    COMPUTE_DENSITIES=17, COPY_T_TO_R=18, COPY_P_TO_T=19, FROM_PARTICLES=32, G2P=33,
    P2G=34, PARTICLE_ADVECT=35, COPY_OLD=36, DIFF=37, UNDIFF=38, COUNT_PARTICLES=39,
    PRUNE=40, SEED=41)


class Config(ctypes.Structure):
    _fields_ = [
        ("level", ctypes.c_int32), ("w", ctypes.c_int32), ("h", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("dot_order", ctypes.c_int32), ("pcg_limit", ctypes.c_int32),
        ("verbose", ctypes.c_int32), ("device", ctypes.c_int32),
        ("density", ctypes.c_double), ("density_soot", ctypes.c_double), ("diffusion", ctypes.c_double),
        ("tau", ctypes.c_double), ("sigma", ctypes.c_double), ("tol", ctypes.c_double),
        ("t_amb", ctypes.c_double), ("gravity", ctypes.c_double), ("flip_alpha", ctypes.c_double),
        ("flip_g2p", ctypes.c_int32), ("reserved", ctypes.c_int32 * 9),
    ]


class BodyC(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("reserved", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in ("pos_x", "pos_y", "scale_x", "scale_y", "theta", "vel_x", "vel_y",
                                       "vel_theta", "cos_neg", "sin_neg", "cos_pos", "sin_pos")]


class SolveStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32), ("max_error", ctypes.c_double)]


class StepStats(ctypes.Structure):
    _fields_ = [("heat", SolveStats), ("pressure", SolveStats), ("particle_count", ctypes.c_int32),
                ("border_solid", ctypes.c_int32), ("nonfinite", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class IflError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libifl.so not built at {LIB_PATH}; run `make -C incremental-fluids_amd`")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.ifl_version.restype = ctypes.c_char_p
    lib.ifl_last_error.restype = ctypes.c_char_p
    lib.ifl_last_error.argtypes = [vp]
    lib.ifl_config_default.argtypes = [ctypes.POINTER(Config), ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ifl_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
    lib.ifl_destroy.argtypes = [vp]
    lib.ifl_set_bodies.argtypes = [vp, ctypes.POINTER(BodyC), ctypes.c_int]
    lib.ifl_add_inflow.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_double)]
    lib.ifl_update.argtypes = [vp, ctypes.c_double, ctypes.POINTER(StepStats)]
    lib.ifl_field_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]
    lib.ifl_upload.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    lib.ifl_download.argtypes = [vp, ctypes.c_int, vp, ctypes.c_size_t]
    lib.ifl_stage.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(SolveStats),
                              ctypes.POINTER(ctypes.c_double)]
    lib.ifl_profile_enable.argtypes = [vp, ctypes.c_int]
    lib.ifl_profile_reset.argtypes = [vp]
    lib.ifl_profile_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    lib.ifl_synchronize.argtypes = [vp]
    lib.ifl_get_particle_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    lib.ifl_set_particle_count.argtypes = [vp, ctypes.c_int]
    lib.ifl_init_particles.argtypes = [vp]
    lib.ifl_max_timestep.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    lib.ifl_to_image.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    lib.ifl_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    lib.ifl_batch_update.argtypes = [vp, ctypes.c_double, ctypes.POINTER(StepStats)]
    lib.ifl_batch_destroy.argtypes = [vp]
    dp = ctypes.POINTER(ctypes.c_double)
    lib.ifl_batch_add_inflow.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp, dp]
    lib.ifl_batch_profile.argtypes = [vp, ctypes.c_int]
    lib.ifl_batch_profile_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.ifl_write_png.argtypes = [ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_int]
    lib.ifl_frand_state.restype = ctypes.c_uint32
    lib.ifl_frand_state.argtypes = []
    lib.ifl_frand_set_state.argtypes = [ctypes.c_uint32]
    return lib


lib = _load()

_DTYPE = {8: np.float64, 4: np.int32, 1: np.uint8}
# the particle state is float in the reduced-precision forms (4-byte elements)
_FLOAT_FIELDS = frozenset(range(80, 86))


def _dtype(fid, e):
    return np.float32 if e == 4 and fid in _FLOAT_FIELDS else _DTYPE[e]


class SolidBody:
    """SolidBody 7:109-167: position, scale, rotation and rigid velocity."""

    type = BODY_BOX

    def __init__(self, x, y, sx, sy, theta, vx, vy, vtheta):
        self.pos_x, self.pos_y, self.scale_x, self.scale_y = x, y, sx, sy
        self.theta, self.vel_x, self.vel_y, self.vel_theta = theta, vx, vy, vtheta

    def update(self, timestep):  # 7:162-166
        self.pos_x += self.vel_x * timestep
        self.pos_y += self.vel_y * timestep
        self.theta += self.vel_theta * timestep

    def to_c(self):
        t = self.theta
        return BodyC(self.type, 0, self.pos_x, self.pos_y, self.scale_x, self.scale_y, t, self.vel_x, self.vel_y,
                     self.vel_theta, math.cos(-t), math.sin(-t), math.cos(t), math.sin(t))


class SolidBox(SolidBody):
    """SolidBox(x, y, sx, sy, t, vx, vy, vt), 7:172-173."""
    type = BODY_BOX


class SolidSphere(SolidBody):
    """SolidSphere(x, y, s, t, vx, vy, vt), 7:223-224."""
    type = BODY_SPHERE

    def __init__(self, x, y, s, theta, vx, vy, vtheta):
        super().__init__(x, y, s, s, theta, vx, vy, vtheta)


def default_config(level, w, h, **kw):
    cfg = Config()
    lib.ifl_config_default(ctypes.byref(cfg), level, w, h)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


class FluidSolver:
    """The MI355X solver for reference solver `level` (3..8).

    params follow the reference constructors: level 3-5 (density,),
    level 6-8 (rho_air, rho_soot, diffusion). `bodies` is a list of
    SolidBody objects the caller keeps and may move between steps.
    """

    def __init__(self, level, w, h, params=(0.1,), bodies=(), **cfg_kw):
        cfg = default_config(level, w, h, **cfg_kw)
        cfg.density = params[0]
        if level >= 6 and len(params) >= 3:
            cfg.density_soot = params[1]
            cfg.diffusion = params[2]
        self.cfg = cfg
        self.level, self.w, self.h = level, w, h
        self.bodies = list(bodies)
        self._h = ctypes.c_void_p()
        self._check(lib.ifl_create(ctypes.byref(cfg), ctypes.byref(self._h)), None)
        self.last_stats = StepStats()
        self._push_bodies()
        if level == 8:
            # the rest of the FLIP constructor (8:1308-1314)
            self._check(lib.ifl_init_particles(self._h))

    def _check(self, rc, handle=True):
        if rc != 0:
            msg = lib.ifl_last_error(self._h if handle else None)
            raise IflError(f"ifl call failed rc={rc}: {msg.decode() if msg else ''}")

    def _push_bodies(self):
        # the bodies are re-read every step (they may have moved); an
        # unchanged list is not sent again
        cs = [b.to_c() for b in self.bodies]
        key = [bytes(c) for c in cs]
        if key == getattr(self, "_pushed", None):
            return
        arr = (BodyC * max(1, len(cs)))(*cs)
        self._check(lib.ifl_set_bodies(self._h, arr, len(cs)))
        self._pushed = key

    def close(self):
        if self._h:
            lib.ifl_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # FluidSolver::addInflow 7:1036-1041
    def add_inflow(self, x, y, w, h, d, t, u, v):
        vals = (ctypes.c_double * 4)(d, t, u, v)
        self._check(lib.ifl_add_inflow(self._h, x, y, w, h, vals))

    # FluidSolver::update 7:995-1034; the bodies are re-read every step
    def update(self, dt):
        self._push_bodies()
        st = StepStats()
        self._check(lib.ifl_update(self._h, dt, ctypes.byref(st)))
        self.last_stats = st
        return st

    def bodies_update(self, dt):
        for b in self.bodies:
            b.update(dt)

    def field_info(self, fid):
        n = ctypes.c_size_t()
        e = ctypes.c_int()
        rc = lib.ifl_field_info(self._h, fid, ctypes.byref(n), ctypes.byref(e))
        if rc != 0:
            return 0, 0
        return n.value, e.value

    def field(self, fid):
        n, e = self.field_info(fid)
        if n == 0:
            return None
        a = np.empty(n, dtype=_dtype(fid, e))
        self._check(lib.ifl_download(self._h, fid, a.ctypes.data, a.nbytes))
        return a

    def set_field(self, fid, values):
        n, e = self.field_info(fid)
        a = np.ascontiguousarray(np.asarray(values, dtype=_dtype(fid, e)).reshape(n))
        self._check(lib.ifl_upload(self._h, fid, a.ctypes.data, a.nbytes))

    def stage(self, stage, iarg=0, darg=0.0, stats=None):
        sid = STAGE[stage] if isinstance(stage, str) else stage
        res = ctypes.c_double(0.0)
        st = stats if stats is not None else SolveStats()
        self._check(lib.ifl_stage(self._h, sid, iarg, darg, ctypes.byref(st), ctypes.byref(res)))
        return res.value

    def profile(self, on=True):
        self._check(lib.ifl_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        self._check(lib.ifl_profile_reset(self._h))

    def profile_read(self, kernel=KERNEL_PCG):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        it = ctypes.c_int64()
        ci = ctypes.c_double()
        self._check(lib.ifl_profile_read(self._h, kernel, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(it),
                                         ctypes.byref(ci)))
        return dict(launches=n.value, ms=ms.value, iterations=it.value, cell_iterations=ci.value)

    def synchronize(self):
        self._check(lib.ifl_synchronize(self._h))

    def to_image(self, render_heat=False):
        """toImage (7:1047): RGBA bytes, shape (h, w, 4) or (h, 2w, 4) with heat."""
        heat = bool(render_heat) and self.level >= 6
        a = np.empty((self.h, self.w * (2 if heat else 1), 4), dtype=np.uint8)
        self._check(lib.ifl_to_image(self._h, a.ctypes.data, a.nbytes, 1 if heat else 0))
        return a

    def max_timestep(self):
        """FluidSolver::maxTimestep (1:310-328): solver 1's CFL bound."""
        dt = ctypes.c_double()
        self._check(lib.ifl_max_timestep(self._h, ctypes.byref(dt)))
        return dt.value

    def particle_count(self, value=None):
        """Level 8: the live particle count (_particleCount 8:701)."""
        if value is not None:
            self._check(lib.ifl_set_particle_count(self._h, value))
            return value
        n = ctypes.c_int()
        self._check(lib.ifl_get_particle_count(self._h, ctypes.byref(n)))
        return n.value


def state_fields(level):
    """The arrays that make up a solver's state between two update() calls
    (the quantities and their solid geometry, the PCG vectors and matrix, the
    face densities at levels 7-8, the particles at level 8): what a caller
    copies to continue a run in another solver."""
    qs = [Q_D, Q_U, Q_V] + ([Q_T] if level >= 6 else [])
    out = []
    for q in qs:
        out.append(q * 16 + QF_SRC)
        if level != 8:
            out.append(q * 16 + QF_DST)
        if level >= 4:
            out += [q * 16 + f for f in (QF_CELL, QF_BODY, QF_NORMALX, QF_NORMALY, QF_MASK)]
        if level >= 5:
            out += [q * 16 + QF_VOLUME, q * 16 + QF_PHI]
    out += [F_R, F_P, F_Z, F_S, F_PRECON, F_ADIAG, F_APLUSX, F_APLUSY]
    if level >= 7:
        out += [F_UDENSITY, F_VDENSITY]
    if level == 8:
        out += [F_POSX, F_POSY, F_PROP_D, F_PROP_T, F_PROP_U, F_PROP_V, F_COUNTS]
    return out


def frand_state():
    """frand()'s process-wide seed (8:38-46)."""
    return lib.ifl_frand_state()


def frand_set_state(state):
    lib.ifl_frand_set_state(state & 0xFFFFFFFF)


def reference_log_line(level, stats):
    """The stdout line the reference prints for one project() (7:860/874/885)."""
    if stats.status == SOLVE_CONVERGED:
        return "Exiting solver after %d iterations, maximum error is %f\n" % (stats.iterations, stats.max_error)
    if stats.status == SOLVE_BUDGET:
        return "Exceeded budget of %d iterations, maximum error was %f\n" % (stats.iterations, stats.max_error)
    if stats.status == SOLVE_INITIAL_SMALL and level >= 6:
        return "Initial guess sufficiently small\n"
    return ""


def write_png(path, rgba):
    """The reference's frame file (7:1122): an RGBA, 8-bit PNG."""
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = a.shape[0], a.shape[1]
    rc = lib.ifl_write_png(os.fsencode(path), a.ctypes.data, w, h)
    if rc != 0:
        raise IflError(f"ifl_write_png failed rc={rc}: {lib.ifl_last_error(None).decode()}")


class Batch:
    """Independent solvers of one level stepped together (ifl_batch_*): each
    PCG iteration of their solves is one launch per phase for all of them."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        self._h = ctypes.c_void_p()
        rc = lib.ifl_batch_create(arr, len(self.solvers), ctypes.byref(self._h))
        if rc != 0:
            raise IflError(f"ifl_batch_create failed rc={rc}: {lib.ifl_last_error(None).decode()}")

    def add_inflow(self, entries):
        """addInflow of many solvers in one call: entries are (index into
        the batch's solvers, x, y, w, h, d, t, u, v); a solver's entries
        apply in their order (ifl_batch_add_inflow)."""
        entries = list(entries)
        if not entries:
            return
        k = len(entries)
        inst = (ctypes.c_int * k)(*[int(e[0]) for e in entries])
        rect = (ctypes.c_double * (4 * k))(*[float(v) for e in entries for v in e[1:5]])
        vals = (ctypes.c_double * (4 * k))(*[float(v) for e in entries for v in e[5:9]])
        rc = lib.ifl_batch_add_inflow(self._h, k, inst, rect, vals)
        if rc != 0:
            raise IflError(f"ifl_batch_add_inflow failed rc={rc}: {lib.ifl_last_error(None).decode()}")

    def update(self, dt):
        for s in self.solvers:
            s._push_bodies()
        st = (StepStats * len(self.solvers))()
        rc = lib.ifl_batch_update(self._h, dt, st)
        if rc != 0:
            raise IflError(f"ifl_batch_update failed rc={rc}: {lib.ifl_last_error(None).decode()}")
        return list(st)

    def profile(self, on=True):
        """Times every fused PCG launch (and clears the totals)."""
        lib.ifl_batch_profile(self._h, 1 if on else 0)

    def profile_read(self, kernel):
        """Totals of one fused kernel (KERNEL_MATVEC, KERNEL_SWEEP_FWD,
        KERNEL_SWEEP_BWD): launches, ms and the cells of the instances that
        worked in them, summed over the launches."""
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        ci = ctypes.c_double()
        rc = lib.ifl_batch_profile_read(self._h, kernel, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(ci))
        if rc != 0:
            raise IflError(f"ifl_batch_profile_read failed rc={rc}")
        return {"launches": n.value, "ms": ms.value, "cell_iterations": ci.value}

    def close(self):
        if self._h:
            lib.ifl_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
