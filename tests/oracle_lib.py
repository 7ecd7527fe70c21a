"""ctypes front-ends for the two CPU checkers (test infrastructure only).

``RefSolver`` drives the reference's own Fluid.cpp (solvers 3..8) compiled
from the reference tree by oracle/Makefile into oracle/_ref/libref<N>.so.
``OracleSolver`` drives the C restatement oracle/liboracle.so. Both expose the
same Python surface so a test can run one scene through either and compare
every field bit for bit.
"""
import ctypes
import os
import shutil
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

# ids shared with include/ifl.h
Q_D, Q_T, Q_U, Q_V = 0, 1, 2, 3
QF_SRC, QF_DST, QF_PHI, QF_VOLUME, QF_NORMALX, QF_NORMALY, QF_CELL, QF_BODY, QF_MASK = range(9)
F_R, F_P, F_Z, F_S, F_PRECON, F_ADIAG, F_APLUSX, F_APLUSY, F_UDENSITY, F_VDENSITY = range(64, 74)
F_POSX, F_POSY, F_PROP_D, F_PROP_T, F_PROP_U, F_PROP_V, F_WEIGHT, F_COUNTS = range(80, 88)
BODY_BOX, BODY_SPHERE = 0, 1

STAGE = dict(
    FILL_SOLID=1, EXTRAPOLATE=2, ADVECT=3, FLIP=4, BUILD_RHS=5, BUILD_PRESSURE=6,
    BUILD_HEAT=7, BUILD_PRECON=8, APPLY_PRECON=9, MATVEC=10, DOT=11, INF_NORM=12,
    PROJECT=13, APPLY_PRESSURE=14, ADD_BUOYANCY=15, SET_BOUNDARY=16,
    COMPUTE_DENSITIES=17, COPY_T_TO_R=18, COPY_P_TO_T=19, FROM_PARTICLES=32, G2P=33,
    P2G=34, PARTICLE_ADVECT=35, COPY_OLD=36, DIFF=37, UNDIFF=38, COUNT_PARTICLES=39,
    PRUNE=40, SEED=41)


def qfield(q, f):
    return q * 16 + f


class BodyPod(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("x", "y", "sx", "sy", "theta", "vx", "vy", "vtheta")]


def make_bodies(bodies):
    """bodies: list of dicts/tuples (type, x, y, sx, sy, theta, vx, vy, vtheta)."""
    arr = (BodyPod * max(1, len(bodies)))()
    for i, b in enumerate(bodies):
        arr[i] = BodyPod(*b)
    return arr, len(bodies)


def ref_available(level=7):
    return os.path.exists(os.path.join(REF_DIR, f"libref{level}.so"))


def _bind_common(lib, prefix):
    f = getattr(lib, prefix + "update")
    f.argtypes = [ctypes.c_void_p, ctypes.c_double] + ([ctypes.POINTER(ctypes.c_int)] if prefix == "orc_" else [])
    f = getattr(lib, prefix + "add_inflow")
    f.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 8
    f = getattr(lib, prefix + "field")
    f.restype = ctypes.c_void_p
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int)]
    f = getattr(lib, prefix + "stage")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    f = getattr(lib, prefix + "bodies_update")
    f.argtypes = [ctypes.c_void_p, ctypes.c_double]
    f = getattr(lib, prefix + "particle_count")
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]


_DT = {8: np.float64, 4: np.int32, 1: np.uint8}


class _Base:
    prefix = ""

    def _ptr(self, fid):
        n = ctypes.c_long()
        e = ctypes.c_int()
        p = getattr(self.lib, self.prefix + "field")(self.h, fid, ctypes.byref(n), ctypes.byref(e))
        if not p or n.value == 0:
            return None
        ctype = {8: ctypes.c_double, 4: ctypes.c_int32, 1: ctypes.c_uint8}[e.value]
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctype)), shape=(n.value,))

    def field(self, fid):
        a = self._ptr(fid)
        return None if a is None else a.copy()

    def view(self, fid):
        return self._ptr(fid)

    def set_field(self, fid, values):
        a = self._ptr(fid)
        a[...] = np.asarray(values, dtype=a.dtype).reshape(a.shape)

    def add_inflow(self, x, y, w, h, d, t, u, v):
        getattr(self.lib, self.prefix + "add_inflow")(self.h, x, y, w, h, d, t, u, v)

    def bodies_update(self, dt):
        getattr(self.lib, self.prefix + "bodies_update")(self.h, dt)

    def stage(self, stage, iarg=0, darg=0.0):
        res = ctypes.c_double(0.0)
        sid = STAGE[stage] if isinstance(stage, str) else stage
        rc = getattr(self.lib, self.prefix + "stage")(self.h, sid, iarg, darg, ctypes.byref(res))
        if rc != 0:
            raise RuntimeError(f"stage {stage} rc={rc}")
        return res.value

    def particle_count(self, value=None):
        f = getattr(self.lib, self.prefix + "particle_count")
        return f(self.h, 0 if value is None else 1, 0 if value is None else value)


class RefSolver(_Base):
    """The reference solver N compiled from its own sources.

    fresh=True loads a private copy of the library so that process-global
    state (level 8's frand() seed, the captured stdout) starts from scratch.
    """
    prefix = "ref_"

    def __init__(self, level, w, h, params, bodies=(), fresh=False):
        path = os.path.join(REF_DIR, f"libref{level}.so")
        if fresh:
            self._tmp = tempfile.mkdtemp(prefix="ifl_ref_")
            p2 = os.path.join(self._tmp, f"libref{level}_{id(self)}.so")
            shutil.copy(path, p2)
            path = p2
        self.lib = ctypes.CDLL(path)
        lib = self.lib
        lib.ref_create.restype = ctypes.c_void_p
        lib.ref_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(BodyPod), ctypes.c_int]
        lib.ref_log.restype = ctypes.c_char_p
        lib.ref_frand.restype = ctypes.c_double
        _bind_common(lib, "ref_")
        lib.ref_destroy.argtypes = [ctypes.c_void_p]
        self.level, self.w, self.hgt = level, w, h
        p = (ctypes.c_double * 3)(*(list(params) + [0.0] * (3 - len(params))))
        arr, n = make_bodies(list(bodies))
        lib.ref_log_clear()
        self.h = lib.ref_create(w, h, p, arr, n)

    def update(self, dt):
        self.lib.ref_update(self.h, dt)

    def log(self):
        return self.lib.ref_log().decode()

    def log_clear(self):
        self.lib.ref_log_clear()

    def close(self):
        if self.level != 8 and self.h:
            self.lib.ref_destroy(self.h)
            self.h = None


_orc_lib = None


def oracle_lib():
    global _orc_lib
    if _orc_lib is None:
        lib = ctypes.CDLL(os.path.join(ORACLE_DIR, "liboracle.so"))
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(BodyPod), ctypes.c_int]
        lib.orc_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_log.restype = ctypes.c_char_p
        lib.orc_log.argtypes = [ctypes.c_void_p]
        lib.orc_log_clear.argtypes = [ctypes.c_void_p]
        lib.orc_frand.restype = ctypes.c_double
        lib.orc_frand_seed.argtypes = [ctypes.c_uint32]
        lib.orc_frand_state.restype = ctypes.c_uint32
        lib.orc_occupancy.restype = ctypes.c_double
        lib.orc_occupancy.argtypes = [ctypes.c_double] * 4
        lib.orc_body_distance_pod.restype = ctypes.c_double
        lib.orc_body_distance_pod.argtypes = [ctypes.POINTER(BodyPod), ctypes.c_double, ctypes.c_double]
        lib.orc_cubic_pulse.restype = ctypes.c_double
        lib.orc_cubic_pulse.argtypes = [ctypes.c_double]
        lib.orc_cerp1.restype = ctypes.c_double
        lib.orc_cerp1.argtypes = [ctypes.c_double] * 5
        lib.orc_set_mic.argtypes = [ctypes.c_double, ctypes.c_double]
        _bind_common(lib, "orc_")
        _orc_lib = lib
    return _orc_lib


class OracleSolver(_Base):
    """The C restatement, same surface as RefSolver."""
    prefix = "orc_"

    def __init__(self, level, w, h, params, bodies=(), fresh=False):
        self.lib = oracle_lib()
        if fresh:
            self.lib.orc_frand_seed(0xBA5EBA11)
        self.level, self.w, self.hgt = level, w, h
        p = (ctypes.c_double * 3)(*(list(params) + [0.0] * (3 - len(params))))
        arr, n = make_bodies(list(bodies))
        self.h = self.lib.orc_create(level, w, h, p, arr, n)
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        self.last_iters = (None, None)

    def update(self, dt):
        it = (ctypes.c_int * 2)()
        self.lib.orc_update(self.h, dt, it)
        self.last_iters = (it[0], it[1])

    def log(self):
        return self.lib.orc_log(self.h).decode()

    def set_flip(self, alpha, cubic=False):
        """Level 8: the FLIP update's alpha and its gridToParticles
        interpolation (cubic: the Catmull-Rom variant, not the reference's)."""
        f = self.lib.orc_set_flip
        f.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        f(self.h, alpha, 1 if cubic else 0)

    def log_clear(self):
        self.lib.orc_log_clear(self.h)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None


# ---------------------------------------------------------------- scenes
# The reference programs' own main() scenes, scaled to test sizes.
import math

REF_SCENES = {
    # 3:468-499
    3: dict(params=(0.1,), dt=0.005, bodies=[],
            inflow=(0.45, 0.2, 0.15, 0.03, 1.0, 0.0, 0.0, 3.0), substeps=4),
    # 4:919-965
    4: dict(params=(0.1,), dt=0.005, bodies=[(BODY_BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0, 0, 0)],
            inflow=(0.45, 0.2, 0.15, 0.03, 1.0, 0.0, 0.0, 3.0), substeps=4),
    # 5:974-1016
    5: dict(params=(0.1,), dt=0.005, bodies=[(BODY_BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0, 0, 0)],
            inflow=(0.45, 0.2, 0.15, 0.03, 1.0, 0.0, 0.0, 3.0), substeps=4),
    # 6:1060-1103
    6: dict(params=(0.1, 0.1, 0.01), dt=0.005, bodies=[(BODY_BOX, 0.3, 0.6, 0.1, 0.5, -math.pi * 0.05, 0, 0, 0)],
            inflow=(0.35, 0.9, 0.1, 0.05, 1.0, 294.0 + 300.0, 0.0, 0.0), substeps=10),
    # 7:1084-1129
    7: dict(params=(0.1, 1.0, 0.01), dt=0.005, bodies=[(BODY_BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0, 0, 0)],
            inflow=(0.45, 0.2, 0.1, 0.05, 1.0, 294.0, 0.0, 0.0), substeps=4),
    # 8:1469-1513 (inflow is inside update, 8:1368)
    8: dict(params=(0.1, 0.25, 0.01), dt=0.0025, bodies=[(BODY_BOX, 0.5, 0.6, 0.7, 0.1, math.pi * 0.25, 0, 0, 0)],
            inflow=None, substeps=4),
}


def run_scene(solver, level, steps, scene=None, on_step=None):
    """Drives `solver` like the reference main(): inflow, update, and body
    update once per frame of `substeps` updates."""
    sc = dict(REF_SCENES[level])
    if scene:
        sc.update(scene)
    for k in range(steps):
        if sc["inflow"] is not None:
            solver.add_inflow(*sc["inflow"])
        solver.update(sc["dt"])
        if (k + 1) % sc["substeps"] == 0:
            solver.bodies_update(sc["dt"])
        if on_step:
            on_step(k, solver)


def state_fields(level):
    """Fields that make up the solver state between steps, per level."""
    qs = [Q_D, Q_U, Q_V] + ([Q_T] if level >= 6 else [])
    out = []
    for q in qs:
        out.append(qfield(q, QF_SRC))
        if level != 8:
            out.append(qfield(q, QF_DST))
        if level >= 4:
            out += [qfield(q, QF_CELL), qfield(q, QF_BODY), qfield(q, QF_NORMALX), qfield(q, QF_NORMALY),
                    qfield(q, QF_MASK)]
        if level >= 5:
            out += [qfield(q, QF_VOLUME), qfield(q, QF_PHI)]
    out += [F_R, F_P, F_Z, F_S, F_PRECON, F_ADIAG, F_APLUSX, F_APLUSY]
    if level >= 7:
        out += [F_UDENSITY, F_VDENSITY]
    if level == 8:
        out += [F_POSX, F_POSY, F_PROP_D, F_PROP_T, F_PROP_U, F_PROP_V, F_COUNTS]
    return out
