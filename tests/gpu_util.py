"""Helpers shared by the GPU parity tests: scene driving and state transfer
between the oracle and the product (test infrastructure)."""
import numpy as np

import oracle_lib as ol


def make_product(level, w, h, params, bodies, **kw):
    import ifl_fluids as F
    objs = []
    for b in bodies:
        t, x, y, sx, sy, th, vx, vy, vt = b
        if t == F.BODY_SPHERE:
            objs.append(F.SolidSphere(x, y, sx, th, vx, vy, vt))
        else:
            objs.append(F.SolidBox(x, y, sx, sy, th, vx, vy, vt))
    return F.FluidSolver(level, w, h, params, objs, **kw)


def compare_fields(a, b, fields, where=""):
    """Bit-for-bit comparison; returns a list of mismatch descriptions."""
    bad = []
    for f in fields:
        x, y = a.field(f), b.field(f)
        if x is None and y is None:
            continue
        if x is None or y is None:
            bad.append(f"{where} field {f}: present in one side only")
            continue
        if x.shape != y.shape:
            bad.append(f"{where} field {f}: shape {x.shape} vs {y.shape}")
            continue
        if x.tobytes() != y.tobytes():
            if x.dtype == np.float64:
                xv, yv = x.view(np.uint64), y.view(np.uint64)
                idx = np.nonzero(xv != yv)[0]
                i = idx[0]
                rel = np.nanmax(np.abs(x[idx] - y[idx]) / np.maximum(np.abs(x[idx]), 1e-300))
                bad.append(f"{where} field {f}: {len(idx)} differ, first {i}: {x[i]!r} vs {y[i]!r}, max rel {rel:.3g}")
            else:
                idx = np.nonzero(x != y)[0]
                bad.append(f"{where} field {f}: {len(idx)} differ, first {idx[0]}: {x[idx[0]]} vs {y[idx[0]]}")
    return bad


def copy_state(src, dst, fields):
    for f in fields:
        v = src.field(f)
        if v is not None:
            dst.set_field(f, v)
