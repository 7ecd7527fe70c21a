"""Pins of the C restatement that do not depend on running the reference:
closed forms, textbook reductions and invariants. Each is chosen so that a
plausible slip (a dropped term, a wrong sign or index, a transposed operand)
fails it.
"""
import ctypes
import math
import struct

import numpy as np
import pytest

import oracle_lib as ol

lib = ol.oracle_lib()


# ---------------------------------------------------------------- occupancy
def _clip_area(phi_fn):
    """Area of {phi < 0} inside the unit square for a linear phi, by polygon
    clipping (Sutherland-Hodgman against the half-plane)."""
    poly = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    out = []
    for i in range(4):
        p, q = poly[i], poly[(i + 1) % 4]
        fp, fq = phi_fn(*p), phi_fn(*q)
        if fp < 0:
            out.append(p)
        if (fp < 0) != (fq < 0):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    a = 0.0
    for i in range(len(out)):
        x0, y0 = out[i]
        x1, y1 = out[(i + 1) % len(out)]
        a += x0 * y1 - x1 * y0
    return abs(a) / 2.0


def test_occupancy_exact_for_half_planes():
    """occupancy 7:60-102 of a linear SDF equals the clipped area. Corner
    convention from fillSolidFields 7:506-509: d11=(0,0), d12=(1,0),
    d21=(0,1), d22=(1,1)."""
    rng = np.random.default_rng(1)
    seen = set()
    for _ in range(4000):
        th = rng.uniform(0, 2 * math.pi)
        c = rng.uniform(-1.2, 1.2)
        nx, ny = math.cos(th), math.sin(th)
        phi = lambda x, y: nx * (x - 0.5) + ny * (y - 0.5) + c
        d11, d12, d21, d22 = phi(0, 0), phi(1, 0), phi(0, 1), phi(1, 1)
        if min(abs(d11), abs(d12), abs(d21), abs(d22)) < 1e-6:
            continue
        got = lib.orc_occupancy(d11, d12, d21, d22)
        want = _clip_area(phi)
        assert abs(got - want) < 1e-12, (d11, d12, d21, d22, got, want)
        seen.add(sum(1 for d in (d11, d12, d21, d22) if d < 0))
    assert seen == {0, 1, 2, 3, 4}


def test_occupancy_saddle_cases_follow_the_reference_formula():
    """Opposite corners inside (0x5 / 0xA, 7:93-96) are summed from the two
    triangles around the *outside* corners (the `in` argument is an outside
    corner there), so occupancy returns the outside area for a saddle: a
    reference quirk kept as is. Checked against the sum of those triangles."""
    tri = lambda o1, i, o2: 0.5 * i * i / ((o1 - i) * (o2 - i))
    rng = np.random.default_rng(5)
    for _ in range(200):
        a, b = -rng.uniform(0.05, 0.5, 2)
        c, d = rng.uniform(0.05, 0.5, 2)
        # 0x5: d11, d22 inside
        assert lib.orc_occupancy(a, c, d, b) == tri(a, c, b) + tri(b, d, a)
        # 0xA: d12, d21 inside
        assert lib.orc_occupancy(c, a, b, d) == tri(b, c, a) + tri(a, d, b)


# ---------------------------------------------------------------- bodies
def test_box_and_sphere_distance_closed_form():
    box = ol.BodyPod(ol.BODY_BOX, 0.5, 0.5, 0.4, 0.2, 0.0, 0, 0, 0)
    d = lambda b, x, y: lib.orc_body_distance_pod(ctypes.byref(b), x, y)
    assert d(box, 0.5, 0.5) == -0.1                      # deepest inside: max(dx, dy)
    assert abs(d(box, 0.9, 0.5) - 0.2) < 1e-15            # outside along x
    assert abs(d(box, 0.9, 0.8) - math.hypot(0.2, 0.2)) < 1e-15  # outside a corner
    assert abs(d(box, 0.5, 0.65) - 0.05) < 1e-15
    rot = ol.BodyPod(ol.BODY_BOX, 0.5, 0.5, 0.4, 0.2, math.pi / 2, 0, 0, 0)  # axes swap
    assert abs(d(rot, 0.5, 0.9) - 0.2) < 1e-15
    sph = ol.BodyPod(ol.BODY_SPHERE, 0.3, 0.4, 0.2, 0.2, 0.0, 0, 0, 0)
    for x, y in [(0.3, 0.4), (0.6, 0.8), (0.1, 0.2)]:
        assert abs(d(sph, x, y) - (math.hypot(x - 0.3, y - 0.4) - 0.1)) < 1e-15


def test_cubic_pulse_closed_form():
    for x in np.linspace(-1.5, 1.5, 301):
        a = min(abs(x), 1.0)
        assert abs(lib.orc_cubic_pulse(x) - (1 - 3 * a * a + 2 * a ** 3)) < 1e-15
    assert lib.orc_cubic_pulse(0.0) == 1.0 and lib.orc_cubic_pulse(1.0) == 0.0


def test_cerp_properties():
    """Catmull-Rom (7:281-295): reproduces b at x = 0, constants exactly,
    linear data to rounding, and never leaves [min, max] of the samples."""
    rng = np.random.default_rng(2)
    for _ in range(500):
        a, b, c, d = rng.normal(size=4)
        assert lib.orc_cerp1(a, b, c, d, 0.0) == b
        k = rng.normal()
        x = rng.uniform(0, 1)
        assert lib.orc_cerp1(k, k, k, k, x) == k
        h = rng.normal()
        v = lib.orc_cerp1(b - h, b, b + h, b + 2 * h, x)
        assert abs(v - (b + h * x)) < 1e-13
        v = lib.orc_cerp1(a, b, c, d, x)
        assert min(a, b, c, d) <= v <= max(a, b, c, d)


def test_frand_is_the_published_lcg():
    """frand 8:38-46: seed = (seed*1103515245 + 12345) mod 2^31, mantissa bits -> [0, 1)."""
    lib.orc_frand_seed(0xBA5EBA11)
    seed = 0xBA5EBA11
    for _ in range(2000):
        seed = (seed * 1103515245 + 12345) & 0x7FFFFFFF
        f = struct.unpack("<f", struct.pack("<I", (seed >> 8) | 0x3F800000))[0] - 1.0
        assert lib.orc_frand() == np.float32(f)


# ---------------------------------------------------------------- MIC(0) / PCG
def _dense_system(s, n_w, n_h):
    ad = s.field(ol.F_ADIAG)
    px = s.field(ol.F_APLUSX)
    py = s.field(ol.F_APLUSY)
    n = n_w * n_h
    A = np.diag(ad.copy())
    for i in range(n):
        x, y = i % n_w, i // n_w
        if x < n_w - 1:
            A[i, i + 1] = A[i + 1, i] = px[i]
        if y < n_h - 1:
            A[i, i + n_w] = A[i + n_w, i] = py[i]
    return A


def _factor(s, n_w, n_h):
    pc = s.field(ol.F_PRECON)
    px = s.field(ol.F_APLUSX)
    py = s.field(ol.F_APLUSY)
    n = n_w * n_h
    L = np.diag(1.0 / pc)
    for i in range(n):
        x, y = i % n_w, i // n_w
        if x < n_w - 1:
            L[i + 1, i] = px[i] * pc[i]
        if y < n_h - 1:
            L[i + n_w, i] = py[i] * pc[i]
    return L


def _solver3(w, h, steps=3):
    s = ol.OracleSolver(3, w, h, (0.1,))
    for _ in range(steps):
        s.add_inflow(0.3, 0.2, 0.4, 0.3, 1.0, 0.0, 0.5, 3.0)
        s.update(0.005)
    s.add_inflow(0.3, 0.2, 0.4, 0.3, 1.0, 0.0, 0.5, 3.0)
    s.stage("BUILD_RHS")
    s.stage("BUILD_PRESSURE", 0, 0.005)
    return s


def test_ic0_factor_reproduces_matrix_on_pattern():
    """With tau = 0 (and sigma = 0) buildPreconditioner 3:247-272 is IC(0):
    (L L^T)_ij = A_ij on the matrix pattern."""
    w, h = 9, 7
    s = _solver3(w, h)
    try:
        lib.orc_set_mic(0.0, 0.0)
        s.stage("BUILD_PRECON")
    finally:
        lib.orc_set_mic(0.97, 0.25)
    A = _dense_system(s, w, h)
    L = _factor(s, w, h)
    M = L @ L.T
    mask = A != 0
    np.testing.assert_allclose(M[mask], A[mask], rtol=1e-12, atol=1e-9 * np.abs(A).max())


def test_mic0_factor_preserves_row_sums():
    """With tau = 1 the modified factor keeps row sums: (L L^T) 1 = A 1."""
    w, h = 9, 7
    s = _solver3(w, h)
    try:
        lib.orc_set_mic(1.0, 0.0)
        s.stage("BUILD_PRECON")
    finally:
        lib.orc_set_mic(0.97, 0.25)
    A = _dense_system(s, w, h)
    L = _factor(s, w, h)
    d = (L @ L.T).sum(1) - A.sum(1)
    # the Neumann matrix is singular: the last pivot is round-off and the
    # safeguard e < sigma*A_ii -> e = A_ii (3:266-267) replaces it
    np.testing.assert_allclose(d[:-1], 0.0, rtol=0, atol=1e-9 * np.abs(A).max())
    pc = s.field(ol.F_PRECON)
    assert pc[-1] == 1.0 / math.sqrt(A[-1, -1])


def test_apply_preconditioner_is_the_factor_solve():
    """applyPreconditioner 3:275-304 returns (L L^T)^-1 r for the factor it built."""
    w, h = 8, 8
    s = _solver3(w, h)
    s.stage("BUILD_PRECON")
    r = s.field(ol.F_R)
    s.stage("APPLY_PRECON")
    z = s.field(ol.F_Z)
    L = _factor(s, w, h)
    want = np.linalg.solve(L @ L.T, r)
    np.testing.assert_allclose(z, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


def test_matvec_is_the_dense_product():
    w, h = 8, 8
    s = _solver3(w, h)
    rng = np.random.default_rng(3)
    sv = rng.normal(size=w * h)
    s.set_field(ol.F_S, sv)
    s.stage("MATVEC")
    A = _dense_system(s, w, h)
    np.testing.assert_allclose(s.field(ol.F_Z), A @ sv, rtol=1e-13, atol=1e-13 * np.abs(A).max())


def test_project_meets_its_residual_bound():
    """project() stops when max|r| < 1e-5; the true residual b - A p agrees."""
    w, h = 24, 20
    s = _solver3(w, h)
    s.stage("BUILD_PRECON")
    b = s.field(ol.F_R)
    it = s.stage("PROJECT", 600)
    assert 0 <= it < 600
    A = _dense_system(s, w, h)
    p = s.field(ol.F_P)
    assert np.abs(b - A @ p).max() < 2e-5


def test_zero_velocity_advection_is_identity():
    """RK3 with u = v = 0 lands on the sample point; cerp at x = 0 returns b."""
    w, h = 20, 16
    s = ol.OracleSolver(3, w, h, (0.1,))
    rng = np.random.default_rng(4)
    d = rng.normal(size=w * h)
    s.set_field(ol.qfield(ol.Q_D, ol.QF_SRC), d)
    s.stage("ADVECT", ol.Q_D, 0.005)
    out = s.field(ol.qfield(ol.Q_D, ol.QF_DST)).reshape(h, w)
    # the last row and column are clamped to w - 1.001 (7:405-406) and interpolate
    assert np.array_equal(out[:-1, :-1], d.reshape(h, w)[:-1, :-1])


def test_rhs_of_uniform_flow_is_zero_inside():
    """buildRhs 3:208-217: -div of a uniform field vanishes."""
    w, h = 16, 12
    s = ol.OracleSolver(3, w, h, (0.1,))
    s.set_field(ol.qfield(ol.Q_U, ol.QF_SRC), np.full((w + 1) * h, 0.7))
    s.set_field(ol.qfield(ol.Q_V, ol.QF_SRC), np.full(w * (h + 1), -0.2))
    s.stage("BUILD_RHS")
    assert np.all(s.field(ol.F_R) == 0.0)
