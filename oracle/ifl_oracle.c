/*
 * ifl_oracle.c -- TEST INFRASTRUCTURE ONLY: a scalar C restatement of the
 * reference's FluidSolver::update() for solvers 3..8 (and solver 1's
 * maxTimestep, the CFL bound), used as the checker for
 * the HIP product and as the timed CPU baseline. Only tests/, the smoke entry
 * and bench.py's CPU-baseline leg load it; the product never does.
 *
 * Each function cites the reference lines it restates (N:line means
 * N-<solver>/Fluid.cpp:line in the reference tree). Expressions keep the
 * reference's operand order and association so results are bit-identical
 * when compiled, like the reference, with -O2 and no FMA contraction.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - bit-identical trajectories and stdout lines against the reference
 *     compiled from its own sources (oracle/_ref, built by oracle/Makefile),
 *     and against the committed fixtures under tests/golden/; maxTimestep
 *     against solver 1 built the same way (oracle/_ref/libref1.so);
 *   - closed forms: occupancy of half-planes, SDFs, cubicPulse, frand's LCG,
 *     IC(0) factor identity at tau=0, dense-solve check of the MIC(0) sweeps
 *     and of the matrix-vector product on small grids.
 * Known, documented divergence: where the reference reads outside its arrays
 * (extrapolation reaching a solid cell on the domain border, 7:548-549 with
 * idx on row 0), the oracle reads 0.0 instead of undefined memory.
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ifl.h"

/* std::min / std::max semantics (NaN handling included), as the reference
 * calls them through <algorithm>. */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }
/* sgn / nsgn, 7:37-43 */
static inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }
static inline int nsgn(double v) { return v < 0.0 ? -1 : 1; }

/* length 7:45-47 */
static double length2(double x, double y) { return sqrt(x * x + y * y); }
/* cubicPulse 7:49-52 */
static double cubic_pulse(double x) {
    x = smin(fabs(x), 1.0);
    return 1.0 - x * x * (3.0 - 2.0 * x);
}
/* rotate 7:54-58 */
static void rotate(double* x, double* y, double phi) {
    double tx = *x, ty = *y;
    *x = cos(phi) * tx + sin(phi) * ty;
    *y = -sin(phi) * tx + cos(phi) * ty;
}

/* occupancy 7:60-102 */
static double tri_occ(double out1, double in, double out2) {
    return 0.5 * in * in / ((out1 - in) * (out2 - in));
}
static double trap_occ(double out1, double out2, double in1, double in2) {
    return 0.5 * (-in1 / (out1 - in1) - in2 / (out2 - in2));
}
double orc_occupancy(double d11, double d12, double d21, double d22) {
    double ds[4] = {d11, d12, d22, d21};
    uint8_t b = 0;
    for (int i = 3; i >= 0; i--) b = (uint8_t)((b << 1) | (ds[i] < 0.0 ? 1 : 0));
    switch (b) {
    case 0x0: return 0.0;
    case 0x1: return tri_occ(d21, d11, d12);
    case 0x2: return tri_occ(d11, d12, d22);
    case 0x4: return tri_occ(d12, d22, d21);
    case 0x8: return tri_occ(d22, d21, d11);
    case 0xE: return 1.0 - tri_occ(-d21, -d11, -d12);
    case 0xD: return 1.0 - tri_occ(-d11, -d12, -d22);
    case 0xB: return 1.0 - tri_occ(-d12, -d22, -d21);
    case 0x7: return 1.0 - tri_occ(-d22, -d21, -d11);
    case 0x3: return trap_occ(d21, d22, d11, d12);
    case 0x6: return trap_occ(d11, d21, d12, d22);
    case 0x9: return trap_occ(d12, d22, d11, d21);
    case 0xC: return trap_occ(d11, d12, d21, d22);
    case 0x5: return tri_occ(d11, d12, d22) + tri_occ(d22, d21, d11);
    case 0xA: return tri_occ(d21, d11, d12) + tri_occ(d12, d22, d21);
    case 0xF: return 1.0;
    }
    return 0.0;
}

/* ---------------- solid bodies, 7:109-257 ---------------- */
typedef struct {
    int type;
    double x, y, sx, sy, theta, vx, vy, vtheta;
} orc_body_pod;

typedef struct {
    int type;
    double pos_x, pos_y, scale_x, scale_y, theta, vel_x, vel_y, vel_theta;
} body_t;

double orc_body_distance(const body_t* b, double x, double y) {
    if (b->type == IFL_BODY_SPHERE) /* SolidSphere::distance 7:226-228 */
        return length2(x - b->pos_x, y - b->pos_y) - b->scale_x * 0.5;
    /* SolidBox::distance 7:175-186 */
    x -= b->pos_x;
    y -= b->pos_y;
    rotate(&x, &y, -b->theta);
    double dx = fabs(x) - b->scale_x * 0.5;
    double dy = fabs(y) - b->scale_y * 0.5;
    if (dx >= 0.0 || dy >= 0.0) return length2(smax(dx, 0.0), smax(dy, 0.0));
    return smax(dx, dy);
}

static void body_closest(const body_t* b, double* x, double* y) {
    if (b->type == IFL_BODY_SPHERE) { /* 7:230-243 with globalToLocal/localToGlobal 7:121-135 */
        *x -= b->pos_x;
        *y -= b->pos_y;
        rotate(x, y, -b->theta);
        *x /= b->scale_x;
        *y /= b->scale_y;
        double r = length2(*x, *y);
        if (r < 1e-4) {
            *x = 0.5;
            *y = 0.0;
        } else {
            *x /= 2.0 * r;
            *y /= 2.0 * r;
        }
        *x *= b->scale_x;
        *y *= b->scale_y;
        rotate(x, y, b->theta);
        *x += b->pos_x;
        *y += b->pos_y;
        return;
    }
    /* SolidBox::closestSurfacePoint 7:188-203 */
    *x -= b->pos_x;
    *y -= b->pos_y;
    rotate(x, y, -b->theta);
    double dx = fabs(*x) - b->scale_x * 0.5;
    double dy = fabs(*y) - b->scale_y * 0.5;
    if (dx > dy)
        *x = nsgn(*x) * 0.5 * b->scale_x;
    else
        *y = nsgn(*y) * 0.5 * b->scale_y;
    rotate(x, y, b->theta);
    *x += b->pos_x;
    *y += b->pos_y;
}

static void body_normal(const body_t* b, double* nx, double* ny, double x, double y) {
    if (b->type == IFL_BODY_SPHERE) { /* 7:245-256; r is a float (Q4) */
        x -= b->pos_x;
        y -= b->pos_y;
        float r = (float)length2(x, y);
        if (r < 1e-4) {
            *nx = 1.0;
            *ny = 0.0;
        } else {
            *nx = x / r;
            *ny = y / r;
        }
        return;
    }
    /* SolidBox::distanceNormal 7:205-217 */
    x -= b->pos_x;
    y -= b->pos_y;
    rotate(&x, &y, -b->theta);
    if (fabs(x) - b->scale_x * 0.5 > fabs(y) - b->scale_y * 0.5) {
        *nx = nsgn(x);
        *ny = 0.0;
    } else {
        *nx = 0.0;
        *ny = nsgn(y);
    }
    rotate(nx, ny, b->theta);
}

/* velocityX / velocityY 7:149-155 */
static double body_velx(const body_t* b, double x, double y) {
    (void)x;
    return (b->pos_y - y) * b->vel_theta + b->vel_x;
}
static double body_vely(const body_t* b, double x, double y) {
    (void)y;
    return (x - b->pos_x) * b->vel_theta + b->vel_y;
}

/* ---------------- FluidQuantity, 7:259-589 / 8:271-689 ---------------- */
typedef struct {
    int w, h;
    double ox, oy, hx;
    double *src, *dst; /* dst is _old at level 8 */
    double *phi, *volume, *normal_x, *normal_y;
    uint8_t *cell, *body, *mask;
} qty_t;

#define MAX_BODIES 255

typedef struct orc_solver {
    int level, w, h;
    double hx;
    double density, density_soot, diffusion, t_amb, g, flip_alpha;
    int g2p_cubic; /* the cubic gridToParticles variant (orc_set_flip) */
    qty_t q[4];
    double *r, *p, *z, *s, *precon, *adiag, *aplusx, *aplusy, *udens, *vdens;
    int nbodies;
    body_t bodies[MAX_BODIES];
    /* level 8 */
    int max_particles, particle_count;
    double *posx, *posy, *prop[4], *weight;
    int* counts;
    char* log;
    size_t log_len, log_cap;
} orc_solver;

/* frand 8:38-46: one process-global LCG, as in the reference. */
static uint32_t g_frand_seed = 0xBA5EBA11u;
static float frand(void) {
    g_frand_seed = (g_frand_seed * 1103515245u + 12345u) & 0x7FFFFFFFu;
    union { uint32_t i; float f; } u;
    u.i = (g_frand_seed >> 8) | 0x3F800000u;
    return (float)(u.f - 1.0);
}
double orc_frand(void) { return frand(); }
void orc_frand_seed(uint32_t seed) { g_frand_seed = seed; }
uint32_t orc_frand_state(void) { return g_frand_seed; }

static void slog(orc_solver* s, const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n < 0) return;
    if (s->log_len + (size_t)n + 1 > s->log_cap) {
        size_t cap = s->log_cap ? s->log_cap * 2 : 4096;
        while (cap < s->log_len + (size_t)n + 1) cap *= 2;
        s->log = (char*)realloc(s->log, cap);
        s->log_cap = cap;
    }
    memcpy(s->log + s->log_len, buf, (size_t)n + 1);
    s->log_len += (size_t)n;
}

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* FluidQuantity ctor 7:318-338 (level 3: 3:92-98, level 4: 4:304-318) */
static void qty_init(qty_t* q, int level, int w, int h, double ox, double oy, double hx) {
    q->w = w; q->h = h; q->ox = ox; q->oy = oy; q->hx = hx;
    size_t n = (size_t)w * h;
    q->src = dalloc(n);
    q->dst = dalloc(n);
    q->phi = dalloc((size_t)(w + 1) * (h + 1));
    q->volume = dalloc(n);
    q->normal_x = dalloc(n);
    q->normal_y = dalloc(n);
    q->cell = (uint8_t*)calloc(n, 1);
    q->body = (uint8_t*)calloc(n, 1);
    q->mask = (uint8_t*)calloc(n, 1);
    for (size_t i = 0; i < n; i++) q->volume[i] = (level >= 5) ? 1.0 : 0.0;
    (void)level;
}

static void qty_free(qty_t* q) {
    free(q->src); free(q->dst); free(q->phi); free(q->volume);
    free(q->normal_x); free(q->normal_y); free(q->cell); free(q->body); free(q->mask);
}

static inline double qat(const qty_t* q, int x, int y) { return q->src[x + y * q->w]; }

/* lerp(a, b, x) 7:277-279 */
static inline double lerp1(double a, double b, double x) { return a * (1.0 - x) + b * x; }

/* cerp(a, b, c, d, x) 7:281-295 */
static double cerp1(double a, double b, double c, double d, double x) {
    double xsq = x * x;
    double xcu = xsq * x;
    double minv = smin(a, smin(b, smin(c, d)));
    double maxv = smax(a, smax(b, smax(c, d)));
    double t = a * (0.0 - 0.5 * x + 1.0 * xsq - 0.5 * xcu) +
               b * (1.0 + 0.0 * x - 2.5 * xsq + 1.5 * xcu) +
               c * (0.0 + 0.5 * x + 2.0 * xsq - 1.5 * xcu) +
               d * (0.0 + 0.0 * x - 0.5 * xsq + 0.5 * xcu);
    return smin(smax(t, minv), maxv);
}

/* lerp(x, y) 7:390-402 */
static double qty_lerp(const qty_t* q, double x, double y) {
    x = smin(smax(x - q->ox, 0.0), q->w - 1.001);
    y = smin(smax(y - q->oy, 0.0), q->h - 1.001);
    int ix = (int)x;
    int iy = (int)y;
    x -= ix;
    y -= iy;
    double x00 = qat(q, ix + 0, iy + 0), x10 = qat(q, ix + 1, iy + 0);
    double x01 = qat(q, ix + 0, iy + 1), x11 = qat(q, ix + 1, iy + 1);
    return lerp1(lerp1(x00, x10, x), lerp1(x01, x11, x), y);
}

/* cerp(x, y) 7:404-422 */
static double qty_cerp(const qty_t* q, double x, double y) {
    x = smin(smax(x - q->ox, 0.0), q->w - 1.001);
    y = smin(smax(y - q->oy, 0.0), q->h - 1.001);
    int ix = (int)x;
    int iy = (int)y;
    x -= ix;
    y -= iy;
    int x0 = imax(ix - 1, 0), x1 = ix, x2 = ix + 1, x3 = imin(ix + 2, q->w - 1);
    int y0 = imax(iy - 1, 0), y1 = iy, y2 = iy + 1, y3 = imin(iy + 2, q->h - 1);
    double q0 = cerp1(qat(q, x0, y0), qat(q, x1, y0), qat(q, x2, y0), qat(q, x3, y0), x);
    double q1 = cerp1(qat(q, x0, y1), qat(q, x1, y1), qat(q, x2, y1), qat(q, x3, y1), x);
    double q2 = cerp1(qat(q, x0, y2), qat(q, x1, y2), qat(q, x2, y2), qat(q, x3, y2), x);
    double q3 = cerp1(qat(q, x0, y3), qat(q, x1, y3), qat(q, x2, y3), qat(q, x3, y3), x);
    return cerp1(q0, q1, q2, q3, y);
}

/* rungeKutta3 7:297-315 (backward in time; lastU not divided by _hx, Q1) */
static void rk3_back(const qty_t* self, double* x, double* y, double dt, const qty_t* u, const qty_t* v) {
    double hx = self->hx;
    double first_u = qty_lerp(u, *x, *y) / hx;
    double first_v = qty_lerp(v, *x, *y) / hx;
    double mid_x = *x - 0.5 * dt * first_u;
    double mid_y = *y - 0.5 * dt * first_v;
    double mid_u = qty_lerp(u, mid_x, mid_y) / hx;
    double mid_v = qty_lerp(v, mid_x, mid_y) / hx;
    double last_x = *x - 0.75 * dt * mid_u;
    double last_y = *y - 0.75 * dt * mid_v;
    double last_u = qty_lerp(u, last_x, last_y);
    double last_v = qty_lerp(v, last_x, last_y);
    *x -= dt * ((2.0 / 9.0) * first_u + (3.0 / 9.0) * mid_u + (4.0 / 9.0) * last_u);
    *y -= dt * ((2.0 / 9.0) * first_v + (3.0 / 9.0) * mid_v + (4.0 / 9.0) * last_v);
}

/* backProject 7:424-435 */
static void back_project(const orc_solver* s, const qty_t* q, double* x, double* y) {
    int rx = imin(imax((int)(*x - q->ox), 0), q->w - 1);
    int ry = imin(imax((int)(*y - q->oy), 0), q->h - 1);
    if (q->cell[rx + ry * q->w] != IFL_CELL_FLUID) {
        *x = (*x - q->ox) * q->hx;
        *y = (*y - q->oy) * q->hx;
        body_closest(&s->bodies[q->body[rx + ry * q->w]], x, y);
        *x = *x / q->hx + q->ox;
        *y = *y / q->hx + q->oy;
    }
}

/* advect 3:154-165 (level 3) / 7:437-454 (levels 4-7) */
static void qty_advect(const orc_solver* s, qty_t* q, double dt) {
    const qty_t* u = &s->q[IFL_Q_U];
    const qty_t* v = &s->q[IFL_Q_V];
    for (int iy = 0, idx = 0; iy < q->h; iy++) {
        for (int ix = 0; ix < q->w; ix++, idx++) {
            if (s->level >= 4 && q->cell[idx] != IFL_CELL_FLUID) continue;
            double x = ix + q->ox;
            double y = iy + q->oy;
            rk3_back(q, &x, &y, dt, u, v);
            if (s->level >= 4) back_project(s, q, &x, &y);
            q->dst[idx] = qty_cerp(q, x, y);
        }
    }
}

/* addInflow 7:456-473 (x bound by _h, Q3) */
static void qty_add_inflow(qty_t* q, double x0, double y0, double x1, double y1, double v) {
    int ix0 = (int)(x0 / q->hx - q->ox);
    int iy0 = (int)(y0 / q->hx - q->oy);
    int ix1 = (int)(x1 / q->hx - q->ox);
    int iy1 = (int)(y1 / q->hx - q->oy);
    long n = (long)q->w * q->h;
    for (int y = imax(iy0, 0); y < imin(iy1, q->h); y++) {
        for (int x = imax(ix0, 0); x < imin(ix1, q->h); x++) {
            double l = length2((2.0 * (x + 0.5) * q->hx - (x0 + x1)) / (x1 - x0),
                               (2.0 * (y + 0.5) * q->hx - (y0 + y1)) / (y1 - y0));
            double vi = cubic_pulse(l) * v;
            long idx = x + (long)y * q->w;
            if (idx >= n) continue; /* the reference writes past its array here */
            if (fabs(q->src[idx]) < fabs(vi)) q->src[idx] = vi;
        }
    }
}

/* fillSolidFields 4:449-480 (level 4) / 7:475-522 (levels 5-8) */
static void qty_fill_solid(const orc_solver* s, qty_t* q) {
    if (s->nbodies == 0) return;
    if (s->level >= 5) {
        for (int iy = 0, idx = 0; iy <= q->h; iy++) {
            for (int ix = 0; ix <= q->w; ix++, idx++) {
                double x = (ix + q->ox - 0.5) * q->hx;
                double y = (iy + q->oy - 0.5) * q->hx;
                q->phi[idx] = orc_body_distance(&s->bodies[0], x, y);
                for (int i = 1; i < s->nbodies; i++)
                    q->phi[idx] = smin(q->phi[idx], orc_body_distance(&s->bodies[i], x, y));
            }
        }
    }
    for (int iy = 0, idx = 0; iy < q->h; iy++) {
        for (int ix = 0; ix < q->w; ix++, idx++) {
            double x = (ix + q->ox) * q->hx;
            double y = (iy + q->oy) * q->hx;
            q->body[idx] = 0;
            double d = orc_body_distance(&s->bodies[0], x, y);
            for (int i = 1; i < s->nbodies; i++) {
                double id = orc_body_distance(&s->bodies[i], x, y);
                if (id < d) {
                    q->body[idx] = (uint8_t)i;
                    d = id;
                }
            }
            if (s->level == 4) {
                q->cell[idx] = d < 0.0 ? IFL_CELL_SOLID : IFL_CELL_FLUID;
                body_normal(&s->bodies[q->body[idx]], &q->normal_x[idx], &q->normal_y[idx], x, y);
                continue;
            }
            int idxp = ix + iy * (q->w + 1);
            q->volume[idx] = 1.0 - orc_occupancy(q->phi[idxp], q->phi[idxp + 1],
                                                 q->phi[idxp + q->w + 1], q->phi[idxp + q->w + 2]);
            if (q->volume[idx] < 0.01) q->volume[idx] = 0.0;
            body_normal(&s->bodies[q->body[idx]], &q->normal_x[idx], &q->normal_y[idx], x, y);
            q->cell[idx] = (q->volume[idx] == 0.0) ? IFL_CELL_SOLID : IFL_CELL_FLUID;
        }
    }
}

/* ---- extrapolation: an explicit LIFO stack, like the reference's std::stack ---- */
typedef struct {
    int* a;
    long n, cap;
} istack;
static void spush(istack* st, int v) {
    if (st->n == st->cap) {
        st->cap = st->cap ? st->cap * 2 : 1024;
        st->a = (int*)realloc(st->a, sizeof(int) * st->cap);
    }
    st->a[st->n++] = v;
}

static inline double qsrc_safe(const qty_t* q, long i) {
    return (i >= 0 && i < (long)q->w * q->h) ? q->src[i] : 0.0;
}
static inline double qnx_safe(const qty_t* q, long i) {
    return (i >= 0 && i < (long)q->w * q->h) ? q->normal_x[i] : 0.0;
}
static inline double qny_safe(const qty_t* q, long i) {
    return (i >= 0 && i < (long)q->w * q->h) ? q->normal_y[i] : 0.0;
}
static inline int qcell_safe(const qty_t* q, long i) {
    return (i >= 0 && i < (long)q->w * q->h) ? q->cell[i] : IFL_CELL_FLUID;
}

/* extrapolateNormal 7:544-552 */
static double extrap_normal(const qty_t* q, long idx) {
    double nx = q->normal_x[idx];
    double ny = q->normal_y[idx];
    double src_x = qsrc_safe(q, idx + sgn(nx));
    double src_y = qsrc_safe(q, idx + (long)sgn(ny) * q->w);
    return (fabs(nx) * src_x + fabs(ny) * src_y) / (fabs(nx) + fabs(ny));
}

/* fillSolidMask 7:524-542 + freeNeighbour 7:554-558 + extrapolate 7:560-588 */
static void qty_extrapolate_solid(qty_t* q) {
    int w = q->w, h = q->h;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 1; x < w - 1; x++) {
            int idx = x + y * w;
            if (q->cell[idx] == IFL_CELL_FLUID) continue;
            double nx = q->normal_x[idx];
            double ny = q->normal_y[idx];
            q->mask[idx] = 0;
            if (nx != 0.0 && q->cell[idx + sgn(nx)] != IFL_CELL_FLUID) q->mask[idx] |= 1;
            if (ny != 0.0 && q->cell[idx + sgn(ny) * w] != IFL_CELL_FLUID) q->mask[idx] |= 2;
        }
    }
    istack st = {0, 0, 0};
    for (int y = 1; y < h - 1; y++)
        for (int x = 1; x < w - 1; x++) {
            int idx = x + y * w;
            if (q->cell[idx] != IFL_CELL_FLUID && q->mask[idx] == 0) spush(&st, idx);
        }
    long n = (long)w * h;
    while (st.n > 0) {
        int idx = st.a[--st.n];
        q->src[idx] = extrap_normal(q, idx);
        long nb[4] = {idx - 1, idx + 1, idx - w, idx + w};
        int bit[4] = {1, 1, 2, 2};
        for (int k = 0; k < 4; k++) {
            long j = nb[k];
            int cond = (k == 0) ? qnx_safe(q, j) > 0.0 : (k == 1) ? qnx_safe(q, j) < 0.0
                     : (k == 2) ? qny_safe(q, j) > 0.0 : qny_safe(q, j) < 0.0;
            if (!cond || j < 0 || j >= n) continue;
            q->mask[j] &= (uint8_t)~bit[k];
            if (q->cell[j] != IFL_CELL_FLUID && q->mask[j] == 0) spush(&st, (int)j);
        }
    }
    free(st.a);
}

/* level 8: fillSolidMask 8:478-513, extrapolateAverage 8:528-545,
 * freeSolidNeighbour / freeEmptyNeighbour 8:547-565,
 * extrapolateEmptyBorders 8:570-609, extrapolate 8:611-651 */
static void qty_extrapolate_flip(qty_t* q) {
    int w = q->w, h = q->h;
    uint8_t* c = q->cell;
    for (int x = 0; x < w; x++) q->mask[x] = q->mask[x + (h - 1) * w] = 0xFF;
    for (int y = 0; y < h; y++) q->mask[y * w] = q->mask[y * w + w - 1] = 0xFF;
    for (int y = 1; y < h - 1; y++) {
        for (int x = 1; x < w - 1; x++) {
            int idx = x + y * w;
            q->mask[idx] = 0;
            if (c[idx] == IFL_CELL_SOLID) {
                double nx = q->normal_x[idx];
                double ny = q->normal_y[idx];
                if (nx != 0.0 && c[idx + sgn(nx)] != IFL_CELL_FLUID) q->mask[idx] |= 1;
                if (ny != 0.0 && c[idx + sgn(ny) * w] != IFL_CELL_FLUID) q->mask[idx] |= 2;
            } else if (c[idx] == IFL_CELL_EMPTY) {
                q->mask[idx] = c[idx - 1] != IFL_CELL_FLUID && c[idx + 1] != IFL_CELL_FLUID &&
                               c[idx - w] != IFL_CELL_FLUID && c[idx + w] != IFL_CELL_FLUID;
            }
        }
    }
    istack st = {0, 0, 0};
    for (int y = 1; y < h - 1; y++)
        for (int x = 1; x < w - 1; x++) {
            int idx = x + y * w;
            if (c[idx] != IFL_CELL_FLUID && q->mask[idx] == 0) spush(&st, idx);
        }
    while (st.n > 0) {
        int idx = st.a[--st.n];
        if (c[idx] == IFL_CELL_EMPTY) {
            double value = 0.0;
            int count = 0;
            if (c[idx - 1] == IFL_CELL_FLUID) { value += q->src[idx - 1]; count++; }
            if (c[idx + 1] == IFL_CELL_FLUID) { value += q->src[idx + 1]; count++; }
            if (c[idx - w] == IFL_CELL_FLUID) { value += q->src[idx - w]; count++; }
            if (c[idx + w] == IFL_CELL_FLUID) { value += q->src[idx + w]; count++; }
            q->src[idx] = value / count;
            c[idx] = IFL_CELL_FLUID;
        } else {
            q->src[idx] = extrap_normal(q, idx);
        }
        long nb[4] = {idx - 1, idx + 1, idx - w, idx + w};
        int bit[4] = {1, 1, 2, 2};
        for (int k = 0; k < 4; k++) {
            long j = nb[k];
            int cond = (k == 0) ? q->normal_x[j] > 0.0 : (k == 1) ? q->normal_x[j] < 0.0
                     : (k == 2) ? q->normal_y[j] > 0.0 : q->normal_y[j] < 0.0;
            if (cond && c[j] == IFL_CELL_SOLID) {
                q->mask[j] &= (uint8_t)~bit[k];
                if (q->mask[j] == 0) spush(&st, (int)j);
            }
        }
        for (int k = 0; k < 4; k++) {
            long j = nb[k];
            if (c[j] == IFL_CELL_EMPTY && q->mask[j] == 1) {
                q->mask[j] = 0;
                spush(&st, (int)j);
            }
        }
    }
    free(st.a);
    double* s = q->src;
    for (int x = 1; x < w - 1; x++) {
        int it = x, ib = x + (h - 1) * w;
        if (c[it] == IFL_CELL_EMPTY) s[it] = s[it + w];
        if (c[ib] == IFL_CELL_EMPTY) s[ib] = s[ib - w];
    }
    for (int y = 1; y < h - 1; y++) {
        int il = y * w, ir = y * w + w - 1;
        if (c[il] == IFL_CELL_EMPTY) s[il] = s[il + 1];
        if (c[ir] == IFL_CELL_EMPTY) s[ir] = s[ir - 1];
    }
    int tl = 0, tr = w - 1, bl = (h - 1) * w, br = h * w - 1;
    if (c[tl] == IFL_CELL_EMPTY) s[tl] = 0.5 * (s[tl + 1] + s[tl + w]);
    if (c[tr] == IFL_CELL_EMPTY) s[tr] = 0.5 * (s[tr - 1] + s[tr + w]);
    if (c[bl] == IFL_CELL_EMPTY) s[bl] = 0.5 * (s[bl + 1] + s[bl - w]);
    if (c[br] == IFL_CELL_EMPTY) s[br] = 0.5 * (s[br - 1] + s[br - w]);
    for (long i = 0; i < (long)w * h; i++)
        if (c[i] == IFL_CELL_EMPTY) c[i] = IFL_CELL_FLUID;
}

static void qty_extrapolate(const orc_solver* s, qty_t* q) {
    if (s->level == 8)
        qty_extrapolate_flip(q);
    else if (s->level >= 4)
        qty_extrapolate_solid(q);
}

/* ---------------- FluidSolver ---------------- */
static inline int fluid(const orc_solver* s, long i) {
    return s->level == 3 || s->q[IFL_Q_D].cell[i] == IFL_CELL_FLUID;
}
static inline int vec_masked(const orc_solver* s) { return s->level >= 6; }

/* buildRhs 3:208-217 / 4:614-627 / 7:624-653 */
static void build_rhs(orc_solver* s) {
    double scale = 1.0 / s->hx;
    const qty_t *u = &s->q[IFL_Q_U], *v = &s->q[IFL_Q_V], *d = &s->q[IFL_Q_D];
    int w = s->w, h = s->h;
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (s->level == 3) {
                s->r[idx] = -scale * (qat(u, x + 1, y) - qat(u, x, y) + qat(v, x, y + 1) - qat(v, x, y));
                continue;
            }
            if (d->cell[idx] != IFL_CELL_FLUID) {
                s->r[idx] = 0.0;
                continue;
            }
            if (s->level == 4) {
                s->r[idx] = -scale * (qat(u, x + 1, y) - qat(u, x, y) + qat(v, x, y + 1) - qat(v, x, y));
                continue;
            }
            double uv0 = u->volume[x + y * u->w], uv1 = u->volume[x + 1 + y * u->w];
            double vv0 = v->volume[x + y * v->w], vv1 = v->volume[x + (y + 1) * v->w];
            s->r[idx] = -scale * (uv1 * qat(u, x + 1, y) - uv0 * qat(u, x, y) +
                                  vv1 * qat(v, x, y + 1) - vv0 * qat(v, x, y));
            double vol = d->volume[idx];
            if (s->nbodies == 0) continue;
            const uint8_t* body = d->body;
            if (x > 0)
                s->r[idx] -= (uv0 - vol) * body_velx(&s->bodies[body[idx - 1]], x * s->hx, (y + 0.5) * s->hx);
            if (y > 0)
                s->r[idx] -= (vv0 - vol) * body_vely(&s->bodies[body[idx - w]], (x + 0.5) * s->hx, y * s->hx);
            if (x < w - 1)
                s->r[idx] += (uv1 - vol) * body_velx(&s->bodies[body[idx + 1]], (x + 1.0) * s->hx, (y + 0.5) * s->hx);
            if (y < h - 1)
                s->r[idx] += (vv1 - vol) * body_vely(&s->bodies[body[idx + w]], (x + 0.5) * s->hx, (y + 1.0) * s->hx);
        }
    }
}

/* computeDensities 7:658-675 */
static void compute_densities(orc_solver* s) {
    double alpha = (s->density_soot - s->density) / s->density;
    int w = s->w, h = s->h;
    memset(s->udens, 0, sizeof(double) * (w + 1) * h);
    memset(s->vdens, 0, sizeof(double) * w * (h + 1));
    const qty_t *t = &s->q[IFL_Q_T], *d = &s->q[IFL_Q_D];
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            double density = s->density * s->t_amb / qat(t, x, y) * (1.0 + alpha * qat(d, x, y));
            density = smax(density, 0.05 * s->density);
            s->udens[x + y * (w + 1)] += 0.5 * density;
            s->vdens[x + y * w] += 0.5 * density;
            s->udens[x + 1 + y * (w + 1)] += 0.5 * density;
            s->vdens[x + (y + 1) * w] += 0.5 * density;
        }
    }
}

/* buildPressureMatrix 3:222-244 / 4:629-654 / 5:686-713 / 6:649-676 / 7:680-707 */
static void build_pressure(orc_solver* s, double dt) {
    int w = s->w, h = s->h;
    size_t n = (size_t)w * h;
    const qty_t *u = &s->q[IFL_Q_U], *v = &s->q[IFL_Q_V];
    if (s->level == 3) {
        double scale = dt / (s->density * s->hx * s->hx);
        memset(s->adiag, 0, sizeof(double) * n);
        for (int y = 0, idx = 0; y < h; y++) {
            for (int x = 0; x < w; x++, idx++) {
                if (x < w - 1) {
                    s->adiag[idx] += scale;
                    s->adiag[idx + 1] += scale;
                    s->aplusx[idx] = -scale;
                } else
                    s->aplusx[idx] = 0.0;
                if (y < h - 1) {
                    s->adiag[idx] += scale;
                    s->adiag[idx + w] += scale;
                    s->aplusy[idx] = -scale;
                } else
                    s->aplusy[idx] = 0.0;
            }
        }
        return;
    }
    double scale = (s->level >= 7) ? dt / (s->hx * s->hx) : dt / (s->density * s->hx * s->hx);
    const uint8_t* cell = s->q[IFL_Q_D].cell;
    memset(s->adiag, 0, sizeof(double) * n);
    memset(s->aplusx, 0, sizeof(double) * n);
    memset(s->aplusy, 0, sizeof(double) * n);
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (cell[idx] != IFL_CELL_FLUID) continue;
            if (x < w - 1 && cell[idx + 1] == IFL_CELL_FLUID) {
                double factor;
                if (s->level == 4)
                    factor = scale;
                else if (s->level <= 6)
                    factor = scale * u->volume[x + 1 + y * u->w];
                else
                    factor = scale * u->volume[x + 1 + y * u->w] / s->udens[x + 1 + y * (w + 1)];
                s->adiag[idx] += factor;
                s->adiag[idx + 1] += factor;
                s->aplusx[idx] = -factor;
            }
            if (y < h - 1 && cell[idx + w] == IFL_CELL_FLUID) {
                double factor;
                if (s->level == 4)
                    factor = scale;
                else if (s->level <= 6)
                    factor = scale * v->volume[x + (y + 1) * v->w];
                else /* u-grid stride on the v density, Q2 (7:700) */
                    factor = scale * v->volume[x + (y + 1) * v->w] / s->vdens[x + (y + 1) * (w + 1)];
                s->adiag[idx] += factor;
                s->adiag[idx + w] += factor;
                s->aplusy[idx] = -factor;
            }
        }
    }
}

/* buildHeatDiffusionMatrix 7:709-737 */
static void build_heat(orc_solver* s, double dt) {
    int w = s->w, h = s->h;
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) s->adiag[i] = 1.0;
    memset(s->aplusx, 0, sizeof(double) * n);
    memset(s->aplusy, 0, sizeof(double) * n);
    const uint8_t* cell = s->q[IFL_Q_D].cell;
    double scale = s->diffusion * dt * 1.0 / (s->hx * s->hx);
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (cell[idx] != IFL_CELL_FLUID) continue;
            if (x < w - 1 && cell[idx + 1] == IFL_CELL_FLUID) {
                s->adiag[idx] += scale;
                s->adiag[idx + 1] += scale;
                s->aplusx[idx] = -scale;
            }
            if (y < h - 1 && cell[idx + w] == IFL_CELL_FLUID) {
                s->adiag[idx] += scale;
                s->adiag[idx + w] += scale;
                s->aplusy[idx] = -scale;
            }
        }
    }
}

static double g_tau = 0.97, g_sigma = 0.25;
void orc_set_mic(double tau, double sigma) { g_tau = tau; g_sigma = sigma; }

/* buildPreconditioner 3:247-272 (all fluid) / 7:739-768 (masked) */
static void build_precon(orc_solver* s) {
    const double tau = g_tau, sigma = g_sigma;
    int w = s->w, h = s->h;
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (!fluid(s, idx)) continue;
            double e = s->adiag[idx];
            if (x > 0 && fluid(s, idx - 1)) {
                double px = s->aplusx[idx - 1] * s->precon[idx - 1];
                double py = s->aplusy[idx - 1] * s->precon[idx - 1];
                e = e - (px * px + tau * px * py);
            }
            if (y > 0 && fluid(s, idx - w)) {
                double px = s->aplusx[idx - w] * s->precon[idx - w];
                double py = s->aplusy[idx - w] * s->precon[idx - w];
                e = e - (py * py + tau * px * py);
            }
            if (e < sigma * s->adiag[idx]) e = s->adiag[idx];
            s->precon[idx] = 1.0 / sqrt(e);
        }
    }
}

/* applyPreconditioner 3:275-304 / 7:770-804 */
static void apply_precon(orc_solver* s, double* dst, const double* a) {
    int w = s->w, h = s->h;
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (!fluid(s, idx)) continue;
            double t = a[idx];
            if (x > 0 && fluid(s, idx - 1)) t -= s->aplusx[idx - 1] * s->precon[idx - 1] * dst[idx - 1];
            if (y > 0 && fluid(s, idx - w)) t -= s->aplusy[idx - w] * s->precon[idx - w] * dst[idx - w];
            dst[idx] = t * s->precon[idx];
        }
    }
    for (int y = h - 1; y >= 0; y--) {
        for (int x = w - 1; x >= 0; x--) {
            int idx = x + y * w;
            if (!fluid(s, idx)) continue;
            double t = dst[idx];
            if (x < w - 1 && fluid(s, idx + 1)) t -= s->aplusx[idx] * s->precon[idx] * dst[idx + 1];
            if (y < h - 1 && fluid(s, idx + w)) t -= s->aplusy[idx] * s->precon[idx] * dst[idx + w];
            dst[idx] = t * s->precon[idx];
        }
    }
}

/* dotProduct 3:307-312 (unmasked, levels 3-5) / 7:806-814 (masked) */
static double dot(const orc_solver* s, const double* a, const double* b) {
    double result = 0.0;
    long n = (long)s->w * s->h;
    for (long i = 0; i < n; i++)
        if (!vec_masked(s) || s->q[IFL_Q_D].cell[i] == IFL_CELL_FLUID) result += a[i] * b[i];
    return result;
}

/* matrixVectorProduct 7:816-833 (unmasked at every level) */
static void matvec(const orc_solver* s, double* dst, const double* b) {
    int w = s->w, h = s->h;
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            double t = s->adiag[idx] * b[idx];
            if (x > 0) t += s->aplusx[idx - 1] * b[idx - 1];
            if (y > 0) t += s->aplusy[idx - w] * b[idx - w];
            if (x < w - 1) t += s->aplusx[idx] * b[idx + 1];
            if (y < h - 1) t += s->aplusy[idx] * b[idx + w];
            dst[idx] = t;
        }
    }
}

/* scaledAdd 3:335-338 / 7:835-841 */
static void scaled_add(const orc_solver* s, double* dst, const double* a, const double* b, double k) {
    long n = (long)s->w * s->h;
    for (long i = 0; i < n; i++)
        if (!vec_masked(s) || s->q[IFL_Q_D].cell[i] == IFL_CELL_FLUID) dst[i] = a[i] + b[i] * k;
}

/* infinityNorm 3:341-346 / 7:843-851 */
static double inf_norm(const orc_solver* s, const double* a) {
    double m = 0.0;
    long n = (long)s->w * s->h;
    for (long i = 0; i < n; i++)
        if (!vec_masked(s) || s->q[IFL_Q_D].cell[i] == IFL_CELL_FLUID) m = smax(m, fabs(a[i]));
    return m;
}

/* project 3:349-380 / 7:853-886. Returns the printed iteration, -1 for the
 * early exit, or `limit` on budget exhaustion. */
static int project(orc_solver* s, int limit) {
    long n = (long)s->w * s->h;
    memset(s->p, 0, sizeof(double) * n);
    apply_precon(s, s->z, s->r);
    memcpy(s->s, s->z, sizeof(double) * n);
    double max_error = inf_norm(s, s->r);
    if (max_error < 1e-5) {
        if (s->level >= 6) slog(s, "Initial guess sufficiently small\n");
        return -1;
    }
    double sigma = dot(s, s->z, s->r);
    for (int iter = 0; iter < limit; iter++) {
        matvec(s, s->z, s->s);
        double alpha = sigma / dot(s, s->z, s->s);
        scaled_add(s, s->p, s->p, s->s, alpha);
        scaled_add(s, s->r, s->r, s->z, -alpha);
        max_error = inf_norm(s, s->r);
        if (max_error < 1e-5) {
            slog(s, "Exiting solver after %d iterations, maximum error is %f\n", iter, max_error);
            return iter;
        }
        apply_precon(s, s->z, s->r);
        double sigma_new = dot(s, s->z, s->r);
        scaled_add(s, s->s, s->z, s->s, sigma_new / sigma);
        sigma = sigma_new;
    }
    slog(s, "Exceeded budget of %d iterations, maximum error was %f\n", limit, max_error);
    return limit;
}

/* applyPressure 3:382-398 / 4:794-809 / 6:863-879 / 7:891-906 */
static void apply_pressure(orc_solver* s, double dt) {
    int w = s->w, h = s->h;
    qty_t *u = &s->q[IFL_Q_U], *v = &s->q[IFL_Q_V];
    int uw = u->w, vw = v->w;
    if (s->level <= 6) {
        double scale = dt / (s->density * s->hx);
        for (int y = 0, idx = 0; y < h; y++) {
            for (int x = 0; x < w; x++, idx++) {
                if (!fluid(s, idx)) continue;
                u->src[x + y * uw] -= scale * s->p[idx];
                u->src[x + 1 + y * uw] += scale * s->p[idx];
                v->src[x + y * vw] -= scale * s->p[idx];
                v->src[x + (y + 1) * vw] += scale * s->p[idx];
            }
        }
        if (s->level == 3) {
            for (int y = 0; y < h; y++) u->src[0 + y * uw] = u->src[w + y * uw] = 0.0;
            for (int x = 0; x < w; x++) v->src[x] = v->src[x + h * vw] = 0.0;
        }
        return;
    }
    double scale = dt / s->hx;
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (!fluid(s, idx)) continue;
            u->src[x + y * uw] -= scale * s->p[idx] / s->udens[x + y * uw];
            v->src[x + y * vw] -= scale * s->p[idx] / s->vdens[x + y * vw];
            u->src[x + 1 + y * uw] += scale * s->p[idx] / s->udens[x + 1 + y * uw];
            v->src[x + (y + 1) * vw] += scale * s->p[idx] / s->vdens[x + (y + 1) * vw];
        }
    }
}

/* addBuoyancy 7:908-919 */
static void add_buoyancy(orc_solver* s, double dt) {
    double alpha = (s->density_soot - s->density) / s->density;
    const qty_t *d = &s->q[IFL_Q_D], *t = &s->q[IFL_Q_T];
    qty_t* v = &s->q[IFL_Q_V];
    for (int y = 0; y < s->h; y++) {
        for (int x = 0; x < s->w; x++) {
            double buoyancy = dt * s->g * (alpha * qat(d, x, y) - (qat(t, x, y) - s->t_amb) / s->t_amb);
            v->src[x + y * v->w] += buoyancy * 0.5;
            v->src[x + (y + 1) * v->w] += buoyancy * 0.5;
        }
    }
}

/* setBoundaryCondition 4:812-833 / 7:921-942 */
static void set_bc(orc_solver* s) {
    int w = s->w, h = s->h;
    qty_t *u = &s->q[IFL_Q_U], *v = &s->q[IFL_Q_V];
    const uint8_t *cell = s->q[IFL_Q_D].cell, *body = s->q[IFL_Q_D].body;
    double hx = s->hx;
    for (int y = 0, idx = 0; y < h; y++) {
        for (int x = 0; x < w; x++, idx++) {
            if (cell[idx] == IFL_CELL_SOLID) {
                const body_t* b = &s->bodies[body[idx]];
                u->src[x + y * u->w] = body_velx(b, x * hx, (y + 0.5) * hx);
                v->src[x + y * v->w] = body_vely(b, (x + 0.5) * hx, y * hx);
                u->src[x + 1 + y * u->w] = body_velx(b, (x + 1.0) * hx, (y + 0.5) * hx);
                v->src[x + (y + 1) * v->w] = body_vely(b, (x + 0.5) * hx, (y + 1.0) * hx);
            }
        }
    }
    for (int y = 0; y < h; y++) u->src[0 + y * u->w] = u->src[w + y * u->w] = 0.0;
    for (int x = 0; x < w; x++) v->src[x] = v->src[x + h * v->w] = 0.0;
}

static void qty_flip(qty_t* q) {
    double* t = q->src;
    q->src = q->dst;
    q->dst = t;
}

/* ---------------- level 8: particles, 8:692-940 ---------------- */
#define MAX_PER_CELL 12
#define MIN_PER_CELL 3
#define AVG_PER_CELL 4

/* pointInBody 8:726-732 */
static int point_in_body(const orc_solver* s, double x, double y) {
    for (int i = 0; i < s->nbodies; i++)
        if (orc_body_distance(&s->bodies[i], x * s->hx, y * s->hx) < 0.0) return 1;
    return 0;
}

/* initParticles 8:735-751 */
static void init_particles(orc_solver* s) {
    int idx = 0;
    for (int y = 0; y < s->h; y++)
        for (int x = 0; x < s->w; x++)
            for (int i = 0; i < AVG_PER_CELL; i++, idx++) {
                s->posx[idx] = x + frand();
                s->posy[idx] = y + frand();
                if (point_in_body(s, s->posx[idx], s->posy[idx])) idx--;
            }
    s->particle_count = idx;
}

/* countParticles 8:754-763 */
static void count_particles(orc_solver* s) {
    memset(s->counts, 0, sizeof(int) * s->w * s->h);
    for (int i = 0; i < s->particle_count; i++) {
        int ix = (int)s->posx[i];
        int iy = (int)s->posy[i];
        if (ix >= 0 && iy >= 0 && ix < s->w && iy < s->h) s->counts[ix + iy * s->w]++;
    }
}

/* pruneParticles 8:766-786 */
static void prune_particles(orc_solver* s) {
    for (int i = 0; i < s->particle_count; i++) {
        int ix = (int)s->posx[i];
        int iy = (int)s->posy[i];
        int idx = ix + iy * s->w;
        if (ix < 0 || iy < 0 || ix >= s->w || iy >= s->h) continue;
        if (s->counts[idx] > MAX_PER_CELL) {
            int j = --s->particle_count;
            s->posx[i] = s->posx[j];
            s->posy[i] = s->posy[j];
            for (int t = 0; t < 4; t++) s->prop[t][i] = s->prop[t][j];
            s->counts[idx]--;
            i--;
        }
    }
}

/* seedParticles 8:789-813 (rejection tests particle `idx`, Q7) */
static void seed_particles(orc_solver* s) {
    for (int y = 0, idx = 0; y < s->h; y++) {
        for (int x = 0; x < s->w; x++, idx++) {
            for (int i = 0; i < MIN_PER_CELL - s->counts[idx]; i++) {
                if (s->particle_count == s->max_particles) return;
                int j = s->particle_count;
                s->posx[j] = x + frand();
                s->posy[j] = y + frand();
                if (point_in_body(s, s->posx[idx], s->posy[idx])) continue;
                for (int t = 0; t < 4; t++) s->prop[t][j] = qty_lerp(&s->q[t], s->posx[j], s->posy[j]);
                s->particle_count++;
            }
        }
    }
}

/* addSample 8:296-303 and fromParticles 8:663-688 */
static void from_particles(orc_solver* s, qty_t* q, const double* property) {
    int w = q->w, h = q->h;
    memset(q->src, 0, sizeof(double) * w * h);
    memset(s->weight, 0, sizeof(double) * w * h);
    for (int i = 0; i < s->particle_count; i++) {
        double x = s->posx[i] - q->ox;
        double y = s->posy[i] - q->oy;
        x = smax(0.5, smin(w - 1.5, x));
        y = smax(0.5, smin(h - 1.5, y));
        int ix = (int)x;
        int iy = (int)y;
        for (int k = 0; k < 4; k++) {
            int cx = ix + (k & 1), cy = iy + (k >> 1);
            if (cx < 0 || cy < 0 || cx >= w || cy >= h) continue;
            double kk = (1.0 - fabs(cx - x)) * (1.0 - fabs(cy - y));
            s->weight[cx + cy * w] += kk;
            q->src[cx + cy * w] += kk * property[i];
        }
    }
    for (long i = 0; i < (long)w * h; i++) {
        if (s->weight[i] != 0.0)
            q->src[i] /= s->weight[i];
        else if (q->cell[i] == IFL_CELL_FLUID)
            q->cell[i] = IFL_CELL_EMPTY;
    }
}

/* gridToParticles 8:904-911 */
/* cubic: the variant that samples with the Catmull-Rom cerp of solvers 3-7
 * (7:404-422) instead of lerp; not in the reference */
static void grid_to_particles(orc_solver* s, double alpha, int cubic) {
    for (int t = 0; t < 4; t++)
        for (int i = 0; i < s->particle_count; i++) {
            s->prop[t][i] *= 1.0 - alpha;
            s->prop[t][i] += cubic ? qty_cerp(&s->q[t], s->posx[i], s->posy[i])
                                   : qty_lerp(&s->q[t], s->posx[i], s->posy[i]);
        }
}

/* particlesToGrid 8:916-927 */
static void particles_to_grid(orc_solver* s) {
    for (int t = 0; t < 4; t++) {
        from_particles(s, &s->q[t], s->prop[t]);
        qty_extrapolate(s, &s->q[t]);
    }
    count_particles(s);
    prune_particles(s);
    seed_particles(s);
    slog(s, "Particle count: %d\n", s->particle_count);
}

/* ParticleQuantities::backProject 8:816-839 */
static void particle_back_project(const orc_solver* s, double* x, double* y) {
    double d = 1e30;
    int closest = -1;
    for (int i = 0; i < s->nbodies; i++) {
        double id = orc_body_distance(&s->bodies[i], *x * s->hx, *y * s->hx);
        if (id < d) {
            d = id;
            closest = i;
        }
    }
    if (d < -1.0) {
        *x *= s->hx;
        *y *= s->hx;
        body_closest(&s->bodies[closest], x, y);
        double nx, ny;
        body_normal(&s->bodies[closest], &nx, &ny, *x, *y);
        *x -= nx * s->hx;
        *y -= ny * s->hx;
        *x /= s->hx;
        *y /= s->hx;
    }
}

/* ParticleQuantities::rungeKutta3 8:844-862 (forward in time) and advect 8:931-939 */
static void particle_advect(orc_solver* s, double dt) {
    const qty_t *u = &s->q[IFL_Q_U], *v = &s->q[IFL_Q_V];
    double hx = s->hx;
    for (int i = 0; i < s->particle_count; i++) {
        double x = s->posx[i], y = s->posy[i];
        double first_u = qty_lerp(u, x, y) / hx;
        double first_v = qty_lerp(v, x, y) / hx;
        double mid_x = x + 0.5 * dt * first_u;
        double mid_y = y + 0.5 * dt * first_v;
        double mid_u = qty_lerp(u, mid_x, mid_y) / hx;
        double mid_v = qty_lerp(v, mid_x, mid_y) / hx;
        double last_x = x + 0.75 * dt * mid_u;
        double last_y = y + 0.75 * dt * mid_v;
        double last_u = qty_lerp(u, last_x, last_y);
        double last_v = qty_lerp(v, last_x, last_y);
        x += dt * ((2.0 / 9.0) * first_u + (3.0 / 9.0) * mid_u + (4.0 / 9.0) * last_u);
        y += dt * ((2.0 / 9.0) * first_v + (3.0 / 9.0) * mid_v + (4.0 / 9.0) * last_v);
        particle_back_project(s, &x, &y);
        s->posx[i] = smax(smin(x, s->w - 0.001), 0.0);
        s->posy[i] = smax(smin(y, s->h - 0.001), 0.0);
    }
}

/* copy / diff / undiff 8:374-388 */
static void qty_copy_old(qty_t* q) { memcpy(q->dst, q->src, sizeof(double) * q->w * q->h); }
static void qty_diff(qty_t* q, double alpha) {
    for (long i = 0; i < (long)q->w * q->h; i++) q->src[i] -= (1.0 - alpha) * q->dst[i];
}
static void qty_undiff(qty_t* q, double alpha) {
    for (long i = 0; i < (long)q->w * q->h; i++) q->src[i] += (1.0 - alpha) * q->dst[i];
}

/* maxTimestep 1:310-328 (solver 1): the largest speed at the cell centres,
 * each component lerp'd from its faces, gives dt = 2h / speed, at most 1. */
static double max_timestep(const orc_solver* s) {
    double max_velocity = 0.0;
    for (int y = 0; y < s->h; y++)
        for (int x = 0; x < s->w; x++) {
            double u = qty_lerp(&s->q[IFL_Q_U], x + 0.5, y + 0.5);
            double v = qty_lerp(&s->q[IFL_Q_V], x + 0.5, y + 0.5);
            double velocity = sqrt(u * u + v * v);
            max_velocity = smax(max_velocity, velocity);
        }
    double max_timestep = 2.0 * s->hx / max_velocity;
    return smin(max_timestep, 1.0);
}

/* (int) of a double as x86-64 converts it: NaN and out-of-range give INT_MIN */
static int x86_int(double v) {
    return (v == v && v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000u;
}

/* toImage 3:455-465 / 4:903-916 / 5:961-972 / 6:1021-1055 / 7:1047-1081 /
 * 8:1426-1467: RGBA bytes, w*h*4, or w*2*h*4 with the heat panel on the left
 * (levels 6-8). Stores of an int into the byte array keep its low 8 bits. */
static void to_image(const orc_solver* s, unsigned char* rgba, int render_heat) {
    int w = s->w, h = s->h, lv = s->level;
    const qty_t* d = &s->q[IFL_Q_D];
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int i = x + y * w;
            if (lv <= 5) {
                int shade = lv == 5 ? x86_int((1.0 - d->src[i]) * d->volume[i] * 255.0)
                                    : x86_int((1.0 - d->src[i]) * 255.0);
                shade = shade < 255 ? shade : 255;
                shade = 0 < shade ? shade : 0;
                if (lv == 4 && d->cell[i] == IFL_CELL_SOLID) shade = 0;
                rgba[i * 4 + 0] = rgba[i * 4 + 1] = rgba[i * 4 + 2] = (unsigned char)shade;
                rgba[i * 4 + 3] = 0xFF;
                continue;
            }
            int idxl = 0, idxr;
            if (render_heat) {
                idxl = 4 * (x + y * w * 2);
                idxr = 4 * (x + y * w * 2 + w);
            } else {
                idxr = 4 * i;
            }
            double volume = d->volume[i];
            double shade = (1.0 - d->src[i]) * volume;
            shade = smin(smax(shade, 0.0), 1.0);
            unsigned char g = (unsigned char)x86_int(shade * 255.0);
            rgba[idxr + 0] = rgba[idxr + 1] = rgba[idxr + 2] = g;
            rgba[idxr + 3] = 0xFF;
            if (lv == 8 && d->cell[i] == IFL_CELL_EMPTY) { /* 8:1445-1448 */
                rgba[idxr] = 0xFF;
                rgba[idxr + 1] = rgba[idxr + 2] = 0;
            }
            if (render_heat) {
                double t = lv == 8 ? fabs(s->q[IFL_Q_T].src[i] - s->t_amb) / 70.0   /* 8:1451 */
                                   : (s->q[IFL_Q_T].src[i] - s->t_amb) / 700.0;
                t = smin(smax(t, 0.0), 1.0);
                double r = 1.0 + volume * (smin(t * 4.0, 1.0) - 1.0);
                double gg = 1.0 + volume * (smin(t * 2.0, 1.0) - 1.0);
                double b = 1.0 + volume * (smax(smin(t * 4.0 - 3.0, 1.0), 0.0) - 1.0);
                rgba[idxl + 0] = (unsigned char)x86_int(r * 255.0);
                rgba[idxl + 1] = (unsigned char)x86_int(gg * 255.0);
                rgba[idxl + 2] = (unsigned char)x86_int(b * 255.0);
                rgba[idxl + 3] = 0xFF;
            }
        }
}

/* ---------------- public API (ctypes) ---------------- */

orc_solver* orc_create(int level, int w, int h, const double* params, const orc_body_pod* bodies, int nbodies) {
    if (level < 3 || level > 8 || w < 3 || h < 3 || nbodies < 0 || nbodies > MAX_BODIES) return NULL;
    orc_solver* s = (orc_solver*)calloc(1, sizeof(orc_solver));
    s->level = level;
    s->w = w;
    s->h = h;
    s->hx = 1.0 / imin(w, h); /* 7:953 */
    s->density = params[0];
    if (level >= 6) {
        s->density_soot = params[1];
        s->diffusion = params[2];
    }
    s->t_amb = 294.0; /* 7:950-951 */
    s->g = 9.81;
    s->flip_alpha = 0.001; /* 8:1296 */
    s->nbodies = level >= 4 ? nbodies : 0;
    for (int i = 0; i < s->nbodies; i++) {
        body_t* b = &s->bodies[i];
        b->type = bodies[i].type;
        b->pos_x = bodies[i].x;
        b->pos_y = bodies[i].y;
        b->scale_x = bodies[i].sx;
        /* SolidSphere(x, y, s, ...) passes s for both scales, 7:223-224 */
        b->scale_y = bodies[i].type == IFL_BODY_SPHERE ? bodies[i].sx : bodies[i].sy;
        b->theta = bodies[i].theta;
        b->vel_x = bodies[i].vx;
        b->vel_y = bodies[i].vy;
        b->vel_theta = bodies[i].vtheta;
    }
    /* FluidSolver ctor 7:953-973 */
    qty_init(&s->q[IFL_Q_D], level, w, h, 0.5, 0.5, s->hx);
    qty_init(&s->q[IFL_Q_T], level, w, h, 0.5, 0.5, s->hx);
    qty_init(&s->q[IFL_Q_U], level, w + 1, h, 0.0, 0.5, s->hx);
    qty_init(&s->q[IFL_Q_V], level, w, h + 1, 0.5, 0.0, s->hx);
    if (level >= 6)
        for (long i = 0; i < (long)w * h; i++) s->q[IFL_Q_T].src[i] = s->t_amb;
    size_t n = (size_t)w * h;
    s->r = dalloc(n); s->p = dalloc(n); s->z = dalloc(n); s->s = dalloc(n);
    s->precon = dalloc(n); s->adiag = dalloc(n); s->aplusx = dalloc(n); s->aplusy = dalloc(n);
    s->udens = dalloc((size_t)(w + 1) * h);
    s->vdens = dalloc((size_t)w * (h + 1));
    if (level == 8) { /* 8:865-878, 8:1308-1314 */
        s->max_particles = w * h * MAX_PER_CELL;
        s->posx = dalloc((size_t)s->max_particles);
        s->posy = dalloc((size_t)s->max_particles);
        for (int t = 0; t < 4; t++) s->prop[t] = dalloc((size_t)s->max_particles);
        s->weight = dalloc((size_t)(w + 1) * (h + 1));
        s->counts = (int*)calloc(n, sizeof(int));
        init_particles(s);
        grid_to_particles(s, 1.0, 0);
    }
    return s;
}

/* level 8: the FLIP update's blend factor and gridToParticles interpolation
 * (0 lerp as the reference, 1 the Catmull-Rom variant) */
void orc_set_flip(orc_solver* s, double alpha, int cubic) {
    s->flip_alpha = alpha;
    s->g2p_cubic = cubic;
}

void orc_destroy(orc_solver* s) {
    if (!s) return;
    for (int i = 0; i < 4; i++) qty_free(&s->q[i]);
    free(s->r); free(s->p); free(s->z); free(s->s); free(s->precon);
    free(s->adiag); free(s->aplusx); free(s->aplusy); free(s->udens); free(s->vdens);
    free(s->posx); free(s->posy);
    for (int t = 0; t < 4; t++) free(s->prop[t]);
    free(s->weight); free(s->counts); free(s->log);
    free(s);
}

/* SolidBody::update 7:162-166 */
void orc_bodies_update(orc_solver* s, double dt) {
    for (int i = 0; i < s->nbodies; i++) {
        body_t* b = &s->bodies[i];
        b->pos_x += b->vel_x * dt;
        b->pos_y += b->vel_y * dt;
        b->theta += b->vel_theta * dt;
    }
}

/* FluidSolver::addInflow 3:449-453 / 7:1036-1041 */
void orc_add_inflow(orc_solver* s, double x, double y, double w, double h, double d, double t, double u, double v) {
    qty_add_inflow(&s->q[IFL_Q_D], x, y, x + w, y + h, d);
    if (s->level >= 6) qty_add_inflow(&s->q[IFL_Q_T], x, y, x + w, y + h, t);
    qty_add_inflow(&s->q[IFL_Q_U], x, y, x + w, y + h, u);
    qty_add_inflow(&s->q[IFL_Q_V], x, y, x + w, y + h, v);
}

/* update 3:433-447 / 4:869-895 / 5:927-953 / 6:966-1008 / 7:995-1034 / 8:1350-1413.
 * iters[0] = heat solve result, iters[1] = pressure solve result (see project). */
void orc_update(orc_solver* s, double dt, int* iters) {
    int lv = s->level;
    int heat = -2, press;
    size_t nb = sizeof(double) * s->w * s->h;
    if (lv == 3) {
        build_rhs(s);
        build_pressure(s, dt);
        build_precon(s);
        press = project(s, 600);
        apply_pressure(s, dt);
        for (int q = 0; q < 4; q++) {
            if (q == IFL_Q_T) continue;
            qty_advect(s, &s->q[q], dt);
        }
        qty_flip(&s->q[IFL_Q_D]); qty_flip(&s->q[IFL_Q_U]); qty_flip(&s->q[IFL_Q_V]);
        if (iters) { iters[0] = heat; iters[1] = press; }
        return;
    }
    qty_fill_solid(s, &s->q[IFL_Q_D]);
    if (lv >= 6) qty_fill_solid(s, &s->q[IFL_Q_T]);
    qty_fill_solid(s, &s->q[IFL_Q_U]);
    qty_fill_solid(s, &s->q[IFL_Q_V]);
    if (lv == 8) {
        particles_to_grid(s);
        for (int q = 0; q < 4; q++) qty_copy_old(&s->q[q]);
        orc_add_inflow(s, 0.45, 0.2, 0.2, 0.05, 1.0, s->t_amb, 0.0, 0.0); /* 8:1368 */
    }
    if (lv >= 6) {
        memcpy(s->r, s->q[IFL_Q_T].src, nb);
        build_heat(s, dt);
        build_precon(s);
        heat = project(s, 2000);
        memcpy(s->q[IFL_Q_T].src, s->p, nb);
        qty_extrapolate(s, &s->q[IFL_Q_T]);
        add_buoyancy(s, dt);
    }
    set_bc(s);
    build_rhs(s);
    if (lv >= 7) compute_densities(s);
    build_pressure(s, dt);
    build_precon(s);
    press = project(s, 2000);
    apply_pressure(s, dt);
    qty_extrapolate(s, &s->q[IFL_Q_D]);
    qty_extrapolate(s, &s->q[IFL_Q_U]);
    qty_extrapolate(s, &s->q[IFL_Q_V]);
    set_bc(s);
    if (lv == 8) {
        for (int q = 0; q < 4; q++) qty_diff(&s->q[q], s->flip_alpha);
        grid_to_particles(s, s->flip_alpha, s->g2p_cubic);
        for (int q = 0; q < 4; q++) qty_undiff(&s->q[q], s->flip_alpha);
        particle_advect(s, dt);
    } else {
        qty_advect(s, &s->q[IFL_Q_D], dt);
        if (lv >= 6) qty_advect(s, &s->q[IFL_Q_T], dt);
        qty_advect(s, &s->q[IFL_Q_U], dt);
        qty_advect(s, &s->q[IFL_Q_V], dt);
        qty_flip(&s->q[IFL_Q_D]);
        if (lv >= 6) qty_flip(&s->q[IFL_Q_T]);
        qty_flip(&s->q[IFL_Q_U]);
        qty_flip(&s->q[IFL_Q_V]);
    }
    if (iters) { iters[0] = heat; iters[1] = press; }
}

const char* orc_log(orc_solver* s) { return s->log ? s->log : ""; }
void orc_log_clear(orc_solver* s) {
    s->log_len = 0;
    if (s->log) s->log[0] = 0;
}

int orc_particle_count(orc_solver* s, int set, int value) {
    if (set) s->particle_count = value;
    return s->particle_count;
}

void* orc_field(orc_solver* s, int field, long* count, int* elem) {
    *count = 0;
    *elem = 8;
    int w = s->w, h = s->h;
    long n = (long)w * h;
    if (field < 64) {
        int qi = field / 16, f = field % 16;
        if (qi > 3 || (qi == IFL_Q_T && s->level < 6)) return NULL;
        qty_t* q = &s->q[qi];
        long qn = (long)q->w * q->h;
        *count = qn;
        switch (f) {
        case IFL_QF_SRC: return q->src;
        case IFL_QF_DST: return q->dst;
        case IFL_QF_PHI: *count = (long)(q->w + 1) * (q->h + 1); return q->phi;
        case IFL_QF_VOLUME: return q->volume;
        case IFL_QF_NORMALX: return q->normal_x;
        case IFL_QF_NORMALY: return q->normal_y;
        case IFL_QF_CELL: *elem = 1; return q->cell;
        case IFL_QF_BODY: *elem = 1; return q->body;
        case IFL_QF_MASK: *elem = 1; return q->mask;
        }
        *count = 0;
        return NULL;
    }
    switch (field) {
    case IFL_FIELD_R: *count = n; return s->r;
    case IFL_FIELD_P: *count = n; return s->p;
    case IFL_FIELD_Z: *count = n; return s->z;
    case IFL_FIELD_S: *count = n; return s->s;
    case IFL_FIELD_PRECON: *count = n; return s->precon;
    case IFL_FIELD_ADIAG: *count = n; return s->adiag;
    case IFL_FIELD_APLUSX: *count = n; return s->aplusx;
    case IFL_FIELD_APLUSY: *count = n; return s->aplusy;
    case IFL_FIELD_UDENSITY: *count = (long)(w + 1) * h; return s->udens;
    case IFL_FIELD_VDENSITY: *count = (long)w * (h + 1); return s->vdens;
    }
    if (s->level != 8) return NULL;
    switch (field) {
    case IFL_FIELD_POSX: *count = s->max_particles; return s->posx;
    case IFL_FIELD_POSY: *count = s->max_particles; return s->posy;
    case IFL_FIELD_PROP_D: case IFL_FIELD_PROP_T: case IFL_FIELD_PROP_U: case IFL_FIELD_PROP_V:
        *count = s->max_particles;
        return s->prop[field - IFL_FIELD_PROP_D];
    case IFL_FIELD_WEIGHT: *count = (long)(w + 1) * (h + 1); return s->weight;
    case IFL_FIELD_COUNTS: *count = n; *elem = 4; return s->counts;
    }
    return NULL;
}

/* Same stage ids as ifl_stage(); returns 0, 1 (bad arg) or 4 (not at this level).
 * For PROJECT, *result receives the value project() returned. */
int orc_stage(orc_solver* s, int stage, int iarg, double darg, double* result) {
    int lv = s->level;
    qty_t* q = (iarg >= 0 && iarg < 4 && !(iarg == IFL_Q_T && lv < 6)) ? &s->q[iarg] : NULL;
    switch (stage) {
    case IFL_STAGE_FILL_SOLID: if (lv < 4) return 4; if (!q) return 1; qty_fill_solid(s, q); return 0;
    case IFL_STAGE_EXTRAPOLATE: if (lv < 4) return 4; if (!q) return 1; qty_extrapolate(s, q); return 0;
    case IFL_STAGE_ADVECT: if (lv == 8) return 4; if (!q) return 1; qty_advect(s, q, darg); return 0;
    case IFL_STAGE_FLIP: if (lv == 8) return 4; if (!q) return 1; qty_flip(q); return 0;
    case IFL_STAGE_BUILD_RHS: build_rhs(s); return 0;
    case IFL_STAGE_BUILD_PRESSURE: build_pressure(s, darg); return 0;
    case IFL_STAGE_BUILD_HEAT: if (lv < 6) return 4; build_heat(s, darg); return 0;
    case IFL_STAGE_BUILD_PRECON: build_precon(s); return 0;
    case IFL_STAGE_APPLY_PRECON: apply_precon(s, s->z, s->r); return 0;
    case IFL_STAGE_MATVEC: matvec(s, s->z, s->s); return 0;
    case IFL_STAGE_DOT: *result = dot(s, s->z, iarg == 0 ? s->r : s->s); return 0;
    case IFL_STAGE_INF_NORM: *result = inf_norm(s, s->r); return 0;
    case IFL_STAGE_PROJECT: { int it = project(s, iarg); if (result) *result = it; return 0; }
    case IFL_STAGE_APPLY_PRESSURE: apply_pressure(s, darg); return 0;
    case IFL_STAGE_ADD_BUOYANCY: if (lv < 6) return 4; add_buoyancy(s, darg); return 0;
    case IFL_STAGE_SET_BOUNDARY: if (lv < 4) return 4; set_bc(s); return 0;
    case IFL_STAGE_COMPUTE_DENSITIES: if (lv < 7) return 4; compute_densities(s); return 0;
    case IFL_STAGE_COPY_T_TO_R:
        if (lv < 6) return 4;
        memcpy(s->r, s->q[IFL_Q_T].src, sizeof(double) * s->w * s->h);
        return 0;
    case IFL_STAGE_COPY_P_TO_T:
        if (lv < 6) return 4;
        memcpy(s->q[IFL_Q_T].src, s->p, sizeof(double) * s->w * s->h);
        return 0;
    }
    if (lv != 8) return 4;
    switch (stage) {
    case IFL_STAGE_FROM_PARTICLES: if (!q) return 1; from_particles(s, q, s->prop[iarg]); return 0;
    case IFL_STAGE_G2P: grid_to_particles(s, darg, 0); return 0;
    case IFL_STAGE_P2G: particles_to_grid(s); return 0;
    case IFL_STAGE_PARTICLE_ADVECT: particle_advect(s, darg); return 0;
    case IFL_STAGE_COPY_OLD: if (!q) return 1; qty_copy_old(q); return 0;
    case IFL_STAGE_DIFF: if (!q) return 1; qty_diff(q, darg); return 0;
    case IFL_STAGE_UNDIFF: if (!q) return 1; qty_undiff(q, darg); return 0;
    case IFL_STAGE_COUNT_PARTICLES: count_particles(s); return 0;
    case IFL_STAGE_PRUNE: prune_particles(s); return 0;
    case IFL_STAGE_SEED: seed_particles(s); return 0;
    }
    return 4;
}

/* Closed-form probes for the oracle's own pin tests. */
double orc_body_distance_pod(const orc_body_pod* p, double x, double y) {
    body_t b = {p->type, p->x, p->y, p->sx, p->type == IFL_BODY_SPHERE ? p->sx : p->sy,
                p->theta, p->vx, p->vy, p->vtheta};
    return orc_body_distance(&b, x, y);
}
double orc_cubic_pulse(double x) { return cubic_pulse(x); }
double orc_cerp1(double a, double b, double c, double d, double x) { return cerp1(a, b, c, d, x); }

double orc_max_timestep(orc_solver* s) { return max_timestep(s); }
void orc_to_image(orc_solver* s, unsigned char* rgba, int render_heat) { to_image(s, rgba, render_heat); }
