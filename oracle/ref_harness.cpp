/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY (the oracle). Never linked into
 * or called by the product library.
 *
 * Compiles one of the reference solvers (N-<name>/Fluid.cpp, read in place
 * from the reference tree, never copied) into a shared library with a flat C
 * API, so tests can run the reference's own update() and its individual
 * stages and read its private arrays. Built by oracle/Makefile into
 * oracle/_ref/libref<N>.so with the reference Makefile's flags (-O2, no -m
 * arch flags, so no FMA contraction on x86-64).
 *
 * What the wrapper changes, and why:
 *   - `class`/`private`/`protected` become `struct`/`public` so the fields
 *     are readable; the reference's code itself is unchanged.
 *   - `main` is renamed so the library has none; it is never called.
 *   - `printf` inside the reference is captured into a per-process log, so
 *     iteration counts can be compared line for line (7:860/874/885).
 *   - `new[]` returns zeroed memory: the reference never initialises _dst,
 *     _z, _s, _mask (e.g. 7:321); zero fill makes runs reproducible.
 * Level-8 solvers are never destroyed: ~ParticleQuantities double-deletes
 * (8:887-888) and crashes.
 */
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stack>
#include <string>
#include <vector>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "lodepng/lodepng.h"
#include "../include/ifl.h"

void* operator new[](std::size_t n) {
    void* p = std::calloc(n ? n : 1, 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

static std::string g_log;
extern "C" int ifl_ref_printf(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_log += buf;
    return n;
}

#define main ifl_reference_main_unused
#define printf ifl_ref_printf
#define class struct
#define private public
#define protected public
#include REF_SRC
#undef class
#undef private
#undef protected
#undef printf
#undef main

#if REF_LEVEL == 3
struct SolidBody; /* solver 3 has no solid bodies */
#endif

namespace {

struct RefBodyPod {
    int type;
    double x, y, sx, sy, theta, vx, vy, vtheta;
};

struct Handle {
    std::vector<SolidBody*> bodies;
    std::vector<const SolidBody*> cbodies;
    FluidSolver* solver;
    int w, h;
};

FluidQuantity* quantity(Handle* h, int q) {
    FluidSolver* s = h->solver;
    switch (q) {
    case IFL_Q_D: return s->_d;
#if REF_LEVEL >= 6
    case IFL_Q_T: return s->_t;
#endif
    case IFL_Q_U: return s->_u;
    case IFL_Q_V: return s->_v;
    }
    return nullptr;
}

}  // namespace

extern "C" {

int ref_level(void) { return REF_LEVEL; }

/* params: level 3-5: {density}; level 6-8: {rhoAir, rhoSoot, diffusion}. */
void* ref_create(int w, int h, const double* params, const RefBodyPod* bodies, int nbodies) {
    Handle* hd = new Handle();
    hd->w = w;
    hd->h = h;
#if REF_LEVEL >= 4
    for (int i = 0; i < nbodies; i++) {
        const RefBodyPod& b = bodies[i];
        SolidBody* sb;
        if (b.type == IFL_BODY_SPHERE)
            sb = new SolidSphere(b.x, b.y, b.sx, b.theta, b.vx, b.vy, b.vtheta);
        else
            sb = new SolidBox(b.x, b.y, b.sx, b.sy, b.theta, b.vx, b.vy, b.vtheta);
        hd->bodies.push_back(sb);
        hd->cbodies.push_back(sb);
    }
#else
    (void)bodies;
    (void)nbodies;
#endif
#if REF_LEVEL == 3
    hd->solver = new FluidSolver(w, h, params[0]);
#elif REF_LEVEL <= 5
    hd->solver = new FluidSolver(w, h, params[0], hd->cbodies);
#else
    hd->solver = new FluidSolver(w, h, params[0], params[1], params[2], hd->cbodies);
#endif
    return hd;
}

void ref_destroy(void* p) {
    Handle* hd = (Handle*)p;
#if REF_LEVEL != 8
    delete hd->solver;
#if REF_LEVEL >= 4
    for (size_t i = 0; i < hd->bodies.size(); i++) delete hd->bodies[i];
#endif
    delete hd;
#else
    (void)hd; /* see header comment: level 8 destructor is unusable */
#endif
}

void ref_bodies_update(void* p, double dt) {
    Handle* hd = (Handle*)p;
#if REF_LEVEL >= 4
    for (size_t i = 0; i < hd->bodies.size(); i++) hd->bodies[i]->update(dt);
#else
    (void)hd; (void)dt;
#endif
}

void ref_add_inflow(void* p, double x, double y, double w, double h, double d, double t, double u, double v) {
    Handle* hd = (Handle*)p;
#if REF_LEVEL <= 5
    (void)t;
    hd->solver->addInflow(x, y, w, h, d, u, v);
#else
    hd->solver->addInflow(x, y, w, h, d, t, u, v);
#endif
}

void ref_update(void* p, double dt) { ((Handle*)p)->solver->update(dt); }

/* toImage 3:455 / 4:903 / 5:961 / 6:1021 / 7:1047 / 8:1426; rgba holds
 * w*h*4 bytes, or w*2*h*4 with render_heat (levels 6-8). */
void ref_to_image(void* p, unsigned char* rgba, int render_heat) {
#if REF_LEVEL <= 5
    (void)render_heat;
    ((Handle*)p)->solver->toImage(rgba);
#else
    ((Handle*)p)->solver->toImage(rgba, render_heat != 0);
#endif
}

const char* ref_log(void) { return g_log.c_str(); }
void ref_log_clear(void) { g_log.clear(); }

/* Pointer to an internal array plus its element count and size. */
void* ref_field(void* p, int field, long* count, int* elem) {
    Handle* hd = (Handle*)p;
    FluidSolver* s = hd->solver;
    int w = hd->w, h = hd->h;
    *count = 0;
    *elem = 8;
    if (field < 64) {
        FluidQuantity* q = quantity(hd, field / 16);
        if (!q) return nullptr;
        int f = field % 16;
        long n = (long)q->_w * q->_h;
        *count = n;
        switch (f) {
        case IFL_QF_SRC: return q->_src;
#if REF_LEVEL == 8
        case IFL_QF_DST: return q->_old;
#else
        case IFL_QF_DST: return q->_dst;
#endif
#if REF_LEVEL >= 5
        case IFL_QF_PHI: *count = (long)(q->_w + 1) * (q->_h + 1); return q->_phi;
        case IFL_QF_VOLUME: return q->_volume;
#endif
#if REF_LEVEL >= 4
        case IFL_QF_NORMALX: return q->_normalX;
        case IFL_QF_NORMALY: return q->_normalY;
        case IFL_QF_CELL: *elem = 1; return q->_cell;
        case IFL_QF_BODY: *elem = 1; return q->_body;
        case IFL_QF_MASK: *elem = 1; return q->_mask;
#endif
        }
        *count = 0;
        return nullptr;
    }
    long n = (long)w * h;
    switch (field) {
    case IFL_FIELD_R: *count = n; return s->_r;
    case IFL_FIELD_P: *count = n; return s->_p;
    case IFL_FIELD_Z: *count = n; return s->_z;
    case IFL_FIELD_S: *count = n; return s->_s;
    case IFL_FIELD_PRECON: *count = n; return s->_precon;
    case IFL_FIELD_ADIAG: *count = n; return s->_aDiag;
    case IFL_FIELD_APLUSX: *count = n; return s->_aPlusX;
    case IFL_FIELD_APLUSY: *count = n; return s->_aPlusY;
#if REF_LEVEL >= 7
    case IFL_FIELD_UDENSITY: *count = (long)(w + 1) * h; return s->_uDensity;
    case IFL_FIELD_VDENSITY: *count = (long)w * (h + 1); return s->_vDensity;
#endif
#if REF_LEVEL == 8
    case IFL_FIELD_POSX: *count = s->_qs->_maxParticles; return s->_qs->_posX;
    case IFL_FIELD_POSY: *count = s->_qs->_maxParticles; return s->_qs->_posY;
    case IFL_FIELD_PROP_D: case IFL_FIELD_PROP_T: case IFL_FIELD_PROP_U: case IFL_FIELD_PROP_V:
        *count = s->_qs->_maxParticles;
        return s->_qs->_properties[field - IFL_FIELD_PROP_D];
    case IFL_FIELD_WEIGHT: *count = (long)(w + 1) * (h + 1); return s->_qs->_weight;
    case IFL_FIELD_COUNTS: *count = n; *elem = 4; return s->_qs->_counts;
#endif
    }
    return nullptr;
}

int ref_particle_count(void* p, int set, int value) {
#if REF_LEVEL == 8
    Handle* hd = (Handle*)p;
    if (set) hd->solver->_qs->_particleCount = value;
    return hd->solver->_qs->_particleCount;
#else
    (void)p; (void)set; (void)value;
    return -1;
#endif
}

/* Runs one stage of update(); same ids as the product's ifl_stage().
 * Returns 0 on success, 4 if the stage does not exist at this level. */
int ref_stage(void* p, int stage, int iarg, double darg, double* result) {
    Handle* hd = (Handle*)p;
    FluidSolver* s = hd->solver;
    FluidQuantity* q = (iarg >= 0 && iarg < 4) ? quantity(hd, iarg) : nullptr;
    switch (stage) {
#if REF_LEVEL >= 4
    case IFL_STAGE_FILL_SOLID: if (!q) return 1; q->fillSolidFields(hd->cbodies); return 0;
    case IFL_STAGE_EXTRAPOLATE: if (!q) return 1; q->extrapolate(); return 0;
    case IFL_STAGE_SET_BOUNDARY: s->setBoundaryCondition(); return 0;
#endif
#if REF_LEVEL == 3
    case IFL_STAGE_ADVECT: if (!q) return 1; q->advect(darg, *s->_u, *s->_v); return 0;
#elif REF_LEVEL < 8
    case IFL_STAGE_ADVECT: if (!q) return 1; q->advect(darg, *s->_u, *s->_v, hd->cbodies); return 0;
#endif
#if REF_LEVEL < 8
    case IFL_STAGE_FLIP: if (!q) return 1; q->flip(); return 0;
#endif
    case IFL_STAGE_BUILD_RHS: s->buildRhs(); return 0;
    case IFL_STAGE_BUILD_PRESSURE: s->buildPressureMatrix(darg); return 0;
#if REF_LEVEL >= 6
    case IFL_STAGE_BUILD_HEAT: s->buildHeatDiffusionMatrix(darg); return 0;
    case IFL_STAGE_ADD_BUOYANCY: s->addBuoyancy(darg); return 0;
    case IFL_STAGE_COPY_T_TO_R: memcpy(s->_r, s->_t->src(), sizeof(double) * hd->w * hd->h); return 0;
    case IFL_STAGE_COPY_P_TO_T: memcpy(s->_t->src(), s->_p, sizeof(double) * hd->w * hd->h); return 0;
#endif
#if REF_LEVEL >= 7
    case IFL_STAGE_COMPUTE_DENSITIES: s->computeDensities(); return 0;
#endif
    case IFL_STAGE_BUILD_PRECON: s->buildPreconditioner(); return 0;
    case IFL_STAGE_APPLY_PRECON: s->applyPreconditioner(s->_z, s->_r); return 0;
    case IFL_STAGE_MATVEC: s->matrixVectorProduct(s->_z, s->_s); return 0;
    case IFL_STAGE_DOT:
        *result = iarg == 0 ? s->dotProduct(s->_z, s->_r) : s->dotProduct(s->_z, s->_s);
        return 0;
    case IFL_STAGE_INF_NORM: *result = s->infinityNorm(s->_r); return 0;
    case IFL_STAGE_PROJECT: s->project(iarg); return 0;
    case IFL_STAGE_APPLY_PRESSURE: s->applyPressure(darg); return 0;
#if REF_LEVEL == 8
    case IFL_STAGE_FROM_PARTICLES: {
        if (!q) return 1;
        ParticleQuantities* qs = s->_qs;
        q->fromParticles(qs->_weight, qs->_particleCount, qs->_posX, qs->_posY, qs->_properties[iarg]);
        return 0;
    }
    case IFL_STAGE_G2P: s->_qs->gridToParticles(darg); return 0;
    case IFL_STAGE_P2G: s->_qs->particlesToGrid(); return 0;
    case IFL_STAGE_PARTICLE_ADVECT: s->_qs->advect(darg, *s->_u, *s->_v); return 0;
    case IFL_STAGE_COPY_OLD: if (!q) return 1; q->copy(); return 0;
    case IFL_STAGE_DIFF: if (!q) return 1; q->diff(darg); return 0;
    case IFL_STAGE_UNDIFF: if (!q) return 1; q->undiff(darg); return 0;
    case IFL_STAGE_COUNT_PARTICLES: s->_qs->countParticles(); return 0;
    case IFL_STAGE_PRUNE: s->_qs->pruneParticles(); return 0;
    case IFL_STAGE_SEED: s->_qs->seedParticles(); return 0;
#endif
    }
    return 4;
}

/* Level 8 frand() (8:38-46) state is a function-local static; expose a draw
 * so tests can check the LCG closed form against it. */
double ref_frand(void) {
#if REF_LEVEL == 8
    return frand();
#else
    return -1.0;
#endif
}

}  // extern "C"
