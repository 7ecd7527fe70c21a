/*
 * ifl.h -- C ABI of the MI355X-native incremental-fluids solver step.
 *
 * The reference (tunabrain/incremental-fluids) is a set of single-file C++
 * programs whose only interface is the FluidSolver / FluidQuantity /
 * SolidBody class API. This header is the flat C boundary under a C++ shim
 * (ifl_fluid.hpp) that keeps those class names, constructors and update().
 * Every entry point below cites the reference member it stands in for
 * (N:line = N-<solver>/Fluid.cpp:line in the reference tree).
 *
 * Conventions: plain pointers and sizes, int status return (IFL_OK == 0),
 * nothing throws across the boundary, device memory never escapes.
 * Field layout is the reference's own row-major x + y*w per quantity.
 */
#ifndef IFL_H
#define IFL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFL_ABI_VERSION 1

/* ---- status codes (the reference has none: new throws, PCG printf's) ---- */
enum {
    IFL_OK = 0,
    IFL_ERR_ARG = 1,          /* bad argument (size, id, level, pointer) */
    IFL_ERR_ALLOC = 2,        /* device allocation failed */
    IFL_ERR_HIP = 3,          /* a HIP runtime call failed */
    IFL_ERR_UNSUPPORTED = 4,  /* operation not available at this level */
    IFL_ERR_NO_DEVICE = 5,    /* no usable gfx950 device */
    IFL_ERR_DEVICE = 6        /* a kernel reported a bounded-wait timeout */
};

/* ---- cell types, 7:104-107 and 8:115-119 ---- */
enum { IFL_CELL_FLUID = 0, IFL_CELL_SOLID = 1, IFL_CELL_EMPTY = 2 };

/* ---- solid bodies, SolidBox 7:169-218 / SolidSphere 7:220-257 ----
 * IFL_BODY_HOST: any other SolidBody subclass. Its three virtuals
 * (distance, closestSurfacePoint, distanceNormal, 7:143-146) are evaluated
 * on the host through the callbacks of ifl_set_host_body_ops; its rigid
 * velocity (velocityX/Y 7:148-154, not virtual) comes from pos/vel below. */
enum { IFL_BODY_BOX = 0, IFL_BODY_SPHERE = 1, IFL_BODY_HOST = 2 };

typedef struct ifl_body {
    int32_t type;          /* IFL_BODY_BOX, IFL_BODY_SPHERE or IFL_BODY_HOST */
    int32_t reserved;
    double pos_x, pos_y;   /* SolidBody::_posX/_posY, 7:111-112 */
    double scale_x, scale_y;
    double theta;
    double vel_x, vel_y, vel_theta;
    /* rotate(x, y, phi) (7:54-58) evaluates libm cos/sin; the caller passes
     * the host values so device results match the host libm bit for bit. */
    double cos_neg, sin_neg;   /* cos(-theta), sin(-theta) */
    double cos_pos, sin_pos;   /* cos(theta),  sin(theta)  */
} ifl_body;

/* ---- configuration ---- */
/* IFL_PREC_FP64: the drop-in, every operation in double as the reference.
 * IFL_PREC_MIXED: advection interpolated in single precision (trace
 * positions in double); project() keeps the conjugate-gradient iteration in
 * double (residual, iterate, directions, A*s, dot products, alpha/beta and
 * the test max|r| < tol) and applies the MIC(0) preconditioner in single
 * precision from a single-precision copy of the double-built factor;
 * pressure solves apply it as P M^-1 P, P taking off the mean over FLUID
 * cells (the pressure matrix is singular with closed walls, and the
 * constant mode is left out of p).
 * IFL_PREC_REFINE: advection as IFL_PREC_MIXED; project() as defect
 * correction in double around a conjugate-gradient iteration entirely in
 * single precision (factor, matrix and vectors in float; dot products,
 * alpha and beta in double), the double residual updated by r -= A e after
 * each inner solve and ended by the reference's test. It converges while
 * the system's condition number times single-precision epsilon stays well
 * below 1 (grids to ~512^2 for the pressure systems of solver 7). */
enum { IFL_PREC_FP64 = 0, IFL_PREC_MIXED = 1, IFL_PREC_REFINE = 2 };
enum { IFL_DOT_TREE = 0, IFL_DOT_SEQUENTIAL = 1 };

typedef struct ifl_config {
    int32_t level;        /* 3..8: which reference solver's update() to run */
    int32_t w, h;         /* grid size in cells */
    int32_t precision;    /* IFL_PREC_FP64 (drop-in) or IFL_PREC_MIXED */
    int32_t dot_order;    /* IFL_DOT_TREE (fast) or IFL_DOT_SEQUENTIAL
                             (reference summation order, for parity runs) */
    int32_t pcg_limit;    /* project(limit): 600 at level 3 (3:437), else 2000 (7:1016) */
    int32_t verbose;      /* 1: print the reference's stdout lines per solve */
    int32_t device;       /* HIP device ordinal */
    double density;       /* level 3-5 fluid density; 6-8 air density (7:605) */
    double density_soot;  /* level 6-8 (7:606) */
    double diffusion;     /* level 6-8 (7:607) */
    double tau, sigma;    /* MIC(0) constants 0.97 / 0.25 (7:740-741) */
    double tol;           /* PCG absolute tolerance on the max residual, 1e-5 (7:859) */
    double t_amb;         /* 294 (7:950) */
    double gravity;       /* 9.81 (7:951) */
    double flip_alpha;    /* 0.001 (8:1296) */
    int32_t flip_g2p;     /* level 8, the FLIP update's gridToParticles (8:1401):
                             0 bilinear lerp, as the reference (8:904-911);
                             1 the Catmull-Rom cerp of solvers 3-7 (7:404-422),
                             a variant with no counterpart in the reference (the
                             constructor's gridToParticles(1.0), 8:1314, stays
                             bilinear) */
    int32_t reserved[9];
} ifl_config;

/* Fills cfg with the reference's literal defaults for `level`. */
int ifl_config_default(ifl_config* cfg, int level, int w, int h);

/* ---- statistics returned in place of the reference's printf lines ---- */
enum {
    IFL_SOLVE_CONVERGED = 0,      /* "Exiting solver after %d iterations ..." 7:874 */
    IFL_SOLVE_INITIAL_SMALL = 1,  /* "Initial guess sufficiently small" 7:860 */
    IFL_SOLVE_BUDGET = 2,         /* "Exceeded budget of %d iterations ..." 7:885 */
    IFL_SOLVE_NOT_RUN = 3
};

typedef struct ifl_solve_stats {
    int32_t status;       /* IFL_SOLVE_* */
    int32_t iterations;   /* the `iter` the reference prints on convergence,
                             or the limit on budget exhaustion */
    double max_error;     /* final max |r| over the reference's norm domain */
} ifl_solve_stats;

typedef struct ifl_step_stats {
    ifl_solve_stats heat;      /* levels 6-8 (7:1001-1005) */
    ifl_solve_stats pressure;  /* all levels */
    int32_t particle_count;    /* level 8, "Particle count: %d" 8:926 */
    int32_t border_solid;      /* 1 if a non-fluid cell touched the domain border */
    int32_t nonfinite;         /* 1 if d, t, u or v holds a NaN or an infinity
                                  after the step (the reference runs on silently) */
    int32_t reserved;
} ifl_step_stats;

/* ---- lifetime: FluidSolver ctor 3:401 / 4:836 / 5:894 / 6:921 / 7:945 / 8:1289,
 *      dtor 7:976 ---- */
typedef struct ifl_solver ifl_solver;
int ifl_create(const ifl_config* cfg, ifl_solver** out);
int ifl_destroy(ifl_solver* s);
const char* ifl_last_error(const ifl_solver* s);   /* s may be NULL */
const char* ifl_version(void);

/* The borrowed `const vector<const SolidBody*>&` (7:622) that the reference
 * re-reads on every update(). n <= 255 (uint8 _body, 7:268). Level 8 needs
 * the bodies before the first update (initParticles 8:735). */
int ifl_set_bodies(ifl_solver* s, const ifl_body* bodies, int n);

/* Host evaluation of the bodies' virtuals, for body lists that hold an
 * IFL_BODY_HOST body (levels 4-7). While such a list is set, every body of it
 * (built-in ones too) is evaluated through these callbacks, in the
 * reference's order: fillSolidFields (7:475-522) takes phi at the cell
 * corners (min over the bodies in list order), the closest body of each
 * cell (strict <, the lowest index wins ties) and its normal from the host;
 * advect's backProject (7:424-435) hands the points that land in a solid
 * cell to `closest` in batches. Each callback gets the body's index in the
 * list and n points; the library does not keep the pointers. Pass NULL to
 * clear. Replaces the virtual calls bodies[i]->distance(x, y) (7:483, 493),
 * ->distanceNormal(nx, ny, x, y) (7:516) and ->closestSurfacePoint(x, y)
 * (7:431). */
typedef struct ifl_host_body_ops {
    void* user;
    void (*distance)(void* user, int body, const double* x, const double* y, double* out, size_t n);
    void (*closest)(void* user, int body, double* x, double* y, size_t n);
    void (*normal)(void* user, int body, const double* x, const double* y, double* nx, double* ny, size_t n);
} ifl_host_body_ops;
int ifl_set_host_body_ops(ifl_solver* s, const ifl_host_body_ops* ops);

/* FluidSolver::addInflow(x, y, w, h, d, [t,] u, v), 3:449 / 7:1036.
 * vals = {d, t, u, v}; t is ignored below level 6. */
int ifl_add_inflow(ifl_solver* s, double x, double y, double w, double h,
                   const double vals[4]);

/* FluidSolver::update(timestep), 3:433 / 4:869 / 5:927 / 6:966 / 7:995 /
 * 8:1350. Blocks until the step is done; stats may be NULL. */
int ifl_update(ifl_solver* s, double dt, ifl_step_stats* stats);

/* ---- state access (the reference's fields are private; the harness and
 *      teacher-forced tests read and write them through these) ---- */
enum { IFL_Q_D = 0, IFL_Q_T = 1, IFL_Q_U = 2, IFL_Q_V = 3 };
enum {
    IFL_QF_SRC = 0, IFL_QF_DST = 1,     /* _src / _dst (_old at level 8) */
    IFL_QF_PHI = 2, IFL_QF_VOLUME = 3,
    IFL_QF_NORMALX = 4, IFL_QF_NORMALY = 5,
    IFL_QF_CELL = 6, IFL_QF_BODY = 7, IFL_QF_MASK = 8
};
#define IFL_FIELD_Q(q, f) ((q) * 16 + (f))
enum {
    IFL_FIELD_R = 64, IFL_FIELD_P = 65, IFL_FIELD_Z = 66, IFL_FIELD_S = 67,
    IFL_FIELD_PRECON = 68, IFL_FIELD_ADIAG = 69, IFL_FIELD_APLUSX = 70,
    IFL_FIELD_APLUSY = 71, IFL_FIELD_UDENSITY = 72, IFL_FIELD_VDENSITY = 73,
    /* level 8 (ParticleQuantities 8:692-940) */
    IFL_FIELD_POSX = 80, IFL_FIELD_POSY = 81,
    IFL_FIELD_PROP_D = 82, IFL_FIELD_PROP_T = 83, IFL_FIELD_PROP_U = 84,
    IFL_FIELD_PROP_V = 85, IFL_FIELD_WEIGHT = 86, IFL_FIELD_COUNTS = 87
};

/* Element count and size (8, 4 or 1 bytes) of a field; 0 count if absent.
 * IFL_FIELD_POSX..IFL_FIELD_PROP_V are doubles in IFL_PREC_FP64 and floats
 * (4 bytes) in the reduced-precision forms; IFL_FIELD_COUNTS is int32. */
int ifl_field_info(const ifl_solver* s, int field, size_t* count, int* elem_bytes);
int ifl_upload(ifl_solver* s, int field, const void* host, size_t bytes);
int ifl_download(ifl_solver* s, int field, void* host, size_t bytes);
/* Level 8: the live particle count (_particleCount 8:701). Setting it hands
 * the particle state to the caller: ifl_update then skips initialisation. */
int ifl_get_particle_count(const ifl_solver* s, int* count);
int ifl_set_particle_count(ifl_solver* s, int count);

/* Level 8: the rest of the FLIP constructor, ParticleQuantities'
 * initParticles (8:735-751, 4 jittered particles per cell from frand, those
 * inside a body dropped) and FluidSolver's gridToParticles(1.0) (8:1314).
 * Call once after ifl_set_bodies; ifl_update calls it if it has not run. */
int ifl_init_particles(ifl_solver* s);

/* ---- batches: independent solvers stepped together ----
 * The solvers (one level, 3..8, one device, FP64 form, tree-ordered dots)
 * keep their own state; ifl_batch_update runs FluidSolver::update(dt) on all
 * of them (the reference steps each in turn, 7:1110-1126), with each PCG
 * iteration of their project() calls one launch per phase for the whole
 * batch; at levels 3..7 with device-evaluated bodies every other stage of
 * update() is one launch for the whole batch too. Level 8 solvers draw
 * frand() in batch order within each step, as the same solvers stepped one
 * after the other would. stats receives n entries. No counterpart in the
 * reference: a caller with many small simulations would otherwise step them
 * one by one. */
typedef struct ifl_batch ifl_batch;
int ifl_batch_create(ifl_solver* const* solvers, int n, ifl_batch** out);
int ifl_batch_update(ifl_batch* batch, double dt, ifl_step_stats* stats);
int ifl_batch_destroy(ifl_batch* batch);
/* FluidQuantity::addInflow (3:167-184, 7:456-473) of many solvers of a batch in
 * one call: entry k applies rect[4k..4k+3] = (x, y, w, h) with the values
 * vals[4k..4k+3] (d, t, u, v) to the batch's solver inst[k] (an index into
 * the list given to ifl_batch_create). A solver's entries apply in their
 * order. The same as ifl_add_inflow on each entry in turn, with one launch
 * per rank (every solver's first entry, then every second ...) and one
 * wait. No counterpart in the reference, which calls addInflow per solver
 * (7:1110-1126). */
int ifl_batch_add_inflow(ifl_batch* batch, int count, const int* inst, const double* rect, const double* vals);
/* Diagnostics, no counterpart in the reference: with profiling on, every
 * fused PCG launch of the batch (IFL_KERNEL_MATVEC, _SWEEP_FWD, _SWEEP_BWD)
 * is timed with events on the batch's stream. ifl_batch_profile(on) also
 * clears the totals; ifl_batch_profile_read returns, per kernel, the timed
 * launches, their total milliseconds and the cells of the instances that
 * worked in them, summed over the launches. */
int ifl_batch_profile(ifl_batch* batch, int on);
int ifl_batch_profile_read(const ifl_batch* batch, int kernel, int64_t* launches, double* ms,
                           double* cell_iterations);

/* FluidSolver::toImage(rgba[, renderHeat]) 3:455 / 4:903 / 5:961 / 6:1021 /
 * 7:1047 / 8:1426, rendered on the device: bytes = w*h*4, or w*h*8 with
 * render_heat at levels 6-8 (the heat panel left of the density panel). */
int ifl_to_image(ifl_solver* s, uint8_t* rgba, size_t bytes, int render_heat);

/* No counterpart in the reference: ifl_to_image for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It renders the same bytes as
 * ifl_to_image, staged in row windows through a 4 MiB device scratch array
 * that every device the library runs on holds for the life of the process;
 * it calls no hipMalloc or hipFree. One image row (w*4 bytes, w*8 with heat)
 * must fit in the scratch. All calls on one device share that scratch:
 * calls that overlap on one device (solvers on other threads) race. */
int ifl_to_image_static(ifl_solver* s, uint8_t* rgba, size_t bytes, int render_heat);

/* The reference's frame file (lodepng_encode32_file, 7:1122): an RGBA, 8-bit
 * PNG of w x h pixels (w doubled for a heat frame), stored deflate blocks. */
int ifl_write_png(const char* path, const uint8_t* rgba, int w, int h);

/* FluidSolver::maxTimestep (1:310-328), solver 1's CFL bound, at any level:
 * 2h over the largest speed at the cell centres (each velocity component
 * lerp'd from its faces), at most 1. Reads the current u, v. */
int ifl_max_timestep(ifl_solver* s, double* dt);

/* frand()'s seed (8:38-46): one per process, 0xBA5EBA11 at load, advanced by
 * every level-8 solver in creation and step order as in the reference. */
uint32_t ifl_frand_state(void);
void ifl_frand_set_state(uint32_t state);

/* ---- single stages of update(), for teacher-forced parity tests ----
 * iarg / darg meaning per stage is given on each line. `result` receives the
 * scalar of DOT / INF_NORM; `stats` receives PROJECT's solve statistics. */
enum {
    IFL_STAGE_FILL_SOLID = 1,       /* iarg=q      FluidQuantity::fillSolidFields 7:475 */
    IFL_STAGE_EXTRAPOLATE = 2,      /* iarg=q      FluidQuantity::extrapolate 7:560 / 8:611 */
    IFL_STAGE_ADVECT = 3,           /* iarg=q darg=dt  FluidQuantity::advect 7:437 */
    IFL_STAGE_FLIP = 4,             /* iarg=q      FluidQuantity::flip 7:354 */
    IFL_STAGE_BUILD_RHS = 5,        /*             buildRhs 7:624 */
    IFL_STAGE_BUILD_PRESSURE = 6,   /* darg=dt     buildPressureMatrix 7:680 */
    IFL_STAGE_BUILD_HEAT = 7,       /* darg=dt     buildHeatDiffusionMatrix 7:709 */
    IFL_STAGE_BUILD_PRECON = 8,     /*             buildPreconditioner 7:739 */
    IFL_STAGE_APPLY_PRECON = 9,     /*             applyPreconditioner(_z, _r) 7:770 */
    IFL_STAGE_MATVEC = 10,          /*             matrixVectorProduct(_z, _s) 7:816 */
    IFL_STAGE_DOT = 11,             /* iarg=0: dot(_z,_r); 1: dot(_z,_s)  7:806 */
    IFL_STAGE_INF_NORM = 12,        /*             infinityNorm(_r) 7:843 */
    IFL_STAGE_PROJECT = 13,         /* iarg=limit  project 7:853 */
    IFL_STAGE_APPLY_PRESSURE = 14,  /* darg=dt     applyPressure 7:891 */
    IFL_STAGE_ADD_BUOYANCY = 15,    /* darg=dt     addBuoyancy 7:908 */
    IFL_STAGE_SET_BOUNDARY = 16,    /*             setBoundaryCondition 7:921 */
    IFL_STAGE_COMPUTE_DENSITIES = 17, /*           computeDensities 7:658 */
    IFL_STAGE_COPY_T_TO_R = 18,     /*             memcpy(_r, _t->src()) 7:1001 */
    IFL_STAGE_COPY_P_TO_T = 19,     /*             memcpy(_t->src(), _p) 7:1005 */
    /* level 8 */
    IFL_STAGE_FROM_PARTICLES = 32,  /* iarg=q      fromParticles 8:663 */
    IFL_STAGE_G2P = 33,             /* darg=alpha  gridToParticles 8:904 */
    IFL_STAGE_P2G = 34,             /*             particlesToGrid 8:916 */
    IFL_STAGE_PARTICLE_ADVECT = 35, /* darg=dt     ParticleQuantities::advect 8:931 */
    IFL_STAGE_COPY_OLD = 36,        /* iarg=q      copy 8:374 */
    IFL_STAGE_DIFF = 37,            /* iarg=q darg=alpha  diff 8:379 */
    IFL_STAGE_UNDIFF = 38,          /* iarg=q darg=alpha  undiff 8:385 */
    IFL_STAGE_COUNT_PARTICLES = 39, /*             countParticles 8:754 */
    IFL_STAGE_PRUNE = 40,           /*             pruneParticles 8:766 */
    IFL_STAGE_SEED = 41             /*             seedParticles 8:789 */
};
int ifl_stage(ifl_solver* s, int stage, int iarg, double darg,
              ifl_solve_stats* stats, double* result);

/* ---- timing of the device work, HIP events on the solver's stream ---- */
enum {
    IFL_KERNEL_PCG = 0,        /* one project() solve, all of its kernels timed as one */
    IFL_KERNEL_PRECON = 1,     /* buildPreconditioner wavefront sweep, one launch per solve */
    IFL_KERNEL_ADVECT = 2,     /* advection of the quantities of one update() */
    IFL_KERNEL_SWEEP_FWD = 3,  /* r -= alpha*q, max|r| and the forward MIC(0) sweep, one launch per PCG iteration */
    IFL_KERNEL_SWEEP_BWD = 4,  /* backward MIC(0) sweep and the z.r partials, one launch per PCG iteration */
    IFL_KERNEL_MATVEC = 5,     /* q = A s and the q.s partials, one launch per PCG iteration */
    IFL_KERNEL_P2G = 6,        /* level 8: fromParticles gather of one grid (d with t, u, v) */
    IFL_KERNEL_G2P = 7,        /* level 8: gridToParticles of the four properties */
    IFL_KERNEL_PARTICLE_ADVECT = 8, /* level 8: ParticleQuantities::advect */
    IFL_KERNEL_COUNT = 9
};
int ifl_profile_enable(ifl_solver* s, int on);
/* launches, summed milliseconds, summed PCG iterations (project() only) and
 * summed cells processed (cells x launches, or x iterations for project()).
 * Only launches that did work are counted: a speculatively queued PCG
 * iteration after convergence is not. */
int ifl_profile_read(const ifl_solver* s, int kernel, int64_t* launches,
                     double* total_ms, int64_t* iterations, double* cell_iterations);
int ifl_profile_reset(ifl_solver* s);
int ifl_synchronize(ifl_solver* s);

/* ---- diagnostics: one MIC(0) sweep (kind 0 forward, 1 backward, 2 factor
 *      build) on the current state, time-stamped per 64-row strip and
 *      8-step block, in 100 MHz ticks: {block begins, its incoming edges
 *      are ready, first half computed, next block's operands ready, second
 *      half computed, results written}. count >= nstrips * blocks_per_strip * 6,
 *      laid out [strip][block][6]. */
int ifl_debug_trace_sweep(ifl_solver* s, int kind, int64_t* host_out, size_t count, int* blocks_per_strip);

#ifdef __cplusplus
}
#endif
#endif /* IFL_H */
