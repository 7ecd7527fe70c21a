// Host-side state of one solver instance and the kernel launchers shared
// between the translation units of libifl.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <atomic>
#include <vector>

#include "ifl_common.hpp"

namespace ifl {

constexpr int kMaxBodies = 255;  // _body is uint8_t (7:268)

// Device-side scalars of one project() call (7:853-886). Written and read
// only by kernels; the host reads `done`, `status`, `iter`, `max_err` once
// per batch of iterations.
struct PcgState {
    double sigma;      // z.r of the current iteration
    double alpha;
    double beta;
    double max_err;
    double dot_tmp;    // scratch for the sequential-order dot
    int iter;          // iterations completed
    int done;          // 1 once converged / early exit / budget
    int status;        // IFL_SOLVE_*
    int printed_iter;  // the `iter` the reference prints
    int timeout;       // a bounded wait gave up (kernel bug, never expected)
    int pad[3];
    double zmean;      // mixed form, pressure solves: the mean over FLUID cells taken off z
    double rmean;      // mixed form, pressure solves: the mean over FLUID cells taken off r
};

struct Profile {
    long long launches = 0;
    double ms = 0;
    long long iterations = 0;
    double cell_iterations = 0;
};

struct Solver {
    ifl_config cfg;
    int level, W, H;
    double hx;
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;

    QtyView q[4];  // d, t, u, v (t only at level >= 6)
    double* udens = nullptr;  // (w+1)*h, row-major
    double* vdens = nullptr;  // padded to (w+1)*(h+1): the reference indexes it with the u stride (Q2)
    size_t vdens_count = 0;

    // PCG arrays, strip-skewed (Skew), each sk.nstrips * sk.strip_sz elements
    Skew sk;
    size_t sk_count = 0;
    double *r = nullptr, *p = nullptr, *z = nullptr, *s = nullptr;
    // levels 6-8, FP64 with tree-ordered dots: the PCG iterations after the
    // first form s = z + beta s inside the matrix-vector product, writing the
    // new s here and in s by turns (null: not used)
    double* s2 = nullptr;
    double *precon = nullptr, *adiag = nullptr, *aplusx = nullptr, *aplusy = nullptr;
    double *cx = nullptr, *cy = nullptr;  // aPlusX*precon, aPlusY*precon (the sweeps' products)
    double *yv = nullptr;                 // forward-sweep output
    // mixed form: the factor and the forward-sweep output in single precision,
    // and the inner conjugate-gradient iteration's vectors (correction e,
    // residual, directions, preconditioned residual, A s) and matrix
    float *cx32 = nullptr, *cy32 = nullptr, *pc32 = nullptr, *y32 = nullptr;
    float *e32 = nullptr, *r32 = nullptr, *s32 = nullptr, *z32 = nullptr, *q32 = nullptr;
    float *ad32 = nullptr, *apx32 = nullptr, *apy32 = nullptr;
    PcgState* st_outer = nullptr;  // the outer (double) residual's norm
    int solve_kind = 2;            // SOLVE_HEAT or SOLVE_PRESSURE: the solve project() runs
    long nfluid = 0;               // FLUID cells of d (the mixed form's deflation)
    double *qv = nullptr;                 // A*s
    uint8_t* dflag = nullptr;             // 1 where the d-cell is FLUID (and inside the grid)
    uint64_t* fmask = nullptr;            // dflag as one 64-lane mask per strip column (nstrips*cols)
    double* edge = nullptr;               // inter-strip hand-off, 2 buffers of nstrips*W
    double* partial = nullptr;            // per-block partial reductions
    int npartial = 0;
    PcgState* st = nullptr;
    PcgState* st_host = nullptr;          // pinned mirror
    double* stage_buf = nullptr;          // row-major staging for layout conversion / upload
    size_t stage_count = 0;

    // extrapolation frontier lists
    int* front = nullptr;  // 2 lists of W*H ints + counters
    int* front_cnt = nullptr;

    // bodies
    int nbodies = 0;
    Body bodies_h[kMaxBodies];
    Body* bodies_d = nullptr;
    // host-evaluated bodies (ifl_set_host_body_ops): while the body list holds
    // an IFL_BODY_HOST body, fillSolidFields and backProject call back to the
    // host; `hreq` holds the back-projection requests of one advection
    ifl_host_body_ops hops{};
    bool host_geom = false;
    struct HostReq* hreq = nullptr;
    int* hreq_cnt = nullptr;  // one counter per quantity
    size_t hreq_cap = 0;      // requests per quantity

    // level 8 particles
    int max_particles = 0;
    int particle_count = 0;
    double *posx = nullptr, *posy = nullptr, *prop[4] = {nullptr, nullptr, nullptr, nullptr};
    // the particle state of the reduced-precision forms (in place of the doubles above)
    float *fposx = nullptr, *fposy = nullptr, *fprop[4] = {nullptr, nullptr, nullptr, nullptr};
    double* weight = nullptr;
    int* counts = nullptr;
    bool particles_ready = false;  // initParticles has run
    // particle work buffers: sort keys / indices and flags (max_particles
    // each), per-cell integer scratch ((w+1)(h+1)+1 each), device scalars
    uint32_t *pkey = nullptr, *pkey2 = nullptr;
    int *pidx = nullptr, *pidx2 = nullptr;
    uint8_t* pflag = nullptr;
    int *cbuf0 = nullptr, *cbuf1 = nullptr, *cbuf2 = nullptr, *cbuf3 = nullptr;
    uint8_t* cflag = nullptr;
    // reduced-precision forms: the binned particle records and the
    // single-precision node accumulators of the LDS-binned fromParticles
    float4* rec32 = nullptr;
    float* rec32b = nullptr;
    float* acc32 = nullptr;
    void* flip_sc = nullptr;    // device FlipScalars
    void* flip_host = nullptr;  // pinned mirror
    void* cub_tmp = nullptr;    // hipcub temporary storage
    size_t cub_bytes = 0;

    // timing
    bool profiling = false;
    bool prof_open = false;  // the last prof_begin recorded an event
    Profile prof[IFL_KERNEL_COUNT];
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    struct Timed {
        int kernel, index;  // index: PCG iteration of the launch, or -1 (always counted)
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> ev_free;  // event pool for per-launch timing
    std::vector<Timed> timed;         // recorded, not yet collected

    int* step_flags = nullptr;   // device: [0] non-finite field values after a step, [1] border flag
    int* step_flags_host = nullptr;  // pinned mirror
    uint8_t* image_d = nullptr;  // toImage output, allocated on first use
    size_t image_bytes = 0;

    int border_solid = 0;
    bool border_pending = false;  // step_flags[1] holds this step's border flag
    size_t lds_bytes = 65536;  // dynamic LDS available to one workgroup
    int num_cus = 1;           // compute units: the strips of one sweep launch that can be resident at once
    std::atomic<int>* busy = nullptr;  // calls stepping solvers on this device right now
    bool batched = false;      // stepped by ifl_batch_update: the batch builds the MIC(0) factors
    // level 8 in a batch: particlesToGrid stops before its first host wait
    // (the prune's count) and the batch finishes every solver's prune, then
    // seeds them in batch order; the overflow check moves to the step's end
    bool flip_split = false;
    bool prune_pending = false;  // a prune's count is still to be read
};

// error helpers
#define IFL_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            s->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return IFL_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)
#define IFL_LAUNCH_CHECK()                                                               \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess) {                                                          \
            s->err = std::string("kernel launch: ") + hipGetErrorString(e_);             \
            return IFL_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)
#define IFL_TRY(call)                  \
    do {                               \
        int rc_ = (call);              \
        if (rc_ != IFL_OK) return rc_; \
    } while (0)

// A point of advect's trace that landed in a solid cell of a host-evaluated
// body: cell index, body, and the point in world units (7:428-431).
struct HostReq {
    int idx;
    int body;
    double x, y;
};

// ---- launchers (k_fields.hip) ----
int k_fill_solid(Solver* s, int qi);
int k_extrapolate(Solver* s, int qi);
int k_advect(Solver* s, int qmask, double dt);
int k_add_inflow(Solver* s, double x0, double y0, double x1, double y1, const double vals[4]);
int k_build_rhs(Solver* s);
int k_compute_densities(Solver* s);
int k_build_pressure(Solver* s, double dt);
int k_build_heat(Solver* s, double dt);
int k_apply_pressure(Solver* s, double dt);
int k_add_buoyancy(Solver* s, double dt);
int k_set_boundary(Solver* s);
int k_copy_t_to_r(Solver* s);
int k_copy_p_to_t(Solver* s);
int k_update_dflag(Solver* s);
int k_max_timestep(Solver* s, double* dt_out);  // solver 1's CFL bound
int k_nonfinite(Solver* s, int* dev_flag);      // any NaN / infinity in d, t, u, v
int k_to_image(Solver* s, uint8_t* dev_rgba, int heat, int y0, int rows);  // toImage rows y0.. into device memory
int edge_reset(Solver* s);  // the sweeps' edge buffers back to "not yet written" (api.hip)
// skewed <-> row-major conversion through s->stage_buf
int k_skew_to_rows(Solver* s, const double* skew, double* rows);
int k_rows_to_skew(Solver* s, const double* rows, double* skew);

// the linear solves of one update() (FluidSolver::project calls)
enum { SOLVE_NONE = 0, SOLVE_HEAT = 1, SOLVE_PRESSURE = 2,
       STAGE_EXTRAPOLATE = 3,  // batched solvers (4-7): d, u, v extrapolated by the batch
       STAGE_PRUNED = 4 };     // batched FLIP solvers: pruned on the device, the batch reads and seeds

// ---- launchers (k_pcg.hip) ----
int k_build_precon(Solver* s);
int k_apply_precon(Solver* s);       // z = M^-1 r (fluid cells), the stage form
int k_matvec(Solver* s);             // z = A s, the stage form
int k_dot(Solver* s, int which, double* out);  // 0: z.r  1: z.s
int k_inf_norm(Solver* s, double* out);
int k_project(Solver* s, int limit, ifl_solve_stats* stats);
struct BatchDev;
struct ExtrapDev;
struct StageDev;
// the batched form's extrapolations (extrapolate 7:560-588) of the quantities
// in qmask, in d, t, u, v order, for solvers of levels 4-7: one launch per
// phase for all of them. With `join` the batch's stream first waits for the
// solvers' own streams and they wait for it afterwards (their stages run on
// them); without, every stage runs on `stream`.
int k_extrapolate_batch(Solver* const* ss, int n, hipStream_t stream, ExtrapDev** dev, int qmask, bool join);
void extrap_dev_free(ExtrapDev* d);
// The batched form's grid stages of update() (levels 3-7, bodies evaluated on
// the device): each stage one launch on `stream` for every solver of the
// batch, the single-solver kernels' bodies over a table of the solvers'
// arguments. kb_begin (re)binds the solver list, makes `stream` wait for the
// solvers' own streams and resets the argument tables of the step.
int kb_begin(Solver* const* ss, int n, hipStream_t stream, StageDev** dev);
int kb_fill_solid(StageDev* d, int qi);  // fillSolidFields of quantity qi (d: + the FLUID flags)
int kb_set_boundary(StageDev* d);
int kb_build_rhs(StageDev* d);
int kb_compute_densities(StageDev* d);
int kb_build_matrix(StageDev* d, int heat, double dt);  // buildPressureMatrix / buildHeatDiffusionMatrix
int kb_apply_pressure(StageDev* d, double dt);
int kb_add_buoyancy(StageDev* d, double dt);
int kb_copy_t_to_r(StageDev* d);
int kb_copy_p_to_t(StageDev* d);
int kb_advect(StageDev* d, int qmask, double dt);
// the end of the step: per solver the non-finite flag and (with `border`) the
// border flag of the last extrapolation, read back with one copy (waits for
// the stream)
int kb_finish(StageDev* d, bool border, int* nonfinite, int* border_solid);
void stage_dev_free(StageDev* d);
// addInflow of several solvers of a batch (entry k: solver ss[inst[k]], the
// rectangle rect[4k..4k+3] = x, y, w, h and vals[4k..4k+3]); a solver's
// entries apply in their order; waits for `stream`
struct InflowDev;
int k_add_inflow_batch(Solver* const* ss, int n, hipStream_t stream, InflowDev** dev, int count, const int* inst,
                       const double* rect, const double* vals);
void inflow_dev_free(InflowDev* d);
// per-kernel totals of a batch's fused PCG launches (when profiling): every
// launch timed with events on the batch's stream; cell_iterations counts the
// cells of the instances that did work in them
struct BatchProfile {
    long long launches[IFL_KERNEL_COUNT] = {};
    double ms[IFL_KERNEL_COUNT] = {};
    double cell_iterations[IFL_KERNEL_COUNT] = {};
};
// project() of several solvers of one level at once, one launch per phase
// of each PCG iteration for all of them (stats: n entries). With `join` the
// solvers' stages run on their own streams, joined by events, and each
// solver's exit tail runs on its stream; without, the tails too run on
// `stream`, as one launch.
int k_project_batch(Solver* const* ss, int n, hipStream_t stream, BatchDev** dev, ifl_solve_stats* stats,
                    BatchProfile* prof, bool join);
void batch_dev_free(BatchDev* d);
// per-launch timing (when profiling): events on the solver's stream around
// a launch; prof_collect() adds the elapsed times of launches whose index is
// below the limit given for their kernel, after the stream has completed.
constexpr int kProfEvery = 8;  // PCG iterations per timed iteration
int prof_begin(Solver* s, int kernel, int index);
int prof_end(Solver* s);
int prof_collect(Solver* s, const int* index_limit);

// sweep diagnostics: time stamps per strip and 8-step block
constexpr int kTraceStamps = 6;
int k_trace_sweep(Solver* s, int kind, long long* dev_trace);

// ---- launchers (k_flip.hip) ----
int flip_alloc(Solver* s);
void flip_free(Solver* s);
int flip_init_particles(Solver* s);
int k_from_particles(Solver* s, int qi, bool with_t);
int k_extrapolate_flip(Solver* s, int qi);
int k_grid_to_particles(Solver* s, double alpha, bool cubic = false);  // cubic: cerp instead of lerp
int k_particles_to_grid(Solver* s);
// particlesToGrid up to the prune's kernels (no host wait), and the prune's
// count read back: k_particles_to_grid_begin, k_prune_finish, k_seed together
// are k_particles_to_grid
int k_particles_to_grid_begin(Solver* s);
int k_prune_finish(Solver* s);
int k_particle_advect(Solver* s, double dt);
int k_copy_old(Solver* s, int qi);
int k_diff(Solver* s, int qi, double alpha, int undo);
int k_count_particles(Solver* s);
int k_prune(Solver* s);
int k_seed(Solver* s);
int flip_check_overflow(Solver* s);
// frand()'s process-wide state (8:39), shared by every level-8 solver
uint32_t frand_state();
void frand_set_state(uint32_t state);

inline int grid_blocks(long n, int threads) { return (int)((n + threads - 1) / threads); }

}  // namespace ifl
