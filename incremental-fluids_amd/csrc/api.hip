// C ABI of libifl.so (declared in include/ifl.h): solver lifetime, state
// access, single stages and FluidSolver::update() for solvers 3..8.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ifl_solver.hpp"

using namespace ifl;

struct ifl_solver {
    Solver s;
};

struct ifl_batch {
    std::vector<ifl_solver*> handles;
    std::vector<Solver*> solvers;
    hipStream_t stream = nullptr;
    int device = 0;
    ifl::BatchDev* dev = nullptr;
    ifl::ExtrapDev* xdev = nullptr;
    ifl::StageDev* sdev = nullptr;
    ifl::InflowDev* idev = nullptr;
    bool profiling = false;
    ifl::BatchProfile prof;
};

namespace ifl {

// Device scratch of the allocation-free entry point ifl_to_image_static: the
// rendered rows are staged here. Each device holds its own copy for the life
// of the process. In host code `scratch` names a host-side stand-in, never
// the device memory: every use goes through g_scratch_dev[device], set once
// per device by scratch_init (hipGetSymbolAddress with that device current).
constexpr size_t kScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices];

// Calls stepping solvers on each device right now (ifl_update, ifl_stage,
// ifl_batch_update, from any thread): a sweep launched while another call
// is under way deals its strips upstream first (strip_of_block).
static std::atomic<int> g_busy[kMaxDevices];
struct BusyGuard {
    std::atomic<int>* c;
    explicit BusyGuard(int device) : c(device >= 0 && device < kMaxDevices ? &g_busy[device] : nullptr) {
        if (c) c->fetch_add(1, std::memory_order_relaxed);
    }
    ~BusyGuard() {
        if (c) c->fetch_sub(1, std::memory_order_relaxed);
    }
};

// With `device` current. Idempotent; the address is kept per device.
static int scratch_init(Solver* s) {
    if (s->device < 0 || s->device >= kMaxDevices) {
        s->err = "device ordinal beyond the scratch table";
        return IFL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[s->device]) return IFL_OK;
    void* ptr = nullptr;
    IFL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[s->device] = ptr;
    return IFL_OK;
}

static void* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
}

// frand() 8:38-46 keeps one static seed for the whole process; every level-8
// solver draws from it in creation and step order, as in the reference.
static std::mutex g_frand_mu;
static uint32_t g_frand = 0xBA5EBA11u;  // 8:39

uint32_t frand_state() {
    std::lock_guard<std::mutex> lock(g_frand_mu);
    return g_frand;
}

void frand_set_state(uint32_t state) {
    std::lock_guard<std::mutex> lock(g_frand_mu);
    g_frand = state;
}

static int ev_take(Solver* s, hipEvent_t* e) {
    if (!s->ev_free.empty()) {
        *e = s->ev_free.back();
        s->ev_free.pop_back();
        return IFL_OK;
    }
    IFL_HIP(hipEventCreate(e));
    return IFL_OK;
}

int prof_begin(Solver* s, int kernel, int index) {
    s->prof_open = false;
    if (!s->profiling) return IFL_OK;
    // Inside the PCG loop every kProfEvery-th iteration is timed: a sample
    // spread evenly over the solve (events around every launch cost ~2.5 %
    // of a step at 2048^2).
    if (index > 0 && index % kProfEvery != 0) return IFL_OK;
    Solver::Timed t = {kernel, index, nullptr, nullptr};
    IFL_TRY(ev_take(s, &t.a));
    IFL_HIP(hipEventRecord(t.a, s->stream));
    s->timed.push_back(t);
    s->prof_open = true;
    return IFL_OK;
}

int prof_end(Solver* s) {
    if (!s->prof_open) return IFL_OK;
    s->prof_open = false;
    if (!s->profiling || s->timed.empty() || s->timed.back().b) return IFL_OK;
    IFL_TRY(ev_take(s, &s->timed.back().b));
    IFL_HIP(hipEventRecord(s->timed.back().b, s->stream));
    return IFL_OK;
}

int prof_collect(Solver* s, const int* index_limit) {
    if (s->timed.empty()) return IFL_OK;
    IFL_HIP(hipStreamSynchronize(s->stream));
    for (auto& t : s->timed) {
        const bool counted = t.b && (t.index < 0 || !index_limit || t.index < index_limit[t.kernel]);
        if (counted) {
            float ms = 0;
            IFL_HIP(hipEventElapsedTime(&ms, t.a, t.b));
            Profile& pr = s->prof[t.kernel];
            pr.launches += 1;
            pr.ms += ms;
            pr.cell_iterations += (double)s->W * s->H;
        }
        s->ev_free.push_back(t.a);
        if (t.b) s->ev_free.push_back(t.b);
    }
    s->timed.clear();
    return IFL_OK;
}

}  // namespace ifl

namespace {

thread_local std::string g_err;

__global__ void fill_u64_kernel(uint64_t* p, long n, uint64_t v) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_f64_kernel(double* p, long n, double v) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int dalloc(Solver* s, double** p, size_t n) {
    n = n ? n : 1;
    IFL_HIP(hipMalloc((void**)p, n * sizeof(double)));
    IFL_HIP(hipMemsetAsync(*p, 0, n * sizeof(double), s->stream));
    return IFL_OK;
}
int falloc(Solver* s, float** p, size_t n) {
    n = n ? n : 1;
    IFL_HIP(hipMalloc((void**)p, n * sizeof(float)));
    IFL_HIP(hipMemsetAsync(*p, 0, n * sizeof(float), s->stream));
    return IFL_OK;
}
int balloc(Solver* s, uint8_t** p, size_t n) {
    n = (n + 7) & ~(size_t)7;  // whole 32-bit words for the mask atomics
    IFL_HIP(hipMalloc((void**)p, n));
    IFL_HIP(hipMemsetAsync(*p, 0, n, s->stream));
    return IFL_OK;
}
int fill(Solver* s, double* p, long n, double v) {
    fill_f64_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(p, n, v);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

bool skewed_field(int f) { return f >= IFL_FIELD_R && f <= IFL_FIELD_APLUSY; }

double* skew_ptr(Solver* s, int f) {
    switch (f) {
    case IFL_FIELD_R: return s->r;
    case IFL_FIELD_P: return s->p;
    case IFL_FIELD_Z: return s->z;
    case IFL_FIELD_S: return s->s;
    case IFL_FIELD_PRECON: return s->precon;
    case IFL_FIELD_ADIAG: return s->adiag;
    case IFL_FIELD_APLUSX: return s->aplusx;
    case IFL_FIELD_APLUSY: return s->aplusy;
    }
    return nullptr;
}

// device pointer, element count and size of a non-skewed field
void* field_ptr(Solver* s, int f, size_t* count, int* elem) {
    *count = 0;
    *elem = 8;
    int lv = s->level;
    if (f < 64) {
        int qi = f / 16, k = f % 16;
        if (qi > 3 || (qi == IFL_Q_T && lv < 6)) return nullptr;
        QtyView& q = s->q[qi];
        size_t n = (size_t)q.w * q.h;
        *count = n;
        switch (k) {
        case IFL_QF_SRC: return q.src;
        case IFL_QF_DST: return q.dst;
        case IFL_QF_PHI: if (lv < 5) break; *count = (size_t)(q.w + 1) * (q.h + 1); return q.phi;
        case IFL_QF_VOLUME: if (lv < 5) break; return q.volume;
        case IFL_QF_NORMALX: if (lv < 4) break; return q.nx;
        case IFL_QF_NORMALY: if (lv < 4) break; return q.ny;
        case IFL_QF_CELL: if (lv < 4) break; *elem = 1; return q.cell;
        case IFL_QF_BODY: if (lv < 4) break; *elem = 1; return q.body;
        case IFL_QF_MASK: if (lv < 4) break; *elem = 1; return q.mask;
        }
        *count = 0;
        return nullptr;
    }
    size_t n = (size_t)s->W * s->H;
    switch (f) {
    case IFL_FIELD_UDENSITY: if (lv < 7) return nullptr; *count = (size_t)(s->W + 1) * s->H; return s->udens;
    case IFL_FIELD_VDENSITY: if (lv < 7) return nullptr; *count = (size_t)s->W * (s->H + 1); return s->vdens;
    }
    if (lv != 8) return nullptr;
    switch (f) {
    case IFL_FIELD_POSX: case IFL_FIELD_POSY:
    case IFL_FIELD_PROP_D: case IFL_FIELD_PROP_T: case IFL_FIELD_PROP_U: case IFL_FIELD_PROP_V: {
        *count = s->max_particles;
        const int k = f == IFL_FIELD_POSX ? 0 : f == IFL_FIELD_POSY ? 1 : 2 + (f - IFL_FIELD_PROP_D);
        if (s->fposx) {  // float particle state
            *elem = 4;
            return k == 0 ? s->fposx : k == 1 ? s->fposy : s->fprop[k - 2];
        }
        return k == 0 ? s->posx : k == 1 ? s->posy : s->prop[k - 2];
    }
    case IFL_FIELD_WEIGHT: *count = (size_t)(s->W + 1) * (s->H + 1); return s->weight;
    case IFL_FIELD_COUNTS: *count = n; *elem = 4; return s->counts;
    }
    return nullptr;
}

void free_all(Solver* s) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    for (int i = 0; i < 4; i++) {
        QtyView& q = s->q[i];
        F(q.src); F(q.dst); F(q.phi); F(q.volume); F(q.nx); F(q.ny); F(q.cell); F(q.body); F(q.mask);
    }
    F(s->udens); F(s->vdens);
    F(s->r); F(s->p); F(s->z); F(s->s); F(s->s2); F(s->precon); F(s->adiag); F(s->aplusx); F(s->aplusy);
    F(s->cx); F(s->cy); F(s->yv); F(s->qv); F(s->dflag); F(s->fmask); F(s->edge);
    F(s->cx32); F(s->cy32); F(s->pc32); F(s->y32); F(s->partial); F(s->st);
    F(s->e32); F(s->r32); F(s->s32); F(s->z32); F(s->q32); F(s->ad32); F(s->apx32); F(s->apy32); F(s->st_outer);
    F(s->stage_buf); F(s->front); F(s->front_cnt); F(s->bodies_d); F(s->hreq); F(s->hreq_cnt);
    F(s->posx); F(s->posy); F(s->fposx); F(s->fposy);
    for (int t = 0; t < 4; t++) { F(s->prop[t]); F(s->fprop[t]); }
    F(s->weight); F(s->counts); F(s->image_d); F(s->step_flags);
    if (s->step_flags_host) (void)hipHostFree(s->step_flags_host);
    flip_free(s);
    if (s->st_host) (void)hipHostFree(s->st_host);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    for (auto& t : s->timed) {
        if (t.a) (void)hipEventDestroy(t.a);
        if (t.b) (void)hipEventDestroy(t.b);
    }
    s->timed.clear();
    for (hipEvent_t e : s->ev_free) (void)hipEventDestroy(e);
    s->ev_free.clear();
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

int init_solver(Solver* s) {
    IFL_HIP(hipSetDevice(s->device));
    IFL_TRY(scratch_init(s));
    IFL_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    IFL_HIP(hipEventCreate(&s->ev0));
    IFL_HIP(hipEventCreate(&s->ev1));
    int lds = 0;
    IFL_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device));
    s->lds_bytes = (size_t)lds;
    int cus = 0;
    IFL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    s->num_cus = cus > 0 ? cus : 1;
    s->busy = s->device >= 0 && s->device < kMaxDevices ? &g_busy[s->device] : nullptr;
    int W = s->W, H = s->H, lv = s->level;
    // FluidSolver ctor 7:953-958: d, t at cell centres, u, v on faces
    const int qw[4] = {W, W, W + 1, W};
    const int qh[4] = {H, H, H, H + 1};
    const double qox[4] = {0.5, 0.5, 0.0, 0.5};
    const double qoy[4] = {0.5, 0.5, 0.5, 0.0};
    for (int i = 0; i < 4; i++) {
        QtyView& q = s->q[i];
        q.w = qw[i];
        q.h = qh[i];
        q.ox = qox[i];
        q.oy = qoy[i];
        q.hx = s->hx;
        size_t n = (size_t)q.w * q.h;
        IFL_TRY(dalloc(s, &q.src, n));
        IFL_TRY(dalloc(s, &q.dst, n));
        IFL_TRY(dalloc(s, &q.phi, (size_t)(q.w + 1) * (q.h + 1)));
        IFL_TRY(dalloc(s, &q.volume, n));
        IFL_TRY(dalloc(s, &q.nx, n));
        IFL_TRY(dalloc(s, &q.ny, n));
        IFL_TRY(balloc(s, &q.cell, n));
        IFL_TRY(balloc(s, &q.body, n));
        IFL_TRY(balloc(s, &q.mask, n));
        if (lv >= 5) IFL_TRY(fill(s, q.volume, (long)n, 1.0));  // 7:332-335
    }
    if (lv >= 6) IFL_TRY(fill(s, s->q[IFL_Q_T].src, (long)W * H, s->cfg.t_amb));  // 7:960-961
    IFL_TRY(dalloc(s, &s->udens, (size_t)(W + 1) * H));
    s->vdens_count = (size_t)(W + 1) * (H + 1);
    IFL_TRY(dalloc(s, &s->vdens, s->vdens_count));

    Skew& sk = s->sk;
    sk.W = W;
    sk.H = H;
    sk.nstrips = (H + kStrip - 1) / kStrip;
    sk.cols = ((W + kStrip - 1) + 15) / 16 * 16;
    sk.strip_sz = (long)sk.cols * kStrip;
    s->sk_count = (size_t)sk.nstrips * sk.strip_sz;
    size_t n = s->sk_count;
    double** arrs[] = {&s->r, &s->p, &s->z, &s->s, &s->precon, &s->adiag, &s->aplusx, &s->aplusy,
                       &s->cx, &s->cy, &s->yv, &s->qv};
    for (double** a : arrs) IFL_TRY(dalloc(s, a, n));
    IFL_TRY(balloc(s, &s->dflag, n));
    // the fused vector update of the PCG iterations (levels 6-8, FP64);
    // IFL_NO_FUSED_UPDATE=1 in the environment at creation keeps the
    // separate update kernel (diagnostics: the two give identical iterates)
    const char* nofuse = getenv("IFL_NO_FUSED_UPDATE");
    if (lv >= 6 && s->cfg.precision == IFL_PREC_FP64 && !(nofuse && nofuse[0] == '1')) IFL_TRY(dalloc(s, &s->s2, n));
    if (s->cfg.precision != IFL_PREC_FP64) {
        float** sarrs[] = {&s->cx32, &s->cy32, &s->pc32, &s->y32, &s->e32, &s->r32, &s->s32,
                           &s->z32, &s->q32, &s->ad32, &s->apx32, &s->apy32};
        for (float** a : sarrs) {
            IFL_HIP(hipMalloc((void**)a, n * sizeof(float)));
            IFL_HIP(hipMemsetAsync(*a, 0, n * sizeof(float), s->stream));
        }
        IFL_HIP(hipMalloc((void**)&s->st_outer, sizeof(PcgState)));
        IFL_HIP(hipMemsetAsync(s->st_outer, 0, sizeof(PcgState), s->stream));
    }
    IFL_HIP(hipMalloc((void**)&s->fmask, (size_t)sk.nstrips * sk.cols * sizeof(uint64_t)));
    IFL_HIP(hipMemsetAsync(s->fmask, 0, (size_t)sk.nstrips * sk.cols * sizeof(uint64_t), s->stream));
    IFL_HIP(hipMalloc((void**)&s->edge, 3L * sk.nstrips * W * sizeof(double)));
    IFL_TRY(edge_reset(s));
    s->npartial = std::max(grid_blocks((long)n, 256), sk.nstrips) + 256;
    IFL_TRY(dalloc(s, &s->partial, s->npartial));
    IFL_HIP(hipMalloc((void**)&s->st, sizeof(PcgState)));
    IFL_HIP(hipMemsetAsync(s->st, 0, sizeof(PcgState), s->stream));
    IFL_HIP(hipHostMalloc((void**)&s->st_host, sizeof(PcgState), hipHostMallocDefault));
    memset(s->st_host, 0, sizeof(PcgState));
    s->stage_count = (size_t)(W + 2) * (H + 2);
    IFL_TRY(dalloc(s, &s->stage_buf, s->stage_count));
    long cap = (long)(W + 1) * (H + 1);
    IFL_HIP(hipMalloc((void**)&s->front, 2 * cap * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->front_cnt, 2 * sizeof(int)));
    IFL_HIP(hipMemsetAsync(s->front_cnt, 0, 2 * sizeof(int), s->stream));
    IFL_HIP(hipMalloc((void**)&s->bodies_d, kMaxBodies * sizeof(Body)));
    IFL_HIP(hipMalloc((void**)&s->hreq_cnt, 4 * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->step_flags, 2 * sizeof(int)));
    IFL_HIP(hipHostMalloc((void**)&s->step_flags_host, 2 * sizeof(int), hipHostMallocDefault));
    s->step_flags_host[0] = s->step_flags_host[1] = 0;
    if (lv == 8) {
        // ParticleQuantities 8:865-878: room for 12 particles per cell; the
        // property arrays start at zero (8:896-897)
        s->max_particles = W * H * 12;
        size_t mp = (size_t)s->max_particles;
        if (s->cfg.precision == IFL_PREC_FP64) {
            IFL_TRY(dalloc(s, &s->posx, mp));
            IFL_TRY(dalloc(s, &s->posy, mp));
            for (int t = 0; t < 4; t++) IFL_TRY(dalloc(s, &s->prop[t], mp));
        } else {  // the reduced-precision forms keep the particles in float
            IFL_TRY(falloc(s, &s->fposx, mp));
            IFL_TRY(falloc(s, &s->fposy, mp));
            for (int t = 0; t < 4; t++) IFL_TRY(falloc(s, &s->fprop[t], mp));
        }
        IFL_TRY(dalloc(s, &s->weight, (size_t)(W + 1) * (H + 1)));
        IFL_HIP(hipMalloc((void**)&s->counts, (size_t)W * H * sizeof(int)));
        IFL_HIP(hipMemsetAsync(s->counts, 0, (size_t)W * H * sizeof(int), s->stream));
        IFL_TRY(flip_alloc(s));
    }
    IFL_TRY(k_update_dflag(s));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int set_err(Solver* s, int rc, const char* msg) {
    s->err = msg;
    return rc;
}

int solve(Solver* s, ifl_solve_stats* st) {
    ifl_solve_stats local;
    return k_project(s, s->cfg.pcg_limit, st ? st : &local);
}

void flip_q(Solver* s, int qi) { std::swap(s->q[qi].src, s->q[qi].dst); }

// The border flag of the last extrapolation goes to step_flags[1] on the
// device and reaches the host with the end-of-step read (no wait here).
int defer_border_flag(Solver* s) {
    IFL_HIP(hipMemcpyAsync(s->step_flags + 1, s->front_cnt + 1, sizeof(int), hipMemcpyDeviceToDevice, s->stream));
    s->border_pending = true;
    return IFL_OK;
}

int extrapolate(Solver* s, int qi) {
    if (s->level == 8) return k_extrapolate_flip(s, qi);
    IFL_TRY(k_extrapolate(s, qi));
    return IFL_OK;
}

int fill_solid(Solver* s, int qi) {
    IFL_TRY(k_fill_solid(s, qi));
    if (qi == IFL_Q_D) IFL_TRY(k_update_dflag(s));
    return IFL_OK;
}

// FluidSolver::update for levels 3..8 (3:433-447, 4:869-895, 5:927-953,
// 6:966-1008, 7:995-1034, 8:1350-1413), split at its linear solves so that a
// batch of solvers can run each solve together: update_part runs from the
// solve `after` (SOLVE_NONE at the start of the step) up to the next one and
// names it in *next (SOLVE_NONE once the step is complete).
int update_part(Solver* s, int after, double dt, int* next) {
    const int lv = s->level;
    *next = SOLVE_NONE;
    if (after == SOLVE_NONE || after == STAGE_PRUNED) {
        if (after == SOLVE_NONE) {
            s->border_solid = 0;
            s->border_pending = false;
            if (lv == 3) {
                IFL_TRY(k_build_rhs(s));
                IFL_TRY(k_build_pressure(s, dt));
                if (!s->batched) IFL_TRY(k_build_precon(s));
                *next = SOLVE_PRESSURE;
                return IFL_OK;
            }
            if (lv == 8 && !s->particles_ready) IFL_TRY(flip_init_particles(s));
            IFL_TRY(fill_solid(s, IFL_Q_D));
            if (lv >= 6) IFL_TRY(fill_solid(s, IFL_Q_T));
            IFL_TRY(fill_solid(s, IFL_Q_U));
            IFL_TRY(fill_solid(s, IFL_Q_V));
            if (lv == 8 && s->flip_split) {
                // the batch reads every solver's prune and seeds them in order
                IFL_TRY(k_particles_to_grid_begin(s));
                *next = STAGE_PRUNED;
                return IFL_OK;
            }
            // 8:1357-1368: particles to grid, the pre-update copies, and the
            // inflow the FLIP solver applies itself
            if (lv == 8) IFL_TRY(k_particles_to_grid(s));
        }
        if (lv == 8) {
            IFL_TRY(k_update_dflag(s));
            if (s->cfg.verbose) printf("Particle count: %d\n", s->particle_count);
            for (int qi = 0; qi < 4; qi++) IFL_TRY(k_copy_old(s, qi));
            const double x = 0.45, y = 0.2, w = 0.2, h = 0.05;
            const double vals[4] = {1.0, s->cfg.t_amb, 0.0, 0.0};
            IFL_TRY(k_add_inflow(s, x, y, x + w, y + h, vals));
        }
        if (lv >= 6) {
            IFL_TRY(k_copy_t_to_r(s));
            IFL_TRY(k_build_heat(s, dt));
            if (!s->batched) IFL_TRY(k_build_precon(s));
            *next = SOLVE_HEAT;
            return IFL_OK;
        }
    }
    if (after == SOLVE_HEAT) {
        IFL_TRY(k_copy_p_to_t(s));
        IFL_TRY(extrapolate(s, IFL_Q_T));
        IFL_TRY(k_add_buoyancy(s, dt));
    }
    if (after != SOLVE_PRESSURE && after != STAGE_EXTRAPOLATE) {
        IFL_TRY(k_set_boundary(s));
        IFL_TRY(k_build_rhs(s));
        if (lv >= 7) IFL_TRY(k_compute_densities(s));
        IFL_TRY(k_build_pressure(s, dt));
        if (!s->batched) IFL_TRY(k_build_precon(s));
        *next = SOLVE_PRESSURE;
        return IFL_OK;
    }
    if (after == SOLVE_PRESSURE) {
        IFL_TRY(k_apply_pressure(s, dt));
        if (lv == 3) {
            IFL_TRY(k_advect(s, (1 << IFL_Q_D) | (1 << IFL_Q_U) | (1 << IFL_Q_V), dt));
            flip_q(s, IFL_Q_D);
            flip_q(s, IFL_Q_U);
            flip_q(s, IFL_Q_V);
            return IFL_OK;
        }
        if (s->batched && lv < 8) {
            *next = STAGE_EXTRAPOLATE;  // k_extrapolate_batch
            return IFL_OK;
        }
        IFL_TRY(extrapolate(s, IFL_Q_D));
        IFL_TRY(extrapolate(s, IFL_Q_U));
        IFL_TRY(extrapolate(s, IFL_Q_V));
    }
    if (lv < 8) IFL_TRY(defer_border_flag(s));
    IFL_TRY(k_set_boundary(s));
    if (lv == 8) {
        // 8:1393-1412: FLIP update of the particles, then their advection
        const double alpha = s->cfg.flip_alpha;
        for (int qi = 0; qi < 4; qi++) IFL_TRY(k_diff(s, qi, alpha, 0));
        IFL_TRY(k_grid_to_particles(s, alpha, s->cfg.flip_g2p == 1));
        for (int qi = 0; qi < 4; qi++) IFL_TRY(k_diff(s, qi, alpha, 1));
        IFL_TRY(k_particle_advect(s, dt));
        if (!s->flip_split) IFL_TRY(flip_check_overflow(s));  // else at the batch's step end
    } else {
        int qm = (1 << IFL_Q_D) | (1 << IFL_Q_U) | (1 << IFL_Q_V) | (lv >= 6 ? (1 << IFL_Q_T) : 0);
        IFL_TRY(k_advect(s, qm, dt));
        flip_q(s, IFL_Q_D);
        if (lv >= 6) flip_q(s, IFL_Q_T);
        flip_q(s, IFL_Q_U);
        flip_q(s, IFL_Q_V);
    }
    return IFL_OK;
}

// update_part for a whole batch of solvers of one level (3-7, bodies
// evaluated on the device): the same stages in the same order, each one
// launch for every solver on the batch's stream (kb_*); the extrapolations
// are STAGE_EXTRAPOLATE (k_extrapolate_batch), except t's after the heat
// solve, which runs here.
int update_part_batch(ifl_batch* b, int after, double dt, int* next) {
    StageDev* d = b->sdev;
    const int lv = b->solvers[0]->level;
    *next = SOLVE_NONE;
    if (after == SOLVE_NONE) {
        for (Solver* s : b->solvers) {
            s->border_solid = 0;
            s->border_pending = false;
        }
        if (lv == 3) {
            IFL_TRY(kb_build_rhs(d));
            IFL_TRY(kb_build_matrix(d, 0, dt));
            *next = SOLVE_PRESSURE;
            return IFL_OK;
        }
        IFL_TRY(kb_fill_solid(d, IFL_Q_D));
        if (lv >= 6) IFL_TRY(kb_fill_solid(d, IFL_Q_T));
        IFL_TRY(kb_fill_solid(d, IFL_Q_U));
        IFL_TRY(kb_fill_solid(d, IFL_Q_V));
        if (lv >= 6) {
            IFL_TRY(kb_copy_t_to_r(d));
            IFL_TRY(kb_build_matrix(d, 1, dt));
            *next = SOLVE_HEAT;
            return IFL_OK;
        }
    }
    if (after == SOLVE_HEAT) {
        IFL_TRY(kb_copy_p_to_t(d));
        IFL_TRY(k_extrapolate_batch(b->solvers.data(), (int)b->solvers.size(), b->stream, &b->xdev, 1 << IFL_Q_T,
                                    false));
        IFL_TRY(kb_add_buoyancy(d, dt));
    }
    if (after != SOLVE_PRESSURE && after != STAGE_EXTRAPOLATE) {
        IFL_TRY(kb_set_boundary(d));
        IFL_TRY(kb_build_rhs(d));
        if (lv >= 7) IFL_TRY(kb_compute_densities(d));
        IFL_TRY(kb_build_matrix(d, 0, dt));
        *next = SOLVE_PRESSURE;
        return IFL_OK;
    }
    if (after == SOLVE_PRESSURE) {
        IFL_TRY(kb_apply_pressure(d, dt));
        if (lv == 3) {
            IFL_TRY(kb_advect(d, (1 << IFL_Q_D) | (1 << IFL_Q_U) | (1 << IFL_Q_V), dt));
            for (Solver* s : b->solvers) {
                flip_q(s, IFL_Q_D);
                flip_q(s, IFL_Q_U);
                flip_q(s, IFL_Q_V);
            }
            return IFL_OK;
        }
        *next = STAGE_EXTRAPOLATE;
        return IFL_OK;
    }
    // after the extrapolations of d, u and v: their border flag is read at
    // the end of the step (kb_finish)
    for (Solver* s : b->solvers) s->border_pending = true;
    IFL_TRY(kb_set_boundary(d));
    const int qm = (1 << IFL_Q_D) | (1 << IFL_Q_U) | (1 << IFL_Q_V) | (lv >= 6 ? (1 << IFL_Q_T) : 0);
    IFL_TRY(kb_advect(d, qm, dt));
    for (Solver* s : b->solvers) {
        flip_q(s, IFL_Q_D);
        if (lv >= 6) flip_q(s, IFL_Q_T);
        flip_q(s, IFL_Q_U);
        flip_q(s, IFL_Q_V);
    }
    return IFL_OK;
}

// the end of a step: the non-finite check, timing collection, statistics
int finish_step(Solver* s, const ifl_solve_stats& heat, const ifl_solve_stats& press, ifl_step_stats* stats) {
    IFL_TRY(k_nonfinite(s, s->step_flags));
    IFL_HIP(hipMemcpyAsync(s->step_flags_host, s->step_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    if (s->border_pending) s->border_solid |= s->step_flags_host[1];
    s->border_pending = false;
    IFL_TRY(prof_collect(s, nullptr));
    if (stats) {
        stats->nonfinite = s->step_flags_host[0];
        stats->reserved = 0;
        stats->heat = heat;
        stats->pressure = press;
        stats->particle_count = s->particle_count;
        stats->border_solid = s->border_solid;
    }
    return IFL_OK;
}

int update_grid(Solver* s, double dt, ifl_step_stats* stats) {
    ifl_solve_stats heat = {IFL_SOLVE_NOT_RUN, 0, 0.0}, press = heat;
    int after = SOLVE_NONE, next = SOLVE_NONE;
    while (true) {
        IFL_TRY(update_part(s, after, dt, &next));
        if (next == SOLVE_NONE) break;
        s->solve_kind = next;
        IFL_TRY(solve(s, next == SOLVE_HEAT ? &heat : &press));
        s->solve_kind = SOLVE_PRESSURE;
        after = next;
    }
    return finish_step(s, heat, press, stats);
}

// PNG (RGBA, 8 bits per channel, what the reference's lodepng_encode32_file
// call writes, 7:1122) with the image data in stored deflate blocks: a valid zlib
// stream (RFC 1950/1951) without compression, CRC-32 per chunk and Adler-32
// over the filtered scanlines (filter byte 0 per row).
struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t k = i;
            for (int b = 0; b < 8; b++) k = (k & 1) ? 0xEDB88320u ^ (k >> 1) : k >> 1;
            t[i] = k;
        }
    }
};

uint32_t crc32_png(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    static const CrcTable table;  // initialised once, thread-safe
    for (size_t i = 0; i < n; i++) c = table.t[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void put32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void put_chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data) {
    put32(o, (uint32_t)data.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, crc32_png(&o[start], o.size() - start) ^ 0xFFFFFFFFu);
}

bool encode_png(const uint8_t* rgba, int w, int h, std::vector<uint8_t>& out) {
    const size_t row = (size_t)w * 4 + 1;
    const size_t raw_n = row * (size_t)h;
    if (raw_n / row != (size_t)h || raw_n > ((size_t)1 << 34)) return false;
    std::vector<uint8_t> raw(raw_n);
    for (int y = 0; y < h; y++) {
        raw[y * row] = 0;  // filter: none
        memcpy(&raw[y * row + 1], rgba + (size_t)y * w * 4, (size_t)w * 4);
    }
    std::vector<uint8_t> z;
    z.push_back(0x78);  // deflate, 32 KB window
    z.push_back(0x01);  // no preset dictionary, fastest; (0x78 << 8 | 0x01) % 31 == 0
    for (size_t off = 0; off < raw_n || off == 0; off += 65535) {
        const size_t len = std::min<size_t>(65535, raw_n - off);
        const bool last = off + len >= raw_n;
        z.push_back(last ? 1 : 0);  // BFINAL, BTYPE = 00 (stored)
        z.push_back((uint8_t)len);
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len);
        z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (size_t i = 0; i < raw_n; i++) {
        a = (a + raw[i]) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    out.assign(sig, sig + 8);
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(6);  // colour type: RGBA
    ihdr.push_back(0);  // compression: deflate
    ihdr.push_back(0);  // filter method 0
    ihdr.push_back(0);  // no interlace
    put_chunk(out, "IHDR", ihdr);
    put_chunk(out, "IDAT", z);
    put_chunk(out, "IEND", std::vector<uint8_t>());
    return true;
}

}  // namespace

// The sweeps' edge buffers (3 kinds x strips x W) back to the sentinel that
// marks an edge not yet written; after a sweep that gave up, edges it left
// behind would otherwise be taken for the next sweep's.
int ifl::edge_reset(Solver* s) {
    long ne = 3L * s->sk.nstrips * s->W;
    fill_u64_kernel<<<grid_blocks(ne, 256), 256, 0, s->stream>>>((uint64_t*)s->edge, ne, 0x7FF4DEADBEEF1234ull);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

extern "C" {

const char* ifl_version(void) { return "ifl 1.0 (gfx950, solvers 3-8)"; }

const char* ifl_last_error(const ifl_solver* h) {
    if (!h) return g_err.c_str();
    return h->s.err.c_str();
}

int ifl_config_default(ifl_config* cfg, int level, int w, int h) {
    if (!cfg) return IFL_ERR_ARG;
    memset(cfg, 0, sizeof *cfg);
    cfg->level = level;
    cfg->w = w;
    cfg->h = h;
    cfg->precision = IFL_PREC_FP64;
    cfg->dot_order = IFL_DOT_TREE;
    cfg->pcg_limit = level == 3 ? 600 : 2000;  // 3:437, 7:1004/1016
    cfg->verbose = 0;
    cfg->device = 0;
    cfg->density = 0.1;          // the reference mains' constants (7:1089-1091)
    cfg->density_soot = level == 8 ? 0.25 : (level == 7 ? 1.0 : 0.1);
    cfg->diffusion = 0.01;
    cfg->tau = 0.97;             // 7:740
    cfg->sigma = 0.25;           // 7:741
    cfg->tol = 1e-5;             // 7:859
    cfg->t_amb = 294.0;          // 7:950
    cfg->gravity = 9.81;         // 7:951
    cfg->flip_alpha = 0.001;     // 8:1296
    return IFL_OK;
}

int ifl_create(const ifl_config* cfg, ifl_solver** out) {
    if (!cfg || !out) return IFL_ERR_ARG;
    *out = nullptr;
    if (cfg->level < 3 || cfg->level > 8 || cfg->w < 3 || cfg->h < 3 || (long)cfg->w * cfg->h > (1L << 30)) {
        g_err = "ifl_create: level must be 3..8 and w, h >= 3";
        return IFL_ERR_ARG;
    }
    if (cfg->level == 8 && (long)cfg->w * cfg->h * 12 > 0x7FFFFFFFL) {
        g_err = "ifl_create: level 8 keeps 12 particles per cell; w*h*12 must fit in an int";
        return IFL_ERR_ARG;
    }
    if (cfg->precision != IFL_PREC_FP64 && cfg->precision != IFL_PREC_MIXED && cfg->precision != IFL_PREC_REFINE) {
        g_err = "ifl_create: precision must be IFL_PREC_FP64, IFL_PREC_MIXED or IFL_PREC_REFINE";
        return IFL_ERR_ARG;
    }
    if (cfg->flip_g2p != 0 && cfg->flip_g2p != 1) {
        g_err = "ifl_create: flip_g2p must be 0 (bilinear, the reference) or 1 (Catmull-Rom)";
        return IFL_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device) {
        g_err = "ifl_create: no HIP device";
        return IFL_ERR_NO_DEVICE;
    }
    ifl_solver* h = new (std::nothrow) ifl_solver();
    if (!h) return IFL_ERR_ALLOC;
    Solver* s = &h->s;
    s->cfg = *cfg;
    s->level = cfg->level;
    s->W = cfg->w;
    s->H = cfg->h;
    s->hx = 1.0 / std::min(cfg->w, cfg->h);  // 7:953
    s->device = cfg->device;
    int rc = init_solver(s);
    if (rc != IFL_OK) {
        g_err = s->err;
        free_all(s);
        delete h;
        return rc;
    }
    *out = h;
    return IFL_OK;
}

int ifl_destroy(ifl_solver* h) {
    if (!h) return IFL_OK;
    (void)hipSetDevice(h->s.device);
    (void)hipStreamSynchronize(h->s.stream);
    free_all(&h->s);
    delete h;
    return IFL_OK;
}

int ifl_set_bodies(ifl_solver* h, const ifl_body* b, int n) {
    if (!h || n < 0 || n > kMaxBodies || (n > 0 && !b)) return IFL_ERR_ARG;
    Solver* s = &h->s;
    if (s->level < 4 && n > 0) return set_err(s, IFL_ERR_UNSUPPORTED, "solver 3 has no solid bodies");
    bool host = false;
    for (int i = 0; i < n; i++) {
        if (b[i].type < IFL_BODY_BOX || b[i].type > IFL_BODY_HOST) return set_err(s, IFL_ERR_ARG, "unknown body type");
        host = host || b[i].type == IFL_BODY_HOST;
    }
    if (host && s->level == 8)
        return set_err(s, IFL_ERR_UNSUPPORTED, "host-evaluated bodies are not available at level 8");
    if (host && !(s->hops.distance && s->hops.closest && s->hops.normal))
        return set_err(s, IFL_ERR_ARG, "a host-evaluated body needs ifl_set_host_body_ops first");
    s->host_geom = host;
    for (int i = 0; i < n; i++) {
        Body& d = s->bodies_h[i];
        d.type = b[i].type;
        d.px = b[i].pos_x;
        d.py = b[i].pos_y;
        d.sx = b[i].scale_x;
        d.sy = b[i].type == IFL_BODY_SPHERE ? b[i].scale_x : b[i].scale_y;  // SolidSphere 7:223-224
        d.theta = b[i].theta;
        d.vx = b[i].vel_x;
        d.vy = b[i].vel_y;
        d.vt = b[i].vel_theta;
        d.cn = b[i].cos_neg;
        d.sn = b[i].sin_neg;
        d.cp = b[i].cos_pos;
        d.sp = b[i].sin_pos;
    }
    s->nbodies = n;
    IFL_HIP(hipSetDevice(s->device));
    if (n > 0) IFL_HIP(hipMemcpyAsync(s->bodies_d, s->bodies_h, n * sizeof(Body), hipMemcpyHostToDevice, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_set_host_body_ops(ifl_solver* h, const ifl_host_body_ops* ops) {
    if (!h) return IFL_ERR_ARG;
    Solver* s = &h->s;
    if (ops && !(ops->distance && ops->closest && ops->normal))
        return set_err(s, IFL_ERR_ARG, "ifl_host_body_ops needs all three callbacks");
    if (!ops && s->host_geom) return set_err(s, IFL_ERR_ARG, "the body list holds a host-evaluated body");
    if (ops)
        s->hops = *ops;
    else
        s->hops = ifl_host_body_ops{};
    return IFL_OK;
}

int ifl_add_inflow(ifl_solver* h, double x, double y, double w, double hh, const double vals[4]) {
    if (!h || !vals) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipSetDevice(s->device));
    IFL_TRY(k_add_inflow(s, x, y, x + w, y + hh, vals));  // 7:1036-1041
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_update(ifl_solver* h, double dt, ifl_step_stats* stats) {
    if (!h) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipSetDevice(s->device));
    BusyGuard busy(s->device);
    return update_grid(s, dt, stats);
}

int ifl_field_info(const ifl_solver* hc, int field, size_t* count, int* elem_bytes) {
    if (!hc || !count || !elem_bytes) return IFL_ERR_ARG;
    Solver* s = const_cast<Solver*>(&hc->s);
    if (skewed_field(field)) {
        *count = (size_t)s->W * s->H;
        *elem_bytes = 8;
        return IFL_OK;
    }
    field_ptr(s, field, count, elem_bytes);
    return *count ? IFL_OK : IFL_ERR_ARG;
}

int ifl_upload(ifl_solver* h, int field, const void* host, size_t bytes) {
    if (!h || !host) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipSetDevice(s->device));
    if (skewed_field(field)) {
        size_t n = (size_t)s->W * s->H;
        if (bytes != n * 8) return set_err(s, IFL_ERR_ARG, "ifl_upload: size mismatch");
        IFL_HIP(hipMemcpyAsync(s->stage_buf, host, bytes, hipMemcpyHostToDevice, s->stream));
        IFL_TRY(k_rows_to_skew(s, s->stage_buf, skew_ptr(s, field)));
        IFL_HIP(hipStreamSynchronize(s->stream));
        return IFL_OK;
    }
    size_t n;
    int e;
    void* p = field_ptr(s, field, &n, &e);
    if (!p) return set_err(s, IFL_ERR_ARG, "ifl_upload: no such field at this level");
    if (bytes != n * (size_t)e) return set_err(s, IFL_ERR_ARG, "ifl_upload: size mismatch");
    IFL_HIP(hipMemcpyAsync(p, host, bytes, hipMemcpyHostToDevice, s->stream));
    if (field == IFL_FIELD_Q(IFL_Q_D, IFL_QF_CELL)) IFL_TRY(k_update_dflag(s));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_download(ifl_solver* h, int field, void* host, size_t bytes) {
    if (!h || !host) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipSetDevice(s->device));
    if (skewed_field(field)) {
        size_t n = (size_t)s->W * s->H;
        if (bytes != n * 8) return set_err(s, IFL_ERR_ARG, "ifl_download: size mismatch");
        IFL_TRY(k_skew_to_rows(s, skew_ptr(s, field), s->stage_buf));
        IFL_HIP(hipMemcpyAsync(host, s->stage_buf, bytes, hipMemcpyDeviceToHost, s->stream));
        IFL_HIP(hipStreamSynchronize(s->stream));
        return IFL_OK;
    }
    size_t n;
    int e;
    void* p = field_ptr(s, field, &n, &e);
    if (!p) return set_err(s, IFL_ERR_ARG, "ifl_download: no such field at this level");
    if (bytes != n * (size_t)e) return set_err(s, IFL_ERR_ARG, "ifl_download: size mismatch");
    IFL_HIP(hipMemcpyAsync(host, p, bytes, hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_get_particle_count(const ifl_solver* h, int* count) {
    if (!h || !count) return IFL_ERR_ARG;
    *count = h->s.particle_count;
    return h->s.level == 8 ? IFL_OK : IFL_ERR_UNSUPPORTED;
}

int ifl_set_particle_count(ifl_solver* h, int count) {
    if (!h) return IFL_ERR_ARG;
    Solver* s = &h->s;
    if (s->level != 8) return IFL_ERR_UNSUPPORTED;
    if (count < 0 || count > s->max_particles) return set_err(s, IFL_ERR_ARG, "particle count out of range");
    s->particle_count = count;
    s->particles_ready = true;  // the caller now owns the particle state
    return IFL_OK;
}

int ifl_init_particles(ifl_solver* h) {
    if (!h) return IFL_ERR_ARG;
    Solver* s = &h->s;
    if (s->level != 8) return IFL_ERR_UNSUPPORTED;
    if (s->particles_ready) return set_err(s, IFL_ERR_ARG, "ifl_init_particles: particles already initialised");
    IFL_HIP(hipSetDevice(s->device));
    IFL_TRY(flip_init_particles(s));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_to_image(ifl_solver* h, uint8_t* rgba, size_t bytes, int render_heat) {
    if (!h || !rgba) return IFL_ERR_ARG;
    Solver* s = &h->s;
    const int heat = render_heat && s->level >= 6;
    const size_t need = (size_t)s->W * s->H * 4 * (heat ? 2 : 1);
    if (bytes != need) return set_err(s, IFL_ERR_ARG, "ifl_to_image: buffer must hold w*h*4 bytes (w*h*8 with heat)");
    IFL_HIP(hipSetDevice(s->device));
    if (s->image_bytes < need) {
        if (s->image_d) IFL_HIP(hipFree(s->image_d));
        s->image_d = nullptr;
        s->image_bytes = 0;
        IFL_HIP(hipMalloc((void**)&s->image_d, need));
        s->image_bytes = need;
    }
    IFL_TRY(k_to_image(s, s->image_d, heat, 0, s->H));
    IFL_HIP(hipMemcpyAsync(rgba, s->image_d, need, hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_to_image_static(ifl_solver* h, uint8_t* rgba, size_t bytes, int render_heat) {
    if (!h || !rgba) return IFL_ERR_ARG;
    Solver* s = &h->s;
    const int heat = render_heat && s->level >= 6;
    const size_t row = (size_t)s->W * 4 * (heat ? 2 : 1);
    if (bytes != row * s->H)
        return set_err(s, IFL_ERR_ARG, "ifl_to_image_static: buffer must hold w*h*4 bytes (w*h*8 with heat)");
    if (row > kScratchBytes)
        return set_err(s, IFL_ERR_ARG, "ifl_to_image_static: one image row exceeds the 4 MiB device scratch "
                                       "(w at most 1048576, 524288 with heat)");
    IFL_HIP(hipSetDevice(s->device));
    uint8_t* ptr = (uint8_t*)scratch_ptr(s->device);
    if (!ptr) return set_err(s, IFL_ERR_HIP, "ifl_to_image_static: device scratch not set up");
    // every byte of a window is written by the kernel: nothing to clear
    const int rows = (int)std::min<size_t>(kScratchBytes / row, (size_t)s->H);
    for (int y0 = 0; y0 < s->H; y0 += rows) {
        const int n = std::min(rows, s->H - y0);
        IFL_TRY(k_to_image(s, ptr, heat, y0, n));
        IFL_HIP(hipMemcpyAsync(rgba + (size_t)y0 * row, ptr, (size_t)n * row, hipMemcpyDeviceToHost, s->stream));
        IFL_HIP(hipStreamSynchronize(s->stream));  // the next window reuses the scratch
    }
    return IFL_OK;
}

int ifl_write_png(const char* path, const uint8_t* rgba, int w, int h) {
    if (!path || !rgba || w <= 0 || h <= 0) return IFL_ERR_ARG;
    std::vector<uint8_t> png;
    if (!encode_png(rgba, w, h, png)) return IFL_ERR_ARG;
    FILE* f = fopen(path, "wb");
    if (!f) {
        g_err = std::string("ifl_write_png: cannot open ") + path;
        return IFL_ERR_ARG;
    }
    const bool ok = fwrite(png.data(), 1, png.size(), f) == png.size();
    if (fclose(f) != 0 || !ok) {
        g_err = std::string("ifl_write_png: write failed for ") + path;
        return IFL_ERR_ARG;
    }
    return IFL_OK;
}

int ifl_batch_create(ifl_solver* const* handles, int n, ifl_batch** out) {
    if (!handles || !out || n < 1) return IFL_ERR_ARG;
    *out = nullptr;
    const Solver* s0 = handles[0] ? &handles[0]->s : nullptr;
    if (!s0) return IFL_ERR_ARG;
    for (int i = 0; i < n; i++) {
        const Solver* s = handles[i] ? &handles[i]->s : nullptr;
        if (!s) return IFL_ERR_ARG;
        if (s->level != s0->level || s->device != s0->device ||
            s->cfg.precision != IFL_PREC_FP64 || s->cfg.dot_order != IFL_DOT_TREE) {
            g_err = "ifl_batch_create: solvers of one level on one device, FP64 form, tree-ordered dots";
            return IFL_ERR_UNSUPPORTED;
        }
        for (int k = 0; k < i; k++)
            if (handles[k] == handles[i]) {
                g_err = "ifl_batch_create: a solver appears twice";
                return IFL_ERR_ARG;
            }
    }
    ifl_batch* b = new (std::nothrow) ifl_batch();
    if (!b) return IFL_ERR_ALLOC;
    b->handles.assign(handles, handles + n);
    for (int i = 0; i < n; i++) b->solvers.push_back(&handles[i]->s);
    b->device = s0->device;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        g_err = "ifl_batch_create: stream creation failed";
        delete b;
        return IFL_ERR_HIP;
    }
    *out = b;
    return IFL_OK;
}

int ifl_batch_update(ifl_batch* b, double dt, ifl_step_stats* stats) {
    if (!b) return IFL_ERR_ARG;
    const int n = (int)b->solvers.size();
    if (hipSetDevice(b->device) != hipSuccess) return IFL_ERR_HIP;
    BusyGuard busy(b->device);
    // The per-solver stages run on each solver's own stream, so that the
    // small kernels of different solvers overlap; the batched iterations run
    // on the batch's stream, joined to the solver streams by events
    // (k_project_batch).
    const ifl_solve_stats none = {IFL_SOLVE_NOT_RUN, 0, 0.0};
    std::vector<ifl_solve_stats> heat(n, none), press(n, none);
    int rc = IFL_OK, after = SOLVE_NONE;
    // buildPreconditioner runs batched inside k_project_batch
    struct Batched {
        std::vector<Solver*>& v;
        explicit Batched(std::vector<Solver*>& x) : v(x) { for (Solver* s : v) s->batched = true; }
        ~Batched() { for (Solver* s : v) s->batched = false; }
    } batched(b->solvers);
    // Levels 3-7 with bodies evaluated on the device run every stage as one
    // launch for the whole batch on its stream; FLIP (particles) and
    // host-evaluated bodies run their stages per solver on the solvers' own
    // streams, joined to the batched solves by events.
    bool staged = b->solvers[0]->level <= 7;
    for (Solver* s : b->solvers) staged = staged && !s->host_geom;
    // Level 8 once every solver has its particles (creation draws frand()
    // in step order): particlesToGrid runs for every solver up to its first
    // host wait, then the batch reads the prunes' counts and seeds the
    // solvers in batch order (the draws of the same solvers stepped one
    // after the other).
    struct Split {
        std::vector<Solver*>& v;
        explicit Split(std::vector<Solver*>& x, bool on) : v(x) { for (Solver* s : v) s->flip_split = on; }
        ~Split() { for (Solver* s : v) s->flip_split = false; }
    } split(b->solvers, [&] {
        bool on = b->solvers[0]->level == 8;
        for (Solver* s : b->solvers) on = on && s->particles_ready;
        return on;
    }());
    const bool flip_split = b->solvers[0]->flip_split;
    if (staged) {
        rc = kb_begin(b->solvers.data(), n, b->stream, &b->sdev);
        if (rc != IFL_OK) g_err = b->solvers[0]->err;
    }
    while (rc == IFL_OK) {
        int next = SOLVE_NONE;
        if (staged) {
            rc = update_part_batch(b, after, dt, &next);
            if (rc != IFL_OK) g_err = b->solvers[0]->err;
        } else {
            for (int i = 0; i < n && rc == IFL_OK; i++) {
                int ni = SOLVE_NONE;
                rc = update_part(b->solvers[i], after, dt, &ni);
                if (rc != IFL_OK) g_err = b->solvers[i]->err;
                next = ni;  // every solver of one level takes the same sequence of solves
            }
        }
        if (rc != IFL_OK || next == SOLVE_NONE) break;
        if (next == STAGE_PRUNED) {
            for (int i = 0; i < n && rc == IFL_OK; i++) {
                rc = k_prune_finish(b->solvers[i]);
                if (rc != IFL_OK) g_err = b->solvers[i]->err;
            }
            for (int i = 0; i < n && rc == IFL_OK; i++) {
                rc = k_seed(b->solvers[i]);
                if (rc != IFL_OK) g_err = b->solvers[i]->err;
            }
            after = next;
            continue;
        }
        if (next == STAGE_EXTRAPOLATE)
            rc = k_extrapolate_batch(b->solvers.data(), n, b->stream, &b->xdev,
                                     (1 << IFL_Q_D) | (1 << IFL_Q_U) | (1 << IFL_Q_V), !staged);
        else
            rc = k_project_batch(b->solvers.data(), n, b->stream, &b->dev,
                                 next == SOLVE_HEAT ? heat.data() : press.data(), b->profiling ? &b->prof : nullptr,
                                 !staged);
        if (rc != IFL_OK) g_err = b->solvers[0]->err;
        after = next;
    }
    if (staged && rc == IFL_OK) {
        std::vector<int> nonfinite(n), border(n);
        rc = kb_finish(b->sdev, b->solvers[0]->level >= 4, nonfinite.data(), border.data());
        if (rc != IFL_OK) g_err = b->solvers[0]->err;
        for (int i = 0; i < n && rc == IFL_OK; i++) {
            Solver* s = b->solvers[i];
            if (s->border_pending) s->border_solid |= border[i];
            s->border_pending = false;
            rc = prof_collect(s, nullptr);
            if (rc != IFL_OK) {
                g_err = s->err;
                break;
            }
            if (stats) {
                stats[i].nonfinite = nonfinite[i];
                stats[i].reserved = 0;
                stats[i].heat = heat[i];
                stats[i].pressure = press[i];
                stats[i].particle_count = s->particle_count;
                stats[i].border_solid = s->border_solid;
            }
        }
    }
    for (int i = 0; i < n && rc == IFL_OK && flip_split; i++) {
        rc = flip_check_overflow(b->solvers[i]);
        if (rc != IFL_OK) g_err = b->solvers[i]->err;
    }
    for (int i = 0; i < n && rc == IFL_OK && !staged; i++) {
        rc = finish_step(b->solvers[i], heat[i], press[i], stats ? &stats[i] : nullptr);
        if (rc != IFL_OK) g_err = b->solvers[i]->err;
    }
    (void)hipStreamSynchronize(b->stream);
    return rc;
}

int ifl_batch_add_inflow(ifl_batch* b, int count, const int* inst, const double* rect, const double* vals) {
    if (!b || count < 0 || (count > 0 && (!inst || !rect || !vals))) return IFL_ERR_ARG;
    const int n = (int)b->solvers.size();
    for (int k = 0; k < count; k++)
        if (inst[k] < 0 || inst[k] >= n) {
            g_err = "ifl_batch_add_inflow: solver index out of range";
            return IFL_ERR_ARG;
        }
    if (count == 0) return IFL_OK;
    if (hipSetDevice(b->device) != hipSuccess) return IFL_ERR_HIP;
    // work queued on the solvers' own streams comes first
    for (Solver* s : b->solvers) IFL_HIP(hipStreamSynchronize(s->stream));
    int rc = k_add_inflow_batch(b->solvers.data(), n, b->stream, &b->idev, count, inst, rect, vals);
    if (rc != IFL_OK) g_err = b->solvers[0]->err;
    return rc;
}

int ifl_batch_profile(ifl_batch* b, int on) {
    if (!b) return IFL_ERR_ARG;
    b->profiling = on != 0;
    b->prof = ifl::BatchProfile();
    return IFL_OK;
}

int ifl_batch_profile_read(const ifl_batch* b, int kernel, int64_t* launches, double* ms, double* cell_iterations) {
    if (!b || kernel < 0 || kernel >= IFL_KERNEL_COUNT) return IFL_ERR_ARG;
    if (launches) *launches = b->prof.launches[kernel];
    if (ms) *ms = b->prof.ms[kernel];
    if (cell_iterations) *cell_iterations = b->prof.cell_iterations[kernel];
    return IFL_OK;
}

int ifl_batch_destroy(ifl_batch* b) {
    if (!b) return IFL_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    batch_dev_free(b->dev);
    extrap_dev_free(b->xdev);
    stage_dev_free(b->sdev);
    inflow_dev_free(b->idev);
    delete b;
    return IFL_OK;
}

int ifl_max_timestep(ifl_solver* h, double* dt) {
    if (!h || !dt) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipSetDevice(s->device));
    return k_max_timestep(s, dt);
}

uint32_t ifl_frand_state(void) { return frand_state(); }

void ifl_frand_set_state(uint32_t state) { frand_set_state(state); }

int ifl_stage(ifl_solver* h, int stage, int iarg, double darg, ifl_solve_stats* stats, double* result) {
    if (!h) return IFL_ERR_ARG;
    Solver* s = &h->s;
    int lv = s->level;
    IFL_HIP(hipSetDevice(s->device));
    BusyGuard busy(s->device);
    bool qok = iarg >= 0 && iarg < 4 && !(iarg == IFL_Q_T && lv < 6);
    int rc = IFL_OK;
    switch (stage) {
    case IFL_STAGE_FILL_SOLID:
        if (lv < 4) return IFL_ERR_UNSUPPORTED;
        if (!qok) return IFL_ERR_ARG;
        rc = fill_solid(s, iarg);
        break;
    case IFL_STAGE_EXTRAPOLATE:
        if (lv < 4) return IFL_ERR_UNSUPPORTED;
        if (!qok) return IFL_ERR_ARG;
        rc = extrapolate(s, iarg);
        break;
    case IFL_STAGE_ADVECT:
        if (lv == 8) return IFL_ERR_UNSUPPORTED;
        if (!qok) return IFL_ERR_ARG;
        rc = k_advect(s, 1 << iarg, darg);
        break;
    case IFL_STAGE_FLIP:
        if (lv == 8) return IFL_ERR_UNSUPPORTED;
        if (!qok) return IFL_ERR_ARG;
        flip_q(s, iarg);
        break;
    case IFL_STAGE_BUILD_RHS: rc = k_build_rhs(s); break;
    case IFL_STAGE_BUILD_PRESSURE: rc = k_build_pressure(s, darg); break;
    case IFL_STAGE_BUILD_HEAT:
        if (lv < 6) return IFL_ERR_UNSUPPORTED;
        rc = k_build_heat(s, darg);
        break;
    case IFL_STAGE_BUILD_PRECON: rc = k_build_precon(s); break;
    case IFL_STAGE_APPLY_PRECON: rc = k_apply_precon(s); break;
    case IFL_STAGE_MATVEC: rc = k_matvec(s); break;
    case IFL_STAGE_DOT: {
        double d = 0;
        rc = k_dot(s, iarg, &d);
        if (result) *result = d;
        break;
    }
    case IFL_STAGE_INF_NORM: {
        double d = 0;
        rc = k_inf_norm(s, &d);
        if (result) *result = d;
        break;
    }
    case IFL_STAGE_PROJECT: {
        ifl_solve_stats st;
        rc = k_project(s, iarg, &st);
        if (stats) *stats = st;
        if (result) *result = st.iterations;
        break;
    }
    case IFL_STAGE_APPLY_PRESSURE: rc = k_apply_pressure(s, darg); break;
    case IFL_STAGE_ADD_BUOYANCY:
        if (lv < 6) return IFL_ERR_UNSUPPORTED;
        rc = k_add_buoyancy(s, darg);
        break;
    case IFL_STAGE_SET_BOUNDARY:
        if (lv < 4) return IFL_ERR_UNSUPPORTED;
        rc = k_set_boundary(s);
        break;
    case IFL_STAGE_COMPUTE_DENSITIES:
        if (lv < 7) return IFL_ERR_UNSUPPORTED;
        rc = k_compute_densities(s);
        break;
    case IFL_STAGE_COPY_T_TO_R:
        if (lv < 6) return IFL_ERR_UNSUPPORTED;
        rc = k_copy_t_to_r(s);
        break;
    case IFL_STAGE_COPY_P_TO_T:
        if (lv < 6) return IFL_ERR_UNSUPPORTED;
        rc = k_copy_p_to_t(s);
        break;
    case IFL_STAGE_FROM_PARTICLES:
    case IFL_STAGE_COPY_OLD:
    case IFL_STAGE_DIFF:
    case IFL_STAGE_UNDIFF:
        if (lv != 8) return IFL_ERR_UNSUPPORTED;
        if (!qok) return IFL_ERR_ARG;
        rc = stage == IFL_STAGE_FROM_PARTICLES ? k_from_particles(s, iarg, false)
             : stage == IFL_STAGE_COPY_OLD     ? k_copy_old(s, iarg)
                                               : k_diff(s, iarg, darg, stage == IFL_STAGE_UNDIFF);
        break;
    case IFL_STAGE_G2P:
    case IFL_STAGE_P2G:
    case IFL_STAGE_PARTICLE_ADVECT:
    case IFL_STAGE_COUNT_PARTICLES:
    case IFL_STAGE_PRUNE:
    case IFL_STAGE_SEED:
        if (lv != 8) return IFL_ERR_UNSUPPORTED;
        rc = stage == IFL_STAGE_G2P               ? k_grid_to_particles(s, darg)
             : stage == IFL_STAGE_P2G             ? k_particles_to_grid(s)
             : stage == IFL_STAGE_PARTICLE_ADVECT ? k_particle_advect(s, darg)
             : stage == IFL_STAGE_COUNT_PARTICLES ? k_count_particles(s)
             : stage == IFL_STAGE_PRUNE           ? k_prune(s)
                                                  : k_seed(s);
        break;
    default:
        return set_err(s, IFL_ERR_UNSUPPORTED, "ifl_stage: unknown stage or not at this level");
    }
    if (rc != IFL_OK) return rc;
    if (lv == 8) {
        // EMPTY marks on the d grid change the cells the pressure solve visits
        if (stage == IFL_STAGE_FROM_PARTICLES || stage == IFL_STAGE_EXTRAPOLATE || stage == IFL_STAGE_P2G)
            IFL_TRY(k_update_dflag(s));
        if (stage == IFL_STAGE_EXTRAPOLATE || stage == IFL_STAGE_P2G) IFL_TRY(flip_check_overflow(s));
    }
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int ifl_profile_enable(ifl_solver* h, int on) {
    if (!h) return IFL_ERR_ARG;
    h->s.profiling = on != 0;
    return IFL_OK;
}

int ifl_profile_reset(ifl_solver* h) {
    if (!h) return IFL_ERR_ARG;
    for (auto& p : h->s.prof) p = Profile();
    return IFL_OK;
}

int ifl_profile_read(const ifl_solver* h, int kernel, int64_t* launches, double* total_ms, int64_t* iterations,
                     double* cell_iterations) {
    if (!h || kernel < 0 || kernel >= IFL_KERNEL_COUNT) return IFL_ERR_ARG;
    const Profile& p = h->s.prof[kernel];
    if (launches) *launches = p.launches;
    if (total_ms) *total_ms = p.ms;
    if (iterations) *iterations = p.iterations;
    if (cell_iterations) *cell_iterations = p.cell_iterations;
    return IFL_OK;
}

int ifl_debug_trace_sweep(ifl_solver* h, int kind, int64_t* host_out, size_t count, int* blocks_per_strip) {
    if (!h || kind < 0 || kind > 2 || !host_out) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipSetDevice(s->device));
    int nblk = s->sk.cols / 8;
    size_t need = (size_t)s->sk.nstrips * nblk * kTraceStamps;
    if (blocks_per_strip) *blocks_per_strip = nblk;
    if (count < need) return set_err(s, IFL_ERR_ARG, "trace buffer too small");
    long long* d = nullptr;
    IFL_HIP(hipMalloc((void**)&d, need * sizeof(long long)));
    IFL_HIP(hipMemsetAsync(d, 0, need * sizeof(long long), s->stream));
    int rc = k_trace_sweep(s, kind, d);
    if (rc == IFL_OK) {
        IFL_HIP(hipMemcpyAsync(host_out, d, need * sizeof(long long), hipMemcpyDeviceToHost, s->stream));
        IFL_HIP(hipStreamSynchronize(s->stream));
    }
    (void)hipFree(d);
    return rc;
}

int ifl_synchronize(ifl_solver* h) {
    if (!h) return IFL_ERR_ARG;
    Solver* s = &h->s;
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

}  // extern "C"
