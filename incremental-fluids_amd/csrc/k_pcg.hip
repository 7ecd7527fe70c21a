// MIC(0)-preconditioned conjugate gradients, FluidSolver::project (3:349-380,
// 7:853-886), on the strip-skewed layout (Skew in ifl_common.hpp).
//
// Sweeps (buildPreconditioner 7:739-768, applyPreconditioner 7:770-804): one
// wave per strip of 64 rows, one lane per row. Lane l runs `x = t - l` at
// step t, so the cells of one anti-diagonal advance together; the value a
// cell passes to the row above moves by one lane through DPP (wave_shr:1)
// and, across strips, through an edge buffer in global memory that the
// consumer polls (a signalling-NaN sentinel marks "not written yet"; no
// arithmetic result can equal it). Every cell evaluates the reference's
// expression with the reference's operand order, so the order of cells does
// not change a single bit (each cell's inputs are final when it runs).
//
// Terms the reference skips (a non-FLUID or missing neighbour) enter as an
// exact +0.0 subtrahend, which leaves every value, -0.0 included, unchanged.
//
// One PCG iteration is split into phases around its global reductions:
//   A1  p += alpha*s, s = z + beta*s           (deferred from the previous iteration)
//   A2  q = A s, partial sums of q.s           (matrixVectorProduct + dotProduct)
//   B   r += q*(-alpha), max|r|, forward sweep (scaledAdd + infinityNorm + L solve)
//   C   backward sweep, partial sums of z.r    (U solve + dotProduct)
// with small single-workgroup kernels finishing each reduction in a fixed
// order (the sums are deterministic run to run; they differ from the
// reference's left-to-right sum only by rounding, unless dot_order selects
// the reference order).
#include <algorithm>
#include <atomic>
#include <type_traits>
#include <vector>

#include "ifl_solver.hpp"
#include "k_sweep.hpp"

namespace ifl {

// Strips grouped by XCD (order 0 of strip_of_block); the upstream-first
// kernels are k_sweep_up.hip's.
template <int KIND, int MODE, int LV, int NG, typename T, bool TR>
__global__ __launch_bounds__(64 * kWaves) void stream_sweep_kernel(StreamArgs a) {
    const int j = strip_of_block(blockIdx.x, a.nlaunch, KIND == SW_BWD, 0);
    if (j < 0) return;
    sweep_body<KIND, MODE, LV, NG, T, TR>(a, j);
}

// ---------------------------------------------------------------- vector phases
struct VecArgs {
    Skew sk;
    double *p, *s, *z, *r, *q;
    const double *adiag, *apx, *apy;
    const uint8_t* flag;
    double* partial;
    PcgState* st;
    int pm;  // levels 4+: the preconditioner writes _z at FLUID cells only
    int vm;  // levels 6+: vector operations are masked to FLUID cells
    int defl;  // the mixed form's pressure solves: s is formed from z less its mean over FLUID cells
};

IFL_DEV bool skew_cell(const Skew& sk, long i, int& x, int& y) {
    long strip = i / sk.strip_sz;
    long rem = i - strip * sk.strip_sz;
    int t = (int)(rem >> 6), l = (int)(rem & 63);
    x = t - l;
    y = (int)strip * kStrip + l;
    return x >= 0 && x < sk.W && y < sk.H;
}

// A1: the tail of the previous iteration, scaledAdd(_p, _p, _s, alpha) 7:869
// and scaledAdd(_s, _z, _s, sigmaNew/sigma) 7:881. mode 1 runs only the s
// update (the budget-exhausted tail), mode 2 only the p update.
// The reference's _z at a non-FLUID cell is still A s from the matrix-vector
// product (applyPreconditioner skips the cell, 4:766-796); this product keeps
// A s in q, so levels 4-5 (unmasked scaledAdd) read it from there.
IFL_DEV void vec_update_body(const VecArgs& a, long blk, int mode, int check_done) {
    PcgState* st = a.st;
    if (check_done && st->done) return;
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blk * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x, y;
    if (!skew_cell(a.sk, i, x, y)) return;
    // every load first, then the selects
    const bool fl = a.flag[i] != 0;
    const double s = a.s[i];
    const double p = mode != 1 ? a.p[i] : 0.0;
    const double zv = mode != 2 ? a.z[i] : 0.0, qv = mode != 2 && a.pm ? a.q[i] : 0.0;
    if (a.vm && !fl) return;
    if (mode != 1) a.p[i] = p + s * st->alpha;
    if (mode != 2) {
        double zi = a.pm && !fl ? qv : zv;
        if (a.defl && fl) zi = zi - st->zmean;
        a.s[i] = zi + s * st->beta;
    }
}
__global__ void vec_update_kernel(VecArgs a, int mode, int check_done) { vec_update_body(a, blockIdx.x, mode, check_done); }

// End of a solve that ran out of its budget: the reference's _z holds A s
// at every non-FLUID cell (levels 4+), the preconditioned residual elsewhere.
__global__ void z_nonfluid_kernel(VecArgs a) {
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x, y;
    if (!skew_cell(a.sk, i, x, y) || a.flag[i]) return;
    a.z[i] = a.q[i];
}

// A2: matrixVectorProduct 7:816-833 (unmasked at every level; the four
// neighbour terms in the reference's order, skipped at the domain border)
// fused with the block partial of dotProduct(q, s).
template <int OUT_Q>
IFL_DEV void matvec_body(const VecArgs& a, long blk, int check_done) {
    __shared__ double red[4];
    PcgState* st = a.st;
    if (check_done && st->done) return;
    const Skew& sk = a.sk;
    long n = (long)sk.nstrips * sk.strip_sz;
    long i = blk * (long)blockDim.x + threadIdx.x;
    double prod = 0.0;
    int x, y;
    if (i < n && skew_cell(sk, i, x, y)) {
        const int l = y & 63;
        const int W = sk.W, H = sk.H;
        const double* b = a.s;
        // every load first (a missing neighbour reads cell i and is not added)
        const bool hw = x > 0, hs = y > 0, he = x < W - 1, hn = y < H - 1;
        const long jw = hw ? i - kStrip : i;
        const long js = hs ? (l > 0 ? i - kStrip - 1 : sk.pos(x, y - 1)) : i;
        const long je = he ? i + kStrip : i;
        const long jn = hn ? (l < kStrip - 1 ? i + kStrip + 1 : sk.pos(x, y + 1)) : i;
        const double b0 = b[i], bw = b[jw], bs = b[js], be = b[je], bn = b[jn];
        const double ad = a.adiag[i], axw = a.apx[jw], ays = a.apy[js], ax = a.apx[i], ay = a.apy[i];
        const bool on = !a.vm || a.flag[i];
        double t = ad * b0;
        t = hw ? t + axw * bw : t;
        t = hs ? t + ays * bs : t;
        t = he ? t + ax * be : t;
        t = hn ? t + ay * bn : t;
        if (OUT_Q)
            a.q[i] = t;
        else
            a.z[i] = t;
        if (on) prod = t * b0;
    }
    prod = wave_sum(prod);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blk] = (red[0] + red[1]) + (red[2] + red[3]);
}
template <int OUT_Q>
__global__ __launch_bounds__(256) void matvec_kernel(VecArgs a, int check_done) {
    matvec_body<OUT_Q>(a, blockIdx.x, check_done);
}

// A1 and A2 in one pass, for vector operations masked to FLUID cells (levels
// 6-8) in the FP64 form: scaledAdd(_p, _p, _s, alpha) 7:869 and
// scaledAdd(_s, _z, _s, beta) 7:881 of the previous iteration, then
// matrixVectorProduct(_z, _s) 7:868 and the block partial of its dot
// product. Each product term takes its neighbour's new s as that update
// forms it (z + s beta at a FLUID cell, s unchanged elsewhere), so the terms
// are vec_update_kernel's and matvec_kernel's values; the new s goes to a
// second array, since neighbouring workgroups still read the old one.
IFL_DEV void matvec_update_body(const VecArgs& a, long blk, const double* __restrict__ so, double* __restrict__ sn) {
    __shared__ double red[4];
    PcgState* st = a.st;
    if (st->done) return;
    const Skew& sk = a.sk;
    const long n = (long)sk.nstrips * sk.strip_sz;
    const long i = blk * (long)blockDim.x + threadIdx.x;
    const double alpha = st->alpha, beta = st->beta;
    double prod = 0.0;
    int x, y;
    if (i < n && skew_cell(sk, i, x, y)) {
        const int l = y & 63;
        const int W = sk.W, H = sk.H;
        // every load first (a missing neighbour reads cell i and is not
        // added), so the loads overlap instead of waiting on the flags
        const bool hw = x > 0, hs = y > 0, he = x < W - 1, hn = y < H - 1;
        const long jw = hw ? i - kStrip : i;
        const long js = hs ? (l > 0 ? i - kStrip - 1 : sk.pos(x, y - 1)) : i;
        const long je = he ? i + kStrip : i;
        const long jn = hn ? (l < kStrip - 1 ? i + kStrip + 1 : sk.pos(x, y + 1)) : i;
        const bool fl = a.flag[i] != 0, fw = a.flag[jw] != 0, fs = a.flag[js] != 0, fe = a.flag[je] != 0,
                   fn = a.flag[jn] != 0;
        const double s0 = so[i], sw = so[jw], ss = so[js], se = so[je], snn = so[jn];
        const double z0 = a.z[i], zw = a.z[jw], zs = a.z[js], ze = a.z[je], zn = a.z[jn];
        const double ad = a.adiag[i], axw = a.apx[jw], ays = a.apy[js], ax = a.apx[i], ay = a.apy[i];
        const double p0 = a.p[i];
        // the new s (z + s beta at a FLUID cell, s elsewhere)
        const double b = fl ? z0 + s0 * beta : s0;
        const double bw = fw ? zw + sw * beta : sw;
        const double bs = fs ? zs + ss * beta : ss;
        const double be = fe ? ze + se * beta : se;
        const double bn = fn ? zn + snn * beta : snn;
        if (fl) a.p[i] = p0 + s0 * alpha;
        sn[i] = b;
        double t = ad * b;
        t = hw ? t + axw * bw : t;
        t = hs ? t + ays * bs : t;
        t = he ? t + ax * be : t;
        t = hn ? t + ay * bn : t;
        a.q[i] = t;
        if (fl) prod = t * b;
    }
    prod = wave_sum(prod);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blk] = (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ __launch_bounds__(256) void matvec_update_kernel(VecArgs a, const double* __restrict__ so,
                                                            double* __restrict__ sn) {
    matvec_update_body(a, blockIdx.x, so, sn);
}

// Fixed-order sum of `n` partials by one workgroup of 256 threads.
IFL_DEV double block_sum_partials(const double* partial, int n) {
    __shared__ double red[256];
    // thread t sums partials t, t + 256, ... in that order; 32 loads are in
    // flight at a time (the adds keep their order)
    constexpr int kInFlight = 32;
    double acc = 0.0;
    int k = threadIdx.x;
    for (; k + (kInFlight - 1) * 256 < n; k += kInFlight * 256) {
        double v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; u++) v[u] = partial[k + u * 256];
#pragma unroll
        for (int u = 0; u < kInFlight; u++) acc += v[u];
    }
    for (; k + 7 * 256 < n; k += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = partial[k + u * 256];
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
    }
    for (; k < n; k += 256) acc += partial[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    return red[0];
}

// dotProduct in the reference's left-to-right order (dot_order = SEQUENTIAL).
// The reference's left-to-right sum (7:806-814, masked from level 6 on) in
// row-major cell order, by one workgroup of 256 threads: waves 1-3 form the
// products of the next chunk of cells into LDS while wave 0 adds the current
// chunk in order, two products per LDS read (every lane of wave 0 holds the
// same running sum). A cell the reference skips contributes +0.0, which
// leaves the sum unchanged bit for bit: the sum starts at +0.0 and a
// round-to-nearest sum is never -0.0 unless both addends are, so x + (+0.0)
// is x for every running sum x. a_nf, when given, replaces a at non-FLUID
// cells. Every thread of the block must call it; the result is valid in
// wave 0.
constexpr int kSeqPer = 8;                  // products per producer thread per chunk
constexpr int kSeqChunk = 192 * kSeqPer;    // cells per chunk
IFL_DEV double seq_dot(const Skew& sk, const double* a, const double* b, const uint8_t* flag, int vm,
                       const double* a_nf = nullptr) {
    __shared__ __align__(16) double buf[2][kSeqChunk];
    const long n = (long)sk.W * sk.H;
    const int tid = threadIdx.x;
    const bool producer = tid >= 64;
    auto produce = [&](long base, double* dst) {
        const int t = tid - 64;
#pragma unroll
        for (int k = 0; k < kSeqPer; k++) {
            const int j = t + 192 * k;
            const long i = base + j;
            double pr = 0.0;
            if (i < n) {
                const int x = (int)(i % sk.W), y = (int)(i / sk.W);
                const long p = sk.pos(x, y);
                if (!vm || flag[p]) pr = (a_nf && !flag[p] ? a_nf[p] : a[p]) * b[p];
            }
            dst[j] = pr;
        }
    };
    double result = 0.0;
    const long nchunks = (n + kSeqChunk - 1) / kSeqChunk;
    if (producer) produce(0, buf[0]);
    __syncthreads();
    for (long c = 0; c < nchunks; c++) {
        if (producer) {
            if (c + 1 < nchunks) produce((c + 1) * kSeqChunk, buf[(c + 1) & 1]);
        } else {
            const vec2d* src = (const vec2d*)buf[c & 1];
            // cells past the end were formed as +0.0: summing the whole chunk
            // changes nothing
#pragma unroll 8
            for (int k = 0; k < kSeqChunk / 2; k++) {
                const vec2d v = src[k];
                result += v.x;
                result += v.y;
            }
        }
        __syncthreads();
    }
    return result;
}

// finish A2: alpha = sigma / dot(q, s)  (7:868)
IFL_DEV void fin_alpha_body(const VecArgs& a, int n, int seq) {
    PcgState* st = a.st;
    if (st->done) return;
    double d;
    if (seq) {
        const double t = seq_dot(a.sk, a.q, a.s, a.flag, a.vm);
        if (threadIdx.x == 0) st->dot_tmp = t;
        __syncthreads();
        d = st->dot_tmp;
    } else {
        d = block_sum_partials(a.partial, n);
    }
    if (threadIdx.x == 0) {
        st->alpha = st->sigma / d;
        st->max_err = 0.0;
    }
}
__global__ __launch_bounds__(256) void fin_alpha_kernel(VecArgs a, int n, int seq) { fin_alpha_body(a, n, seq); }

// finish B: the convergence test of 7:872-876
IFL_DEV void fin_check_body(PcgState* st, double tol) {
    if (st->done) return;
    if (st->max_err < tol) {
        st->done = 1;
        st->status = IFL_SOLVE_CONVERGED;
        st->printed_iter = st->iter;
    }
}
__global__ void fin_check_kernel(PcgState* st, double tol) { fin_check_body(st, tol); }

// finish C: sigmaNew = dot(z, r), beta = sigmaNew/sigma (7:880-882).
// init=1 is the entry of project(): sigma = dot(z, r) and the early-exit
// test of 7:858-862.
// The convergence test of 7:872-876 where the PCG loop has no fin_check
// launch: the backward sweep returns at once when it holds (as it does once
// fin_check has set `done`), and the finish after it records the exit here;
// `iter` has not moved since the test's place in the loop.
IFL_DEV bool fin_converged(PcgState* st, double tol) {
    if (!(st->max_err < tol)) return false;
    if (threadIdx.x == 0) fin_check_body(st, tol);
    return true;
}

IFL_DEV void fin_sigma_body(const VecArgs& a, int n, int seq, int init, double tol, int limit) {
    PcgState* st = a.st;
    if (!init && st->done) return;
    if (!init && fin_converged(st, tol)) return;
    double d;
    if (seq) {
        // inside the loop the reference's _z at a non-FLUID cell is A s
        // (levels 4-5 take every cell; this product keeps A s in q)
        const double t = seq_dot(a.sk, a.z, a.r, a.flag, a.vm, !init && a.pm && !a.vm ? a.q : nullptr);
        if (threadIdx.x == 0) st->dot_tmp = t;
        __syncthreads();
        d = st->dot_tmp;
    } else {
        d = block_sum_partials(a.partial, n);
    }
    if (threadIdx.x != 0) return;
    if (init) {
        st->sigma = d;
        st->iter = 0;
        if (st->max_err < tol) {
            st->done = 1;
            st->status = IFL_SOLVE_INITIAL_SMALL;
            st->printed_iter = -1;
        } else if (limit <= 0) {
            st->done = 1;
            st->status = IFL_SOLVE_BUDGET;
            st->printed_iter = limit;
        }
        return;
    }
    st->beta = d / st->sigma;
    st->sigma = d;
    st->iter += 1;
    if (st->iter >= limit) {
        st->done = 1;
        st->status = IFL_SOLVE_BUDGET;
        st->printed_iter = limit;
    }
}
__global__ __launch_bounds__(256) void fin_sigma_kernel(VecArgs a, int n, int seq, int init, double tol, int limit) {
    fin_sigma_body(a, n, seq, init, tol, limit);
}

__global__ void state_reset_kernel(PcgState* st) {
    st->zmean = 0.0;
    st->rmean = 0.0;
    st->sigma = st->alpha = st->beta = 0.0;
    st->max_err = 0.0;
    st->dot_tmp = 0.0;
    st->iter = 0;
    st->done = 0;
    st->status = IFL_SOLVE_NOT_RUN;
    st->printed_iter = 0;
}

__global__ void inf_norm_kernel(VecArgs a) {
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double m = 0.0;
    int x, y;
    if (i < n && skew_cell(a.sk, i, x, y) && (!a.vm || a.flag[i])) m = fabs(a.r[i]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long*)&a.st->max_err, (unsigned long long)as_bits(m));
}

__global__ void dot_partial_kernel(VecArgs a, const double* u, const double* v) {
    __shared__ double red[4];
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double prod = 0.0;
    int x, y;
    if (i < n && skew_cell(a.sk, i, x, y) && (!a.vm || a.flag[i])) prod = u[i] * v[i];
    prod = wave_sum(prod);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void dot_finish_kernel(VecArgs a, const double* u, const double* v, int n, int seq, double* out) {
    double d;
    if (seq) {
        const double t = seq_dot(a.sk, u, v, a.flag, a.vm);
        if (threadIdx.x == 0) *out = t;
        return;
    }
    d = block_sum_partials(a.partial, n);
    if (threadIdx.x == 0) *out = d;
}

// ---------------------------------------------------------------- launchers
static VecArgs vec_args(Solver* s) {
    VecArgs a;
    a.sk = s->sk;
    a.p = s->p; a.s = s->s; a.z = s->z; a.r = s->r; a.q = s->qv;
    a.adiag = s->adiag; a.apx = s->aplusx; a.apy = s->aplusy;
    a.flag = s->dflag;
    a.partial = s->partial;
    a.st = s->st;
    a.pm = s->level >= 4;
    a.vm = s->level >= 6;
    a.defl = s->cfg.precision == IFL_PREC_MIXED && s->solve_kind == SOLVE_PRESSURE;
    return a;
}

static StreamArgs stream_args(Solver* s, int kind) {
    StreamArgs a;
    memset(&a, 0, sizeof a);
    a.trace = nullptr;
    a.sk = s->sk;
    a.fmask = s->fmask;
    a.edge = s->edge + (long)kind * s->sk.nstrips * s->W;
    a.partial = s->partial;
    a.st = s->st;
    a.timeout = &s->st->timeout;
    a.tau = s->cfg.tau;
    a.sigma = s->cfg.sigma;
    return a;
}

template <int KIND, int MODE, int LV, int NG, typename T>
static int ring_setup(Solver* s, StreamArgs& a, size_t* shmem, int max_nb = 12) {
    a.slot_bytes = SweepShape<KIND, MODE, NG, T>::SLOT;
    size_t fixed = sizeof(RingHdr) + kERing * sizeof(double);
    int nb = (int)((s->lds_bytes - fixed) / a.slot_bytes);
    if (nb > max_nb) nb = max_nb;
    if (nb < 3) {
        s->err = "not enough LDS for the sweep ring";
        return IFL_ERR_UNSUPPORTED;
    }
    a.nb = nb;
    *shmem = fixed + (size_t)nb * a.slot_bytes;
    return IFL_OK;
}

template <int KIND, int MODE, int LV, int NG, typename T, bool TR>
static int launch_stream_k(Solver* s, StreamArgs a) {
    size_t shmem = 0;
    IFL_TRY((ring_setup<KIND, MODE, LV, NG, T>(s, a, &shmem)));
    // once per kernel; solvers stepped from several host threads may race
    // here, which only repeats the same call
    static std::atomic<bool> attr_set{false};
    if (!attr_set.load(std::memory_order_acquire)) {
        IFL_HIP(hipFuncSetAttribute((const void*)stream_sweep_kernel<KIND, MODE, LV, NG, T, TR>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
        attr_set.store(true, std::memory_order_release);
    }
    a.nlaunch = s->sk.nstrips;
    a.strip0 = 0;
    // the traced sweeps (diagnostics) always run alone
    if (!TR && (a.nlaunch > s->num_cus || (s->busy && s->busy->load(std::memory_order_relaxed) > 1)))
        return launch_sweep_up<KIND, MODE, LV, NG, T>(s, a, shmem);
    stream_sweep_kernel<KIND, MODE, LV, NG, T, TR><<<sweep_blocks(a.nlaunch, 0), 64 * kWaves, shmem, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

template <int KIND, int MODE, int LV, int NG, typename T>
static int launch_stream(Solver* s, StreamArgs a) {
    if (a.trace) return launch_stream_k<KIND, MODE, LV, NG, T, true>(s, a);
    return launch_stream_k<KIND, MODE, LV, NG, T, false>(s, a);
}

static bool mixed(const Solver* s) { return s->cfg.precision == IFL_PREC_MIXED; }

// forward sweep: the drop-in reads the double factor, the mixed form its
// single-precision copy (and writes y in single precision)
template <int MODE, typename T>
static StreamArgs fwd_args(Solver* s) {
    StreamArgs a = stream_args(s, SW_FWD);
    const bool sp = sizeof(T) == 4;
    a.in[0] = s->r;
    a.in[1] = sp ? (const void*)s->cx32 : s->cx;
    a.in[2] = sp ? (const void*)s->cy32 : s->cy;
    a.in[3] = sp ? (const void*)s->pc32 : s->precon;
    a.in[4] = s->qv;
    a.out[0] = sp ? (void*)s->y32 : s->yv;
    a.out[1] = s->r;
    return a;
}
template <int MODE, typename T>
static int launch_fwd_t(Solver* s, long long* trace) {
    StreamArgs a = fwd_args<MODE, T>(s);
    a.trace = trace;
    constexpr int NG = MODE == 1 ? 5 : 4;
    if (s->level == 3) return launch_stream<SW_FWD, MODE, 3, NG, T>(s, a);
    if (s->level <= 5) return launch_stream<SW_FWD, MODE, 4, NG, T>(s, a);
    return launch_stream<SW_FWD, MODE, 6, NG, T>(s, a);
}
template <int MODE>
static int launch_fwd(Solver* s, long long* trace = nullptr) {
    return mixed(s) ? launch_fwd_t<MODE, float>(s, trace) : launch_fwd_t<MODE, double>(s, trace);
}

template <int MODE, typename T>
static StreamArgs bwd_args(Solver* s) {
    StreamArgs a = stream_args(s, SW_BWD);
    const bool sp = sizeof(T) == 4;
    a.in[0] = sp ? (const void*)s->y32 : s->yv;
    a.in[1] = sp ? (const void*)s->cx32 : s->cx;
    a.in[2] = sp ? (const void*)s->cy32 : s->cy;
    a.in[3] = sp ? (const void*)s->pc32 : s->precon;
    a.in[4] = s->r;
    a.in[5] = MODE == 1 ? s->qv : s->z;  // the reference's _z at non-FLUID cells (levels 4-5 dot)
    a.out[0] = s->z;
    if (MODE == 1) a.sigma = s->cfg.tol;  // the PCG loop's convergence test (sweep_body)
    return a;
}
template <int MODE, typename T>
static int launch_bwd_t(Solver* s, long long* trace) {
    StreamArgs a = bwd_args<MODE, T>(s);
    a.trace = trace;
    if (s->level == 3) return launch_stream<SW_BWD, MODE, 3, 5, T>(s, a);
    if (s->level <= 5) return launch_stream<SW_BWD, MODE, 4, 6, T>(s, a);
    return launch_stream<SW_BWD, MODE, 6, 5, T>(s, a);
}
template <int MODE>
static int launch_bwd(Solver* s, long long* trace = nullptr) {
    return mixed(s) ? launch_bwd_t<MODE, float>(s, trace) : launch_bwd_t<MODE, double>(s, trace);
}

static long skew_n(Solver* s) { return (long)s->sk.nstrips * s->sk.strip_sz; }

static StreamArgs precon_args(Solver* s) {
    StreamArgs a = stream_args(s, SW_PRECON);
    a.in[0] = s->adiag; a.in[1] = s->aplusx; a.in[2] = s->aplusy;
    a.out[0] = s->precon;
    a.out[1] = s->cx;
    a.out[2] = s->cy;
    if (s->cfg.precision != IFL_PREC_FP64) {  // both reduced-precision forms sweep the float copies
        a.out[3] = s->pc32;
        a.out[4] = s->cx32;
        a.out[5] = s->cy32;
    }
    return a;
}

static int launch_precon(Solver* s, long long* trace) {
    StreamArgs a = precon_args(s);
    a.trace = trace;
    return launch_stream<SW_PRECON, 0, 3, 3, double>(s, a);
}

int k_build_precon(Solver* s) {
    IFL_TRY(prof_begin(s, IFL_KERNEL_PRECON, -1));
    IFL_TRY(launch_precon(s, nullptr));
    return prof_end(s);
}

// Diagnostics: one sweep of `kind` on the current state, time-stamped per
// block (100 MHz clock): wave 0 starts the block, its operands for the
// second half are ready, the writer has drained it.
int k_trace_sweep(Solver* s, int kind, long long* dev_trace) {
    if (kind == SW_PRECON) return launch_precon(s, dev_trace);
    if (kind == SW_FWD) return launch_fwd<0>(s, dev_trace);
    return launch_bwd<0>(s, dev_trace);
}

int k_apply_precon(Solver* s) {
    IFL_TRY(launch_fwd<0>(s));
    IFL_TRY(launch_bwd<0>(s));
    return IFL_OK;
}

int k_matvec(Solver* s) {
    VecArgs a = vec_args(s);
    matvec_kernel<0><<<grid_blocks(skew_n(s), 256), 256, 0, s->stream>>>(a, 0);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

int k_dot(Solver* s, int which, double* out) {
    VecArgs a = vec_args(s);
    const double* u = s->z;
    const double* v = which == 0 ? s->r : s->s;
    int nb = grid_blocks(skew_n(s), 256);
    int seq = s->cfg.dot_order == IFL_DOT_SEQUENTIAL;
    double* d_out = &s->st->dot_tmp;
    if (!seq) {
        dot_partial_kernel<<<nb, 256, 0, s->stream>>>(a, u, v);
        IFL_LAUNCH_CHECK();
    }
    dot_finish_kernel<<<1, 256, 0, s->stream>>>(a, u, v, nb, seq, d_out);
    IFL_LAUNCH_CHECK();
    IFL_HIP(hipMemcpyAsync(out, d_out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int k_inf_norm(Solver* s, double* out) {
    VecArgs a = vec_args(s);
    IFL_HIP(hipMemsetAsync(&s->st->max_err, 0, sizeof(double), s->stream));
    inf_norm_kernel<<<grid_blocks(skew_n(s), 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    IFL_HIP(hipMemcpyAsync(out, &s->st->max_err, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

struct VecArgs32 {
    Skew sk;
    float *e, *s, *z, *r, *q;
    const float *adiag, *apx, *apy;
    const uint8_t* flag;
    double* partial;
    PcgState* st;
    int pm, vm;
};

// e += alpha*s, s = z + beta*s (mode 0); mode 2 only the e update
__global__ void vec_update32_kernel(VecArgs32 a, int mode, int check_done) {
    PcgState* st = a.st;
    if (check_done && st->done) return;
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x, y;
    if (!skew_cell(a.sk, i, x, y)) return;
    const bool fl = a.flag[i] != 0;
    if (a.vm && !fl) return;
    const float sv = a.s[i];
    if (mode != 1) a.e[i] = a.e[i] + sv * (float)st->alpha;
    if (mode != 2) a.s[i] = (a.pm && !fl ? a.q[i] : a.z[i] - (float)st->zmean) + sv * (float)st->beta;
}

// the first direction: s = z (less its mean over FLUID cells when deflating)
__global__ void mixed_init_s_kernel(VecArgs32 a) {
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float zm = a.flag[i] ? (float)a.st->zmean : 0.0f;
    a.s[i] = a.z[i] - zm;
}

// sums over FLUID cells of r (double) and the FLUID count, per block
__global__ __launch_bounds__(256) void mixed_rsum_kernel(VecArgs a, long n) {
    __shared__ double red[4], cnt[4];
    double v = 0.0, c = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int x, y;
        if (skew_cell(a.sk, i, x, y) && a.flag[i]) {
            v += a.r[i];
            c += 1.0;
        }
    }
    v = wave_sum(v);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = v;
        cnt[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.partial[2 * blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        a.partial[2 * blockIdx.x + 1] = (cnt[0] + cnt[1]) + (cnt[2] + cnt[3]);
    }
}

// mean of r over FLUID cells into outer->zmean (nb blocks of partial pairs),
// and into st->rmean when st is given
__global__ __launch_bounds__(256) void mixed_rmean_kernel(const double* partial, int nb, PcgState* outer,
                                                          PcgState* st) {
    __shared__ double red[256], cnt[256];
    double v = 0.0, c = 0.0;
    for (int k = threadIdx.x; k < nb; k += 256) {
        v += partial[2 * k];
        c += partial[2 * k + 1];
    }
    red[threadIdx.x] = v;
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (threadIdx.x < off) {
            red[threadIdx.x] += red[threadIdx.x + off];
            cnt[threadIdx.x] += cnt[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        outer->zmean = cnt[0] > 0 ? red[0] / cnt[0] : 0.0;
        outer->dot_tmp = cnt[0];
        if (st) st->rmean = outer->zmean;
    }
}

// finish C of the inner loop: with deflation, the mean of z over FLUID cells
// (z's partial sums follow the z.r ones) before sigma and beta
__global__ __launch_bounds__(256) void fin_sigma32_kernel(VecArgs a, int n, int init, double tol, int limit,
                                                          int deflate, const PcgState* outer) {
    PcgState* st = a.st;
    if (!init && st->done) return;
    if (!init && fin_converged(st, tol)) return;
    if (deflate) {
        const double zs = block_sum_partials(a.partial + n, n);
        if (threadIdx.x == 0) st->zmean = zs / outer->dot_tmp;
        __syncthreads();
    }
    fin_sigma_body(a, n, 0, init, tol, limit);
}

// q = A s in single precision (the matrixVectorProduct stencil), q.s partials in double
__global__ __launch_bounds__(256) void matvec32_kernel(VecArgs32 a, int check_done) {
    __shared__ double red[4];
    PcgState* st = a.st;
    if (check_done && st->done) return;
    const Skew& sk = a.sk;
    long n = (long)sk.nstrips * sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double prod = 0.0;
    int x, y;
    if (i < n && skew_cell(sk, i, x, y)) {
        const int l = y & 63;
        const float* b = a.s;
        float t = a.adiag[i] * b[i];
        if (x > 0) t += a.apx[i - kStrip] * b[i - kStrip];
        if (y > 0) {
            long js = l > 0 ? i - kStrip - 1 : sk.pos(x, y - 1);
            t += a.apy[js] * b[js];
        }
        if (x < sk.W - 1) t += a.apx[i] * b[i + kStrip];
        if (y < sk.H - 1) {
            long jn = l < kStrip - 1 ? i + kStrip + 1 : sk.pos(x, y + 1);
            t += a.apy[i] * b[jn];
        }
        a.q[i] = t;
        if (!a.vm || a.flag[i]) prod = (double)t * (double)b[i];
    }
    prod = wave_sum(prod);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// single-precision copies: the matrix (what = 0), or the start of a round:
// r32 = r, e32 = 0, q32 = 0 (what = 1)
__global__ void mixed_copy_kernel(VecArgs a, VecArgs32 b, int what) {
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (what == 0) {
        ((float*)b.adiag)[i] = (float)a.adiag[i];
        ((float*)b.apx)[i] = (float)a.apx[i];
        ((float*)b.apy)[i] = (float)a.apy[i];
    } else {
        // with deflation (pressure solves), r less its mean over FLUID cells
        b.r[i] = (float)(a.flag[i] ? a.r[i] - a.st->zmean : a.r[i]);
        b.e[i] = 0.0f;
        b.q[i] = 0.0f;
    }
}

// the end of a round, in double: p += e, r -= A e (the matrixVectorProduct
// stencil on e), max|r| over the norm domain into st_outer
__global__ __launch_bounds__(256) void mixed_outer_kernel(VecArgs a, const float* __restrict__ e32, PcgState* outer) {
    const Skew& sk = a.sk;
    long n = (long)sk.nstrips * sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double m = 0.0;
    int x, y;
    if (i < n && skew_cell(sk, i, x, y)) {
        const bool fl = a.flag[i] != 0;
        const int l = y & 63;
        const double ei = (double)e32[i];
        double t = a.adiag[i] * ei;
        if (x > 0) t += a.apx[i - kStrip] * (double)e32[i - kStrip];
        if (y > 0) {
            long js = l > 0 ? i - kStrip - 1 : sk.pos(x, y - 1);
            t += a.apy[js] * (double)e32[js];
        }
        if (x < sk.W - 1) t += a.apx[i] * (double)e32[i + kStrip];
        if (y < sk.H - 1) {
            long jn = l < kStrip - 1 ? i + kStrip + 1 : sk.pos(x, y + 1);
            t += a.apy[i] * (double)e32[jn];
        }
        if (!a.vm || fl) {
            a.p[i] = a.p[i] + ei;
            const double rn = a.r[i] - t;
            a.r[i] = rn;
            m = fabs(rn);
        }
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax((unsigned long long*)&outer->max_err, (unsigned long long)as_bits(m));
}

// r -= rmean (sign -1, the start of a deflated solve) or r += rmean (sign +1,
// its end) at FLUID cells
__global__ void defl_shift_kernel(VecArgs a, double sign) {
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || !a.flag[i]) return;
    a.r[i] = a.r[i] + sign * a.st->rmean;
}
__global__ void rmean_clear_kernel(PcgState* st) { st->rmean = 0.0; }

// the first direction of a deflated solve: s = z less its mean over FLUID cells (7:856)
__global__ void defl_init_s_kernel(VecArgs a) {
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    a.s[i] = a.flag[i] ? a.z[i] - a.st->zmean : a.z[i];
}

// Iteration `it` of the solve; each kernel returns at once when the solve is
// already done.
// iterations fuse the vector update into the product (matvec_update_kernel)
static bool fused_update(const Solver* s, const VecArgs& a) {
    return s->s2 && s->cfg.dot_order == IFL_DOT_TREE && a.vm && !a.defl;
}
// the array holding s after iteration `it` (s and s2 by turns when fused)
static double* s_after(Solver* s, int it) { return s->s2 && it > 0 && (it & 1) ? s->s2 : s->s; }

static int launch_iteration(Solver* s, int it, int limit) {
    VecArgs a = vec_args(s);
    long n = skew_n(s);
    int nbv = grid_blocks(n, 256);
    int seq = s->cfg.dot_order == IFL_DOT_SEQUENTIAL;
    const bool fuse = fused_update(s, a);
    if (it > 0 && !fuse) {
        vec_update_kernel<<<nbv, 256, 0, s->stream>>>(a, 0, 1);
        IFL_LAUNCH_CHECK();
    }
    IFL_TRY(prof_begin(s, IFL_KERNEL_MATVEC, it));
    if (it > 0 && fuse) {
        matvec_update_kernel<<<nbv, 256, 0, s->stream>>>(a, s_after(s, it - 1), s_after(s, it));
    } else {
        matvec_kernel<1><<<nbv, 256, 0, s->stream>>>(a, 1);
    }
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_end(s));
    fin_alpha_kernel<<<1, 256, 0, s->stream>>>(a, nbv, seq);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_begin(s, IFL_KERNEL_SWEEP_FWD, it));
    IFL_TRY(launch_fwd<1>(s));
    IFL_TRY(prof_end(s));
    // the convergence test (7:872-876) is in the backward sweep and the finish after it
    IFL_TRY(prof_begin(s, IFL_KERNEL_SWEEP_BWD, it));
    IFL_TRY(launch_bwd<1>(s));
    IFL_TRY(prof_end(s));
    if (a.defl)
        fin_sigma32_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, 0, s->cfg.tol, limit, 1, s->st_outer);
    else
        fin_sigma_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, seq, 0, s->cfg.tol, limit);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ---------------------------------------------------------------- the mixed form
// project() of the mixed form: defect correction in double around an inner
// conjugate-gradient iteration in single precision. The outer loop keeps the
// reference's double residual r and solution p; each round solves A e = r in
// single precision (the MIC(0) factor, the matrix and every vector in float;
// dot products, alpha and beta in double) until its residual has fallen by
// kInnerDrop, then adds e to p and takes A e off r in double. The stopping
// test is the reference's, max|r| < tol on that double residual (updated
// recursively, as the reference's is), and the budget counts inner
// iterations.
#ifndef IFL_INNER_DROP
#define IFL_INNER_DROP 1e-4
#endif
#ifndef IFL_MAX_OUTER
#define IFL_MAX_OUTER 16
#endif
constexpr double kInnerDrop = IFL_INNER_DROP;
constexpr int kMaxOuter = IFL_MAX_OUTER;

static VecArgs32 vec_args32(Solver* s) {
    VecArgs32 a;
    a.sk = s->sk;
    a.e = s->e32; a.s = s->s32; a.z = s->z32; a.r = s->r32; a.q = s->q32;
    a.adiag = s->ad32; a.apx = s->apx32; a.apy = s->apy32;
    a.flag = s->dflag;
    a.partial = s->partial;
    a.st = s->st;
    a.pm = s->level >= 4;
    a.vm = s->level >= 6;
    return a;
}

static StreamArgs fwd_args2(Solver* s) {
    StreamArgs a = stream_args(s, SW_FWD);
    a.in[0] = s->r32;
    a.in[1] = s->cx32;
    a.in[2] = s->cy32;
    a.in[3] = s->pc32;
    a.in[4] = s->q32;
    a.out[0] = s->y32;
    a.out[1] = s->r32;
    return a;
}
static StreamArgs bwd_args2(Solver* s) {
    StreamArgs a = stream_args(s, SW_BWD);
    a.in[0] = s->y32;
    a.in[1] = s->cx32;
    a.in[2] = s->cy32;
    a.in[3] = s->pc32;
    a.in[4] = s->r32;
    a.in[5] = s->q32;
    a.out[0] = s->z32;
    return a;
}
static int launch_fwd2(Solver* s) {
    StreamArgs a = fwd_args2(s);
    if (s->level == 3) return launch_stream<SW_FWD, 2, 3, 5, float>(s, a);
    if (s->level <= 5) return launch_stream<SW_FWD, 2, 4, 5, float>(s, a);
    return launch_stream<SW_FWD, 2, 6, 5, float>(s, a);
}
static int launch_bwd2(Solver* s) {
    StreamArgs a = bwd_args2(s);
    if (s->level == 3) return launch_stream<SW_BWD, 2, 3, 5, float>(s, a);
    if (s->level <= 5) return launch_stream<SW_BWD, 2, 4, 6, float>(s, a);
    return launch_stream<SW_BWD, 2, 6, 5, float>(s, a);
}

// one inner iteration (the phases of launch_iteration on the single-precision vectors)
static int launch_iteration32(Solver* s, int it, int limit, double tol, int gidx, int deflate) {
    VecArgs a = vec_args(s);
    VecArgs32 b = vec_args32(s);
    long n = skew_n(s);
    int nbv = grid_blocks(n, 256);
    if (it > 0) {
        vec_update32_kernel<<<nbv, 256, 0, s->stream>>>(b, 0, 1);
        IFL_LAUNCH_CHECK();
    }
    IFL_TRY(prof_begin(s, IFL_KERNEL_MATVEC, gidx));
    matvec32_kernel<<<nbv, 256, 0, s->stream>>>(b, 1);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_end(s));
    fin_alpha_kernel<<<1, 256, 0, s->stream>>>(a, nbv, 0);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_begin(s, IFL_KERNEL_SWEEP_FWD, gidx));
    IFL_TRY(launch_fwd2(s));
    IFL_TRY(prof_end(s));
    fin_check_kernel<<<1, 1, 0, s->stream>>>(s->st, tol);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_begin(s, IFL_KERNEL_SWEEP_BWD, gidx));
    IFL_TRY(launch_bwd2(s));
    IFL_TRY(prof_end(s));
    fin_sigma32_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, 0, tol, limit, deflate, s->st_outer);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

static int read_state(Solver* s, PcgState* dev, PcgState* host) {
    IFL_HIP(hipMemcpyAsync(host, dev, sizeof(PcgState), hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

int k_project_mixed(Solver* s, int limit, ifl_solve_stats* stats) {
    VecArgs a = vec_args(s);
    VecArgs32 b = vec_args32(s);
    long n = skew_n(s);
    int nbv = grid_blocks(n, 256);
    const double tol = s->cfg.tol;
    // The pressure matrix is singular (A 1 = 0 over the FLUID cells, closed
    // walls), and MIC(0) amplifies the component of z along the constant
    // vector, which single precision cannot carry: pressure solves deflate
    // it, taking the mean over FLUID cells off the round's right-hand side
    // and off every z. A e then has no constant part and the double residual
    // keeps the reference's (its null-space part is never changed by A e).
    const int deflate = s->solve_kind == SOLVE_PRESSURE;
    const int nbr = std::max(1, std::min(nbv, s->npartial / 2 - 1));
    if (s->profiling) IFL_HIP(hipEventRecord(s->ev0, s->stream));
    IFL_HIP(hipMemsetAsync(s->p, 0, n * sizeof(double), s->stream));  // 7:854
    mixed_copy_kernel<<<nbv, 256, 0, s->stream>>>(a, b, 0);
    IFL_LAUNCH_CHECK();
    // max|r| of the right-hand side (7:858)
    VecArgs ao = a;
    ao.st = s->st_outer;
    state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st_outer);
    IFL_LAUNCH_CHECK();
    inf_norm_kernel<<<nbv, 256, 0, s->stream>>>(ao);
    IFL_LAUNCH_CHECK();
    PcgState ho;
    IFL_TRY(read_state(s, s->st_outer, &ho));
    double rmax = ho.max_err;
    int status = IFL_SOLVE_NOT_RUN, printed = 0, total = 0;
    if (rmax < tol) {
        status = IFL_SOLVE_INITIAL_SMALL;
        printed = -1;
    } else if (limit <= 0) {
        status = IFL_SOLVE_BUDGET;
        printed = limit;
    }
    for (int round = 0; status == IFL_SOLVE_NOT_RUN; round++) {
        // a round: r32 = r, e32 = 0, then PCG in single precision from e = 0
        const double tin = std::max(kInnerDrop * rmax, 0.25 * tol);
        const int lim_in = limit - total;
        state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st_outer);
        IFL_LAUNCH_CHECK();
        if (deflate) {
            mixed_rsum_kernel<<<nbr, 256, 0, s->stream>>>(a, n);
            IFL_LAUNCH_CHECK();
            mixed_rmean_kernel<<<1, 256, 0, s->stream>>>(s->partial, nbr, s->st_outer, nullptr);
            IFL_LAUNCH_CHECK();
        }
        mixed_copy_kernel<<<nbv, 256, 0, s->stream>>>(ao, b, 1);
        IFL_LAUNCH_CHECK();
        state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st);  // alpha = 0: the first sweeps apply M^-1 to r32
        IFL_LAUNCH_CHECK();
        IFL_TRY(launch_fwd2(s));
        IFL_TRY(launch_bwd2(s));
        fin_sigma32_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, 1, tin, lim_in, deflate, s->st_outer);
        IFL_LAUNCH_CHECK();
        mixed_init_s_kernel<<<nbv, 256, 0, s->stream>>>(b);
        IFL_LAUNCH_CHECK();
        int launched = 0, batch = 4;
        PcgState h;
        while (true) {
            IFL_TRY(read_state(s, s->st, &h));
            if (h.done || h.timeout || launched >= lim_in) break;
            int nb = std::min(batch, lim_in - launched);
            for (int k = 0; k < nb; k++)
                IFL_TRY(launch_iteration32(s, launched + k, lim_in, tin, total + launched + k, deflate));
            launched += nb;
            batch = std::min(batch * 2, 64);
        }
        IFL_TRY(read_state(s, s->st, &h));
        if (h.timeout) {
            s->err = "PCG sweep hand-off timed out";
            (void)edge_reset(s);
            return IFL_ERR_DEVICE;
        }
        int its = 0;
        if (h.status == IFL_SOLVE_CONVERGED) {
            its = h.printed_iter + 1;
            vec_update32_kernel<<<nbv, 256, 0, s->stream>>>(b, 2, 0);  // the last e += alpha*s
            IFL_LAUNCH_CHECK();
        } else if (h.status == IFL_SOLVE_BUDGET) {
            its = h.iter;
            if (its > 0) {
                vec_update32_kernel<<<nbv, 256, 0, s->stream>>>(b, 2, 0);
                IFL_LAUNCH_CHECK();
            }
        }
        total += its;
        // p += e, r -= A e in double; the new max|r|
        state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st_outer);
        IFL_LAUNCH_CHECK();
        mixed_outer_kernel<<<nbv, 256, 0, s->stream>>>(a, s->e32, s->st_outer);
        IFL_LAUNCH_CHECK();
        IFL_TRY(read_state(s, s->st_outer, &ho));
        rmax = ho.max_err;
        if (rmax < tol) {
            status = IFL_SOLVE_CONVERGED;
            printed = std::max(total - 1, 0);
        } else if (total >= limit || round + 1 >= kMaxOuter || its == 0) {
            status = IFL_SOLVE_BUDGET;
            printed = limit;
        }
    }
    if (s->profiling) {
        IFL_HIP(hipEventRecord(s->ev1, s->stream));
        IFL_HIP(hipEventSynchronize(s->ev1));
        float ms = 0;
        IFL_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        Profile& pr = s->prof[IFL_KERNEL_PCG];
        pr.launches += 1;
        pr.ms += ms;
        pr.iterations += total;
        pr.cell_iterations += (double)total * s->W * s->H;
        int lim[IFL_KERNEL_COUNT] = {0};
        lim[IFL_KERNEL_MATVEC] = lim[IFL_KERNEL_SWEEP_FWD] = lim[IFL_KERNEL_SWEEP_BWD] = total;
        IFL_TRY(prof_collect(s, lim));
    } else {
        IFL_HIP(hipStreamSynchronize(s->stream));
    }
    if (stats) {
        stats->status = status;
        stats->iterations = printed;
        stats->max_error = rmax;
    }
    if (s->cfg.verbose) {
        if (status == IFL_SOLVE_CONVERGED)
            printf("Exiting solver after %d iterations, maximum error is %f\n", printed, rmax);
        else if (status == IFL_SOLVE_BUDGET)
            printf("Exceeded budget of %d iterations, maximum error was %f\n", limit, rmax);
        else if (status == IFL_SOLVE_INITIAL_SMALL && s->level >= 6)
            printf("Initial guess sufficiently small\n");
    }
    return IFL_OK;
}

// project(limit) 7:853-886
int k_project(Solver* s, int limit, ifl_solve_stats* stats) {
    if (s->cfg.precision == IFL_PREC_REFINE) return k_project_mixed(s, limit, stats);
    VecArgs a = vec_args(s);
    long n = skew_n(s);
    int seq = s->cfg.dot_order == IFL_DOT_SEQUENTIAL;
    if (s->profiling) IFL_HIP(hipEventRecord(s->ev0, s->stream));
    state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st);
    IFL_LAUNCH_CHECK();
    IFL_HIP(hipMemsetAsync(s->p, 0, n * sizeof(double), s->stream));  // 7:854
    if (a.defl) {
        // the mixed form's pressure solve: the FLUID count for the means of z
        const int nbr = std::max(1, std::min(grid_blocks(n, 256), s->npartial / 2 - 1));
        state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st_outer);
        IFL_LAUNCH_CHECK();
        mixed_rsum_kernel<<<nbr, 256, 0, s->stream>>>(a, n);
        IFL_LAUNCH_CHECK();
        mixed_rmean_kernel<<<1, 256, 0, s->stream>>>(s->partial, nbr, s->st_outer, s->st);
        IFL_LAUNCH_CHECK();
        // The preconditioner is applied as P M^-1 P, with P taking off the
        // mean over FLUID cells: r is kept less its mean (which A s, summing
        // to zero over FLUID cells, does not change) and z is taken less its
        // mean. The closed-wall pressure matrix is singular along the
        // constant, and a right-hand side off by a constant makes the
        // one-sided form lose the conjugate-gradient properties.
        defl_shift_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a, -1.0);
        IFL_LAUNCH_CHECK();
        // applyPreconditioner(_z, _r) 7:855 in the loop's form (its z sums),
        // then s = z less its mean (7:856)
        IFL_TRY(launch_fwd<0>(s));
        IFL_TRY(launch_bwd<0>(s));
        fin_sigma32_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, 1, s->cfg.tol, limit, 1, s->st_outer);
        IFL_LAUNCH_CHECK();
        defl_init_s_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
        IFL_LAUNCH_CHECK();
    } else {
        IFL_TRY(launch_fwd<0>(s));                                         // applyPreconditioner(_z, _r) 7:855
        IFL_TRY(launch_bwd<0>(s));
        IFL_HIP(hipMemcpyAsync(s->s, s->z, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));  // 7:856
        fin_sigma_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, seq, 1, s->cfg.tol, limit);  // 7:858-864
        IFL_LAUNCH_CHECK();
    }
    int launched = 0;
    int batch = 4;
    while (true) {
        IFL_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, s->stream));
        IFL_HIP(hipStreamSynchronize(s->stream));
        if (s->st_host->done || s->st_host->timeout || launched >= limit) break;
        int nb = std::min(batch, limit - launched);
        for (int k = 0; k < nb; k++) IFL_TRY(launch_iteration(s, launched + k, limit));
        launched += nb;
        batch = std::min(batch * 2, 64);
    }
    PcgState h = *s->st_host;
    if (h.timeout) {
        s->err = "PCG sweep hand-off timed out";
        (void)edge_reset(s);
        return IFL_ERR_DEVICE;
    }
    // with the fused update, s is where the last product that ran read it;
    // the tails below work on that array, and it ends in s
    double* s_last = nullptr;
    if (fused_update(s, a)) {
        const int worked = h.status == IFL_SOLVE_CONVERGED ? h.printed_iter + 1 : h.iter;
        s_last = s_after(s, worked - 1);
        a.s = s_last;
    }
    // deferred tails of the last iteration
    if (h.status == IFL_SOLVE_CONVERGED) {
        vec_update_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a, 2, 0);  // p += alpha*s (7:869)
        IFL_LAUNCH_CHECK();
        IFL_HIP(hipMemcpyAsync(s->z, s->qv, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));  // _z = A s
    } else if (h.status == IFL_SOLVE_BUDGET && limit > 0) {
        vec_update_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a, 0, 0);
        IFL_LAUNCH_CHECK();
        if (a.pm) {
            z_nonfluid_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
            IFL_LAUNCH_CHECK();
        }
    }
    if (s_last && s_last != s->s)
        IFL_HIP(hipMemcpyAsync(s->s, s_last, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    if (a.defl) {
        // r back to the residual itself
        defl_shift_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a, 1.0);
        IFL_LAUNCH_CHECK();
        rmean_clear_kernel<<<1, 1, 0, s->stream>>>(s->st);
        IFL_LAUNCH_CHECK();
    }
    if (s->profiling) {
        IFL_HIP(hipEventRecord(s->ev1, s->stream));
        IFL_HIP(hipEventSynchronize(s->ev1));
        float ms = 0;
        IFL_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        Profile& pr = s->prof[IFL_KERNEL_PCG];
        pr.launches += 1;
        pr.ms += ms;
        int its = h.status == IFL_SOLVE_CONVERGED ? h.printed_iter + 1 : h.iter;
        pr.iterations += its;
        pr.cell_iterations += (double)its * s->W * s->H;
        // iterations whose sweeps did work: the converging iteration stops
        // after its forward sweep
        int lim[IFL_KERNEL_COUNT] = {0};
        lim[IFL_KERNEL_MATVEC] = lim[IFL_KERNEL_SWEEP_FWD] = its;
        lim[IFL_KERNEL_SWEEP_BWD] = h.status == IFL_SOLVE_CONVERGED ? its - 1 : its;
        IFL_TRY(prof_collect(s, lim));
    } else {
        IFL_HIP(hipStreamSynchronize(s->stream));
    }
    if (stats) {
        stats->status = h.status;
        stats->iterations = h.printed_iter;
        stats->max_error = h.max_err;
    }
    if (s->cfg.verbose) {
        if (h.status == IFL_SOLVE_CONVERGED)
            printf("Exiting solver after %d iterations, maximum error is %f\n", h.printed_iter, h.max_err);
        else if (h.status == IFL_SOLVE_BUDGET)
            printf("Exceeded budget of %d iterations, maximum error was %f\n", limit, h.max_err);
        else if (h.status == IFL_SOLVE_INITIAL_SMALL && s->level >= 6)
            printf("Initial guess sufficiently small\n");
    }
    return IFL_OK;
}

// ---------------------------------------------------------------- batched solves
// project() of several independent solvers of one level together: each phase
// of a PCG iteration is one launch for all of them, its blocks dealt from a
// table of (solver, block) pairs. Sweep strips are listed by rank from the
// sweep's start (ascending strips for the forward sweep, descending for the
// backward one), so a strip's upstream neighbour always has the lower block
// id: it is dispatched first, and the strips need not all be resident at
// once however many solvers the batch holds. Solvers that have converged
// leave the tables at the host's next read of the convergence flags, and
// return at once from every launch until then.
struct BatchInst {
    VecArgs v;
    int nbv;      // vector blocks of 256 cells
    int nstrips;  // strips, = the backward sweep's partials
    double tol;
    int limit;
    int pad;
    double* s2;  // the second s when the batch folds the vector update into the product, else null
};

// The array holding an instance's s after iteration `it` (s and s2 by turns
// when the batch folds the update into the product, as for one solver).
IFL_DEV double* batch_s_after(const BatchInst& b, int it) { return b.s2 && it > 0 && (it & 1) ? b.s2 : b.v.s; }
// The array holding it at the end of a solve: where the last product read it.
IFL_DEV double* batch_s_last(const BatchInst& b) {
    const PcgState* st = b.v.st;
    const int worked = st->status == IFL_SOLVE_CONVERGED ? st->printed_iter + 1 : st->iter;
    return batch_s_after(b, worked - 1);
}

struct BatchDev {
    int n = 0, nvblk = 0, nsblk = 0;
    size_t shmem_f = 0, shmem_b = 0, shmem_p = 0, shmem_f0 = 0, shmem_b0 = 0;
    BatchInst* inst = nullptr;    // n
    StreamArgs* fwd = nullptr;    // n: the sweeps inside the PCG loop
    StreamArgs* bwd = nullptr;    // n
    StreamArgs* pre = nullptr;    // n: buildPreconditioner
    StreamArgs* fwd0 = nullptr;   // n: applyPreconditioner at the entry of project()
    StreamArgs* bwd0 = nullptr;   // n
    int2* vmap = nullptr;         // nvblk
    int2* fmap = nullptr;         // nsblk
    int2* bmap = nullptr;         // nsblk
    PcgState* states = nullptr;   // n gathered states
    PcgState* states_host = nullptr;
    PcgState** st_ptrs = nullptr; // n
    std::vector<Solver*> solvers; // the list the tables were built for
    std::vector<hipEvent_t> ready;  // per solver: its entry into the solve is queued
    hipEvent_t joined = nullptr;    // the batched iterations are queued
    // the full block tables; the device tables hold those of the instances
    // still iterating (rebuilt when some have finished)
    std::vector<int2> vmap_h, fmap_h, bmap_h, vmap_c, fmap_c, bmap_c;
    // the matrix-vector product and the vector update take kVecBlocks
    // consecutive blocks of an instance per workgroup: (instance, group)
    std::vector<int2> vmap4_h, vmap4_c;
    int2* vmap4 = nullptr;
    int nvblk4 = 0;
    int nvblk_full = 0, nsblk_full = 0;
    bool compacted = false;
    bool fused = false;  // the vector update folded into the product (every instance: levels 6-8 with s2)
    // profiling: events around the fused launches (matvec, forward and
    // backward sweep) of the current solve
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timed;
};

void batch_dev_free(BatchDev* d) {
    if (!d) return;
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    F(d->inst); F(d->fwd); F(d->bwd); F(d->pre); F(d->fwd0); F(d->bwd0); F(d->vmap); F(d->fmap); F(d->bmap); F(d->states); F(d->st_ptrs);
    F(d->vmap4);
    if (d->states_host) (void)hipHostFree(d->states_host);
    for (hipEvent_t e : d->ready) (void)hipEventDestroy(e);
    for (hipEvent_t e : d->ev_pool) (void)hipEventDestroy(e);
    for (auto& t : d->timed) {
        (void)hipEventDestroy(t.second.first);
        (void)hipEventDestroy(t.second.second);
    }
    if (d->joined) (void)hipEventDestroy(d->joined);
    delete d;
}

namespace {

// blocks of 256 cells per workgroup in the batch's vector update and
// matrix-vector product: an instance's arguments are loaded once for all of
// them (each block's partial is the single-solver kernel's)
constexpr int kVecBlocks = 4;

// vec_update_body (mode 0) over kVecBlocks blocks at once, every load of the
// workgroup's cells issued before the first store; each cell's values are
// vec_update_body's.
__global__ __launch_bounds__(256) void batch_vec_update_kernel(const BatchInst* t, const int2* map) {
    constexpr int K = kVecBlocks;
    const int2 m = map[blockIdx.x];
    const VecArgs a = t[m.x].v;
    const int nbv = t[m.x].nbv;
    const PcgState* st = a.st;
    if (st->done) return;
    const long n = (long)a.sk.nstrips * a.sk.strip_sz;
    bool on[K], fl[K];
    long ci[K];
    double sv[K], pv[K], zv[K], qv[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const long blk = (long)m.y * K + k;
        long i = blk * 256 + threadIdx.x;
        int x, y;
        on[k] = blk < nbv && i < n && skew_cell(a.sk, i, x, y);
        if (!on[k]) i = 0;  // cell 0 stands in for the loads; nothing is stored
        ci[k] = i;
        fl[k] = a.flag[i] != 0;
        sv[k] = a.s[i];
        pv[k] = a.p[i];
        zv[k] = a.z[i];
        qv[k] = a.pm ? a.q[i] : 0.0;
    }
    const double alpha = st->alpha, beta = st->beta, zmean = st->zmean;
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (!on[k] || (a.vm && !fl[k])) continue;
        a.p[ci[k]] = pv[k] + sv[k] * alpha;
        double zi = a.pm && !fl[k] ? qv[k] : zv[k];
        if (a.defl && fl[k]) zi = zi - zmean;
        a.s[ci[k]] = zi + sv[k] * beta;
    }
}

// matvec_body<1> over kVecBlocks blocks at once: every load of the
// workgroup's cells is issued before the first product. Each block's partial
// is matvec_body's: the same terms, the same wave sums over the same lanes
// and the same combination of the block's four wave sums.
__global__ __launch_bounds__(256) void batch_matvec_kernel(const BatchInst* t, const int2* map) {
    constexpr int K = kVecBlocks;
    __shared__ double red[K][4];
    const int2 m = map[blockIdx.x];
    const VecArgs a = t[m.x].v;
    const int nbv = t[m.x].nbv;
    if (a.st->done) return;
    const Skew& sk = a.sk;
    const long n = (long)sk.nstrips * sk.strip_sz;
    const int W = sk.W, H = sk.H;
    bool on[K], dt[K], hw[K], hs[K], he[K], hn[K];
    long ci[K];
    double b0[K], bw[K], bs[K], be[K], bn[K], ad[K], axw[K], ays[K], ax[K], ay[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const long blk = (long)m.y * K + k;
        long i = blk * 256 + threadIdx.x;
        int x = 0, y = 0;
        on[k] = blk < nbv && i < n && skew_cell(sk, i, x, y);
        if (!on[k]) i = 0;  // cell 0 stands in for the loads; nothing is kept
        ci[k] = i;
        const int l = y & 63;
        hw[k] = on[k] && x > 0;
        hs[k] = on[k] && y > 0;
        he[k] = on[k] && x < W - 1;
        hn[k] = on[k] && y < H - 1;
        // a missing neighbour reads cell i and is not added
        const long jw = hw[k] ? i - kStrip : i;
        const long js = hs[k] ? (l > 0 ? i - kStrip - 1 : sk.pos(x, y - 1)) : i;
        const long je = he[k] ? i + kStrip : i;
        const long jn = hn[k] ? (l < kStrip - 1 ? i + kStrip + 1 : sk.pos(x, y + 1)) : i;
        b0[k] = a.s[i]; bw[k] = a.s[jw]; bs[k] = a.s[js]; be[k] = a.s[je]; bn[k] = a.s[jn];
        ad[k] = a.adiag[i]; axw[k] = a.apx[jw]; ays[k] = a.apy[js]; ax[k] = a.apx[i]; ay[k] = a.apy[i];
        // matvec_body's dot term: every cell unless the vector operations are masked
        dt[k] = on[k] && (!a.vm || a.flag[i]);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double tv = ad[k] * b0[k];
        tv = hw[k] ? tv + axw[k] * bw[k] : tv;
        tv = hs[k] ? tv + ays[k] * bs[k] : tv;
        tv = he[k] ? tv + ax[k] * be[k] : tv;
        tv = hn[k] ? tv + ay[k] * bn[k] : tv;
        if (on[k]) a.q[ci[k]] = tv;
        double prod = dt[k] ? tv * b0[k] : 0.0;
        prod = wave_sum(prod);
        if (lane == 0) red[k][w] = prod;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        const long blk = (long)m.y * K + k;
        if (blk < nbv) a.partial[blk] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
    }
}

// matvec_update_body for every instance: the vector update of iteration
// it - 1 and the product of iteration it in one pass (levels 6-8), one block
// of an instance per workgroup; s is read from the array iteration it - 1
// left it in and the new s goes to the other. (Four blocks per workgroup, the
// block reduction's LDS reused between them: 25.3 against 23.7 us per launch
// on 16 FLIP instances at 256^2.)
__global__ __launch_bounds__(256) void batch_matvec_update_kernel(const BatchInst* t, const int2* map, int it) {
    const int2 m = map[blockIdx.x];
    const BatchInst& b = t[m.x];
    const VecArgs a = b.v;
    matvec_update_body(a, m.y, batch_s_after(b, it - 1), batch_s_after(b, it));
}

__global__ __launch_bounds__(256) void batch_fin_alpha_kernel(const BatchInst* t) {
    fin_alpha_body(t[blockIdx.x].v, t[blockIdx.x].nbv, 0);
}

__global__ void batch_fin_check_kernel(const BatchInst* t, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) fin_check_body(t[i].v.st, t[i].tol);
}

__global__ __launch_bounds__(256) void batch_fin_sigma_kernel(const BatchInst* t) {
    const BatchInst& b = t[blockIdx.x];
    fin_sigma_body(b.v, b.nstrips, 0, 0, b.tol, b.limit);
}

template <int KIND, int MODE, int LV, int NG>
__global__ __launch_bounds__(64 * kWaves) void batch_sweep_kernel(const StreamArgs* t, const int2* map) {
    const int2 m = map[blockIdx.x];
    const StreamArgs a = t[m.x];
    sweep_body<KIND, MODE, LV, NG, double, false>(a, m.y);
}

// the entry of project() (7:854-864) for every instance: the state reset,
// p = 0, s = z after the entry sweeps, then sigma and the early-exit test
__global__ void batch_reset_kernel(const BatchInst* t, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PcgState* st = t[i].v.st;
    st->zmean = 0.0;
    st->rmean = 0.0;
    st->sigma = st->alpha = st->beta = 0.0;
    st->max_err = 0.0;
    st->dot_tmp = 0.0;
    st->iter = 0;
    st->done = 0;
    st->status = IFL_SOLVE_NOT_RUN;
    st->printed_iter = 0;
}
__global__ __launch_bounds__(256) void batch_zero_p_kernel(const BatchInst* t, const int2* map) {
    const int2 m = map[blockIdx.x];
    const VecArgs& a = t[m.x].v;
    const long i = (long)m.y * blockDim.x + threadIdx.x;
    if (i < (long)a.sk.nstrips * a.sk.strip_sz) a.p[i] = 0.0;
}
__global__ __launch_bounds__(256) void batch_copy_s_kernel(const BatchInst* t, const int2* map) {
    const int2 m = map[blockIdx.x];
    const VecArgs& a = t[m.x].v;
    const long i = (long)m.y * blockDim.x + threadIdx.x;
    if (i < (long)a.sk.nstrips * a.sk.strip_sz) a.s[i] = a.z[i];
}
__global__ __launch_bounds__(256) void batch_fin_sigma_init_kernel(const BatchInst* t) {
    const BatchInst& b = t[blockIdx.x];
    fin_sigma_body(b.v, b.nstrips, 0, 1, b.tol, b.limit);
}

// The exits of project() for every solver of the batch, as k_project does
// them: converged, p += alpha s (7:869) and _z = A s; out of budget (with a
// budget), the deferred s update and _z = A s at non-FLUID cells (levels 4+).
__global__ __launch_bounds__(256) void batch_tail_kernel(const BatchInst* t, const int2* map) {
    const int2 m = map[blockIdx.x];
    const BatchInst& b = t[m.x];
    VecArgs a = b.v;
    a.s = batch_s_last(b);  // with the folded update: the tails work on that array, and it ends in s
    const int status = a.st->status;
    const long n = (long)a.sk.nstrips * a.sk.strip_sz;
    const long i = (long)m.y * blockDim.x + threadIdx.x;
    if (status == IFL_SOLVE_CONVERGED) {
        vec_update_body(a, m.y, 2, 0);
        if (i < n) a.z[i] = a.q[i];
    } else if (status == IFL_SOLVE_BUDGET && b.limit > 0) {
        vec_update_body(a, m.y, 0, 0);
        int x, y;
        if (a.pm && i < n && skew_cell(a.sk, i, x, y) && !a.flag[i]) a.z[i] = a.q[i];
    }
    if (a.s != b.v.s && i < n) b.v.s[i] = a.s[i];  // the same thread wrote it above
}

__global__ void batch_gather_kernel(PcgState* const* st, int n, PcgState* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *st[i];
}

template <typename X>
int upload(Solver* s, X** dst, const std::vector<X>& v) {
    IFL_HIP(hipMalloc((void**)dst, std::max<size_t>(1, v.size()) * sizeof(X)));
    if (!v.empty()) IFL_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
    return IFL_OK;
}

template <int LV, int NGF, int NGB>
int batch_attrs(Solver* s, BatchDev* d) {
    static std::atomic<bool> set{false};
    if (!set.load(std::memory_order_acquire)) {
        IFL_HIP(hipFuncSetAttribute((const void*)batch_sweep_kernel<SW_FWD, 1, LV, NGF>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)d->shmem_f));
        IFL_HIP(hipFuncSetAttribute((const void*)batch_sweep_kernel<SW_BWD, 1, LV, NGB>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)d->shmem_b));
        IFL_HIP(hipFuncSetAttribute((const void*)batch_sweep_kernel<SW_PRECON, 0, 3, 3>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)d->shmem_p));
        IFL_HIP(hipFuncSetAttribute((const void*)batch_sweep_kernel<SW_FWD, 0, LV, NGF - 1>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)d->shmem_f0));
        IFL_HIP(hipFuncSetAttribute((const void*)batch_sweep_kernel<SW_BWD, 0, LV, NGB>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)d->shmem_b0));
        set.store(true, std::memory_order_release);
    }
    return IFL_OK;
}

template <int LV, int NGF, int NGB>
int batch_build(Solver* const* ss, int n, BatchDev* d) {
    Solver* s0 = ss[0];
    std::vector<BatchInst> inst(n);
    std::vector<StreamArgs> fwd(n), bwd(n), pre(n), fwd0(n), bwd0(n);
    std::vector<int2> vmap, fmap, bmap;
    std::vector<PcgState*> sts(n);
    for (int i = 0; i < n; i++) {
        Solver* s = ss[i];
        BatchInst& b = inst[i];
        b.v = vec_args(s);
        b.nbv = grid_blocks(skew_n(s), 256);
        b.nstrips = s->sk.nstrips;
        b.tol = s->cfg.tol;
        b.limit = s->cfg.pcg_limit;
        b.pad = 0;
        b.s2 = nullptr;
        fwd[i] = fwd_args<1, double>(s);
        bwd[i] = bwd_args<1, double>(s);
        size_t shf = 0, shb = 0;
        IFL_TRY((ring_setup<SW_FWD, 1, LV, NGF, double>(s, fwd[i], &shf)));
        IFL_TRY((ring_setup<SW_BWD, 1, LV, NGB, double>(s, bwd[i], &shb)));
        d->shmem_f = shf;
        d->shmem_b = shb;
        pre[i] = precon_args(s);
        fwd0[i] = fwd_args<0, double>(s);
        bwd0[i] = bwd_args<0, double>(s);
        IFL_TRY((ring_setup<SW_PRECON, 0, 3, 3, double>(s, pre[i], &shf)));
        d->shmem_p = shf;
        IFL_TRY((ring_setup<SW_FWD, 0, LV, NGF - 1, double>(s, fwd0[i], &shf)));
        d->shmem_f0 = shf;
        IFL_TRY((ring_setup<SW_BWD, 0, LV, NGB, double>(s, bwd0[i], &shb)));
        d->shmem_b0 = shb;
        for (int k = 0; k < b.nbv; k++) vmap.push_back(make_int2(i, k));
        for (int k = 0; k * kVecBlocks < b.nbv; k++) d->vmap4_h.push_back(make_int2(i, k));
        sts[i] = s->st;
    }
    // Strips are dealt by their distance from the end of their instance's
    // sweep: every instance's last strip at rank `most - 1`, its first at
    // `most - nstrips` (rank by rank, larger instances first within a rank).
    // A strip's upstream neighbour thus has a lower block id, so the strips
    // need not be resident together; a strip is dispatched after the previous
    // rank, so it seldom holds a CU while it waits for its upstream
    // neighbour; and the largest instances, whose chains of strips are the
    // longest, start first. Forward / backward us per launch on the
    // 256-instance draw: instance-major order 230 / 206; ranked from each
    // sweep's start 182 / 170; this order 176 / 164 (ranks weighted further
    // toward the large instances: 182-192 / 168-176).
    {
        std::vector<int> order(n);
        for (int i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return (long)ss[x]->W * ss[x]->H > (long)ss[y]->W * ss[y]->H;
        });
        int most = 0;
        for (int i = 0; i < n; i++) most = std::max(most, ss[i]->sk.nstrips);
        for (int r = 0; r < most; r++)
            for (int i : order) {
                const int ns = ss[i]->sk.nstrips;
                const int k = r - (most - ns);  // the strip of instance i at this rank
                if (k < 0) continue;
                fmap.push_back(make_int2(i, k));
                bmap.push_back(make_int2(i, ns - 1 - k));
            }
    }
    Solver* s = s0;  // error reporting for the uploads
    d->n = n;
    d->nvblk = d->nvblk_full = (int)vmap.size();
    d->nsblk = d->nsblk_full = (int)fmap.size();
    d->vmap_h = vmap;
    d->fmap_h = fmap;
    d->bmap_h = bmap;
    // the vector update folds into the product when every instance can, as
    // for one solver (fused_update: the second s, tree dots, masked vectors)
    d->fused = true;
    for (int i = 0; i < n; i++) d->fused = d->fused && fused_update(ss[i], inst[i].v);
    for (int i = 0; i < n; i++) inst[i].s2 = d->fused ? ss[i]->s2 : nullptr;
    IFL_TRY(upload(s, &d->inst, inst));
    IFL_TRY(upload(s, &d->fwd, fwd));
    IFL_TRY(upload(s, &d->bwd, bwd));
    IFL_TRY(upload(s, &d->pre, pre));
    IFL_TRY(upload(s, &d->fwd0, fwd0));
    IFL_TRY(upload(s, &d->bwd0, bwd0));
    IFL_TRY(upload(s, &d->vmap, vmap));
    IFL_TRY(upload(s, &d->vmap4, d->vmap4_h));
    d->nvblk4 = (int)d->vmap4_h.size();
    IFL_TRY(upload(s, &d->fmap, fmap));
    IFL_TRY(upload(s, &d->bmap, bmap));
    IFL_TRY(upload(s, &d->st_ptrs, sts));
    IFL_HIP(hipMalloc((void**)&d->states, n * sizeof(PcgState)));
    IFL_HIP(hipHostMalloc((void**)&d->states_host, n * sizeof(PcgState), hipHostMallocDefault));
    d->ready.assign(n, nullptr);
    for (int i = 0; i < n; i++) IFL_HIP(hipEventCreateWithFlags(&d->ready[i], hipEventDisableTiming));
    IFL_HIP(hipEventCreateWithFlags(&d->joined, hipEventDisableTiming));
    d->solvers.assign(ss, ss + n);
    return batch_attrs<LV, NGF, NGB>(s, d);
}

// events around one fused launch when profiling
int batch_time(Solver* s, BatchDev* d, hipStream_t stream, int kernel, bool begin) {
    if (begin) {
        hipEvent_t a;
        if (d->ev_pool.empty()) {
            IFL_HIP(hipEventCreate(&a));
        } else {
            a = d->ev_pool.back();
            d->ev_pool.pop_back();
        }
        IFL_HIP(hipEventRecord(a, stream));
        d->timed.push_back({kernel, {a, nullptr}});
        return IFL_OK;
    }
    hipEvent_t b;
    if (d->ev_pool.empty()) {
        IFL_HIP(hipEventCreate(&b));
    } else {
        b = d->ev_pool.back();
        d->ev_pool.pop_back();
    }
    IFL_HIP(hipEventRecord(b, stream));
    d->timed.back().second.second = b;
    return IFL_OK;
}

template <int LV, int NGF, int NGB>
int batch_iteration(Solver* s, BatchDev* d, hipStream_t stream, int it, bool prof) {
    const bool fuse = it > 0 && d->fused;
    if (it > 0 && !fuse) {
        batch_vec_update_kernel<<<d->nvblk4, 256, 0, stream>>>(d->inst, d->vmap4);
        IFL_LAUNCH_CHECK();
    }
    if (prof) IFL_TRY(batch_time(s, d, stream, IFL_KERNEL_MATVEC, true));
    if (fuse)
        batch_matvec_update_kernel<<<d->nvblk, 256, 0, stream>>>(d->inst, d->vmap, it);
    else
        batch_matvec_kernel<<<d->nvblk4, 256, 0, stream>>>(d->inst, d->vmap4);
    IFL_LAUNCH_CHECK();
    if (prof) IFL_TRY(batch_time(s, d, stream, IFL_KERNEL_MATVEC, false));
    batch_fin_alpha_kernel<<<d->n, 256, 0, stream>>>(d->inst);
    IFL_LAUNCH_CHECK();
    if (prof) IFL_TRY(batch_time(s, d, stream, IFL_KERNEL_SWEEP_FWD, true));
    batch_sweep_kernel<SW_FWD, 1, LV, NGF><<<d->nsblk, 64 * kWaves, d->shmem_f, stream>>>(d->fwd, d->fmap);
    IFL_LAUNCH_CHECK();
    if (prof) IFL_TRY(batch_time(s, d, stream, IFL_KERNEL_SWEEP_FWD, false));
    // the convergence test is in the backward sweep and the finish after it
    if (prof) IFL_TRY(batch_time(s, d, stream, IFL_KERNEL_SWEEP_BWD, true));
    batch_sweep_kernel<SW_BWD, 1, LV, NGB><<<d->nsblk, 64 * kWaves, d->shmem_b, stream>>>(d->bwd, d->bmap);
    IFL_LAUNCH_CHECK();
    if (prof) IFL_TRY(batch_time(s, d, stream, IFL_KERNEL_SWEEP_BWD, false));
    batch_fin_sigma_kernel<<<d->n, 256, 0, stream>>>(d->inst);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// The block tables of the instances in `active` (all when null). The host
// copies stay in BatchDev until the next rebuild; every launch that reads the
// device tables has completed (the caller has synchronised the stream).
int batch_tables(Solver* s, BatchDev* d, hipStream_t stream, const std::vector<char>* active) {
    if (!active) {
        if (!d->compacted) return IFL_OK;
        d->vmap_c = d->vmap_h;
        d->vmap4_c = d->vmap4_h;
        d->fmap_c = d->fmap_h;
        d->bmap_c = d->bmap_h;
    } else {
        auto pick = [&](const std::vector<int2>& full, std::vector<int2>& out) {
            out.clear();
            for (const int2& m : full)
                if ((*active)[m.x]) out.push_back(m);
        };
        pick(d->vmap_h, d->vmap_c);
        pick(d->vmap4_h, d->vmap4_c);
        pick(d->fmap_h, d->fmap_c);
        pick(d->bmap_h, d->bmap_c);
    }
    d->nvblk = (int)d->vmap_c.size();
    d->nvblk4 = (int)d->vmap4_c.size();
    d->nsblk = (int)d->fmap_c.size();
    IFL_HIP(hipMemcpyAsync(d->vmap, d->vmap_c.data(), d->vmap_c.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
    IFL_HIP(hipMemcpyAsync(d->vmap4, d->vmap4_c.data(), d->vmap4_c.size() * sizeof(int2), hipMemcpyHostToDevice,
                           stream));
    IFL_HIP(hipMemcpyAsync(d->fmap, d->fmap_c.data(), d->fmap_c.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
    IFL_HIP(hipMemcpyAsync(d->bmap, d->bmap_c.data(), d->bmap_c.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
    IFL_HIP(hipStreamSynchronize(stream));  // the host tables are rebuilt in place next time
    d->compacted = active != nullptr;
    return IFL_OK;
}

int batch_gather(Solver* s, BatchDev* d, hipStream_t stream) {
    batch_gather_kernel<<<grid_blocks(d->n, 256), 256, 0, stream>>>(d->st_ptrs, d->n, d->states);
    IFL_LAUNCH_CHECK();
    IFL_HIP(hipMemcpyAsync(d->states_host, d->states, d->n * sizeof(PcgState), hipMemcpyDeviceToHost, stream));
    IFL_HIP(hipStreamSynchronize(stream));
    return IFL_OK;
}

}  // namespace

// The batched project(): the entry (7:854-864) and the exit tails per solver,
// exactly as k_project does them, and the iterations in between batched.
int k_project_batch(Solver* const* ss, int n, hipStream_t stream, BatchDev** dev, ifl_solve_stats* stats,
                    BatchProfile* prof, bool join) {
    Solver* s0 = ss[0];
    const int lv = s0->level;
    if (s0->cfg.precision != IFL_PREC_FP64 || s0->cfg.dot_order != IFL_DOT_TREE) {
        s0->err = "batched solves take the FP64 form and tree-ordered dots";
        return IFL_ERR_UNSUPPORTED;
    }
    BatchDev* d = *dev;
    if (!d || d->solvers.size() != (size_t)n || !std::equal(ss, ss + n, d->solvers.begin())) {
        batch_dev_free(d);
        d = *dev = new BatchDev();
        Solver* s = s0;
        IFL_HIP(hipSetDevice(s0->device));
        int rc = lv == 3 ? batch_build<3, 5, 5>(ss, n, d) : lv <= 5 ? batch_build<4, 5, 6>(ss, n, d)
                                                                     : batch_build<6, 5, 5>(ss, n, d);
        if (rc != IFL_OK) return rc;
    }
    // the solvers' own stages up to the solve are queued on their streams;
    // the batch's stream waits for all of them, then builds every factor and
    // runs every entry of project() (7:854-864) with one launch per phase,
    // and the batched iterations
    int limit = 0;
    Solver* s = s0;  // error reporting
    for (int i = 0; i < n; i++) {
        if (join) {
            IFL_HIP(hipEventRecord(d->ready[i], ss[i]->stream));
            IFL_HIP(hipStreamWaitEvent(stream, d->ready[i], 0));
        }
        limit = std::max(limit, ss[i]->cfg.pcg_limit);
    }
    IFL_TRY(batch_tables(s, d, stream, nullptr));  // every instance iterates at first
    {
        batch_sweep_kernel<SW_PRECON, 0, 3, 3><<<d->nsblk, 64 * kWaves, d->shmem_p, stream>>>(d->pre, d->fmap);
        IFL_LAUNCH_CHECK();
        batch_reset_kernel<<<grid_blocks(n, 256), 256, 0, stream>>>(d->inst, n);
        IFL_LAUNCH_CHECK();
        batch_zero_p_kernel<<<d->nvblk, 256, 0, stream>>>(d->inst, d->vmap);  // 7:854
        IFL_LAUNCH_CHECK();
        if (lv == 3) {                                                        // 7:855
            batch_sweep_kernel<SW_FWD, 0, 3, 4><<<d->nsblk, 64 * kWaves, d->shmem_f0, stream>>>(d->fwd0, d->fmap);
            IFL_LAUNCH_CHECK();
            batch_sweep_kernel<SW_BWD, 0, 3, 5><<<d->nsblk, 64 * kWaves, d->shmem_b0, stream>>>(d->bwd0, d->bmap);
        } else if (lv <= 5) {
            batch_sweep_kernel<SW_FWD, 0, 4, 4><<<d->nsblk, 64 * kWaves, d->shmem_f0, stream>>>(d->fwd0, d->fmap);
            IFL_LAUNCH_CHECK();
            batch_sweep_kernel<SW_BWD, 0, 4, 6><<<d->nsblk, 64 * kWaves, d->shmem_b0, stream>>>(d->bwd0, d->bmap);
        } else {
            batch_sweep_kernel<SW_FWD, 0, 6, 4><<<d->nsblk, 64 * kWaves, d->shmem_f0, stream>>>(d->fwd0, d->fmap);
            IFL_LAUNCH_CHECK();
            batch_sweep_kernel<SW_BWD, 0, 6, 5><<<d->nsblk, 64 * kWaves, d->shmem_b0, stream>>>(d->bwd0, d->bmap);
        }
        IFL_LAUNCH_CHECK();
        batch_copy_s_kernel<<<d->nvblk, 256, 0, stream>>>(d->inst, d->vmap);  // 7:856
        IFL_LAUNCH_CHECK();
        batch_fin_sigma_init_kernel<<<n, 256, 0, stream>>>(d->inst);           // 7:858-864
        IFL_LAUNCH_CHECK();
    }
    int launched = 0, batch = 4, iterating = n;
    std::vector<char> active(n, 1);
    while (true) {
        IFL_TRY(batch_gather(s, d, stream));
        int still = 0;
        for (int i = 0; i < n; i++) {
            if (d->states_host[i].timeout) {
                ss[i]->err = "PCG sweep hand-off timed out";
                for (int k = 0; k < n; k++) (void)edge_reset(ss[k]);
                s0->err = ss[i]->err;
                return IFL_ERR_DEVICE;
            }
            active[i] = !d->states_host[i].done;
            still += active[i];
        }
        if (still == 0 || launched >= limit) break;
        // the finished instances' blocks leave the tables (they would return
        // at once, but each still takes a dispatch slot and, in the sweeps, a
        // whole CU's worth of LDS while it is resident)
        if (still < iterating) {
            IFL_TRY(batch_tables(s, d, stream, &active));
            iterating = still;
        }
        const int nb = std::min(batch, limit - launched);
        for (int k = 0; k < nb; k++) {
            int rc = lv == 3 ? batch_iteration<3, 5, 5>(s, d, stream, launched + k, prof != nullptr)
                   : lv <= 5 ? batch_iteration<4, 5, 6>(s, d, stream, launched + k, prof != nullptr)
                             : batch_iteration<6, 5, 5>(s, d, stream, launched + k, prof != nullptr);
            if (rc != IFL_OK) return rc;
        }
        launched += nb;
        batch = std::min(batch * 2, 64);
    }
    if (prof) {
        IFL_HIP(hipStreamSynchronize(stream));
        for (auto& t : d->timed) {
            float ms = 0;
            IFL_HIP(hipEventElapsedTime(&ms, t.second.first, t.second.second));
            prof->launches[t.first] += 1;
            prof->ms[t.first] += ms;
            d->ev_pool.push_back(t.second.first);
            d->ev_pool.push_back(t.second.second);
        }
        d->timed.clear();
        // the cells of the instances that worked in each fused launch: an
        // instance that converged in iteration m ran m + 1 products and
        // forward sweeps and m backward sweeps; one out of budget ran all
        for (int i = 0; i < n; i++) {
            const PcgState& h = d->states_host[i];
            const double cells = (double)ss[i]->W * ss[i]->H;
            const int f = h.status == IFL_SOLVE_CONVERGED ? h.printed_iter + 1 : h.iter;
            const int b = h.status == IFL_SOLVE_CONVERGED ? h.printed_iter : h.iter;
            prof->cell_iterations[IFL_KERNEL_MATVEC] += cells * f;
            prof->cell_iterations[IFL_KERNEL_SWEEP_FWD] += cells * f;
            prof->cell_iterations[IFL_KERNEL_SWEEP_BWD] += cells * b;
        }
    }
    if (!join) {
        // every tail in one launch over the whole tables
        IFL_TRY(batch_tables(s, d, stream, nullptr));
        batch_tail_kernel<<<d->nvblk, 256, 0, stream>>>(d->inst, d->vmap);
        IFL_LAUNCH_CHECK();
        for (int i = 0; i < n; i++) {
            const PcgState& h = d->states_host[i];
            if (stats) {
                stats[i].status = h.status;
                stats[i].iterations = h.printed_iter;
                stats[i].max_error = h.max_err;
            }
        }
        return IFL_OK;
    }
    // the tails back on the solvers' streams, after the batched iterations
    IFL_HIP(hipEventRecord(d->joined, stream));
    for (int i = 0; i < n; i++) {
        Solver* si = ss[i];
        const PcgState& h = d->states_host[i];
        VecArgs a = vec_args(si);
        long nn = skew_n(si);
        s = si;
        IFL_HIP(hipStreamWaitEvent(si->stream, d->joined, 0));
        // with the folded update, s is where the last product read it; the
        // tails work on that array, and it ends in s
        double* s_last = nullptr;
        if (d->fused) {
            const int worked = h.status == IFL_SOLVE_CONVERGED ? h.printed_iter + 1 : h.iter;
            s_last = s_after(si, worked - 1);
            a.s = s_last;
        }
        if (h.status == IFL_SOLVE_CONVERGED) {
            vec_update_kernel<<<grid_blocks(nn, 256), 256, 0, si->stream>>>(a, 2, 0);  // p += alpha*s (7:869)
            IFL_LAUNCH_CHECK();
            IFL_HIP(hipMemcpyAsync(si->z, si->qv, nn * sizeof(double), hipMemcpyDeviceToDevice, si->stream));
        } else if (h.status == IFL_SOLVE_BUDGET && si->cfg.pcg_limit > 0) {
            vec_update_kernel<<<grid_blocks(nn, 256), 256, 0, si->stream>>>(a, 0, 0);
            IFL_LAUNCH_CHECK();
            if (a.pm) {
                z_nonfluid_kernel<<<grid_blocks(nn, 256), 256, 0, si->stream>>>(a);
                IFL_LAUNCH_CHECK();
            }
        }
        if (s_last && s_last != si->s)
            IFL_HIP(hipMemcpyAsync(si->s, s_last, nn * sizeof(double), hipMemcpyDeviceToDevice, si->stream));
        if (stats) {
            stats[i].status = h.status;
            stats[i].iterations = h.printed_iter;
            stats[i].max_error = h.max_err;
        }
    }
    return IFL_OK;
}

}  // namespace ifl
