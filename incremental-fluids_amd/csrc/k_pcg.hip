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
#include "ifl_solver.hpp"

namespace ifl {

constexpr int kU = 8;  // sweep steps per unrolled block
constexpr uint64_t kSentinel = 0x7FF4DEADBEEF1234ull;  // signalling NaN
constexpr long kSpinLimit = 1L << 18;  // ~0.3 s of polling; a legitimate wait is at most ~1 ms

IFL_DEV uint64_t ld_relaxed(const double* p) {
    return __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
IFL_DEV void st_relaxed(double* p, uint64_t v) {
    __hip_atomic_store((uint64_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
IFL_DEV double as_double(uint64_t v) { return __longlong_as_double((long long)v); }
IFL_DEV uint64_t as_bits(double v) { return (uint64_t)__double_as_longlong(v); }

// Strips are assigned so that consecutive strips share an XCD (workgroups
// are dealt round-robin over the 8 XCDs): the edge hand-off then stays in
// one L2.
IFL_DEV int strip_of_block(int b, int nstrips) {
    int per = (nstrips + 7) >> 3;
    int s = (b & 7) * per + (b >> 3);
    return ((b >> 3) < per && s < nstrips) ? s : -1;
}
static int sweep_blocks(int nstrips) { return 8 * ((nstrips + 7) / 8); }

// Lanes [0, kU) fetch the kU edge values of one block; each value is reset
// to the sentinel once read, ready for the next sweep.
IFL_DEV double edge_fetch(double* e, int x, bool active, int* timeout) {
    uint64_t v = active ? ld_relaxed(e + x) : 0;
    long spins = 0;
    if (__hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) spins = kSpinLimit;
    while (__any(active && v == kSentinel)) {
        if (++spins > kSpinLimit) {
            if (active) atomicOr(timeout, 1);
            v = 0;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (active && v == kSentinel) v = ld_relaxed(e + x);
    }
    if (active) st_relaxed(e + x, kSentinel);
    return as_double(v);
}

IFL_DEV double lane_value(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

struct SweepArgs {
    Skew sk;
    double* r;
    const double* q;
    double* y;
    double* z;
    const double* cx;
    const double* cy;
    const double* pc;
    const uint8_t* flag;
    double* edge;     // nstrips * W
    double* partial;  // per strip
    PcgState* st;
    int mode;         // runtime copy of MODE for logging
};

// ---------------------------------------------------------------- forward
// applyPreconditioner forward loop 7:773-787, fused (MODE 1) with
// scaledAdd(_r, _r, _z, -alpha) 7:870 and infinityNorm(_r) 7:872.
// PM: precondition only FLUID cells (levels 4-8). VM: vector ops only on
// FLUID cells (levels 6-8).
template <bool PM, bool VM, int MODE>
__global__ __launch_bounds__(64) void fwd_sweep_kernel(SweepArgs a) {
    const int j = strip_of_block(blockIdx.x, a.sk.nstrips);
    if (j < 0) return;
    PcgState* st = a.st;
    if ((MODE == 1 && st->done) || st->timeout) return;
    const int l = threadIdx.x;
    const int W = a.sk.W, H = a.sk.H, cols = a.sk.cols;
    const bool row_ok = j * kStrip + l < H;
    const long base = (long)j * a.sk.strip_sz + l;
    const double alpha_neg = MODE == 1 ? -st->alpha : 0.0;
    double* ein = j > 0 ? a.edge + (long)j * W : nullptr;
    double* eout = j + 1 < a.sk.nstrips ? a.edge + (long)(j + 1) * W : nullptr;
    double Pw = 0.0, Pn = 0.0, rmax = 0.0;

    double nr[kU], nq[kU], ncx[kU], ncy[kU], npc[kU];
    uint8_t nf[kU];
    auto load = [&](int T) {
#pragma unroll
        for (int k = 0; k < kU; k++) {
            long p = base + (long)(T + k) * kStrip;
            nr[k] = a.r[p];
            if (MODE == 1) nq[k] = a.q[p];
            ncx[k] = a.cx[p];
            ncy[k] = a.cy[p];
            npc[k] = a.pc[p];
            if (PM || VM) nf[k] = a.flag[p];
        }
    };
    load(0);
    for (int T = 0; T < cols; T += kU) {
        double cr[kU], cq[kU], ccx[kU], ccy[kU], cpc[kU];
        uint8_t cf[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            cr[k] = nr[k]; cq[k] = nq[k]; ccx[k] = ncx[k]; ccy[k] = ncy[k]; cpc[k] = npc[k];
            cf[k] = (PM || VM) ? nf[k] : 1;
        }
        if (T + kU < cols) load(T + kU);
        double eblk = 0.0;
        if (ein) eblk = edge_fetch(ein, T + l, l < kU && T + l < W, &st->timeout);
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int t = T + k;
            const int x = t - l;
            const bool valid = row_ok && x >= 0 && x < W;
            const double ev = ein ? lane_value(eblk, k) : 0.0;
            const double Ps = wave_shr1(Pn, ev);
            double r = cr[k];
            const bool vec_on = valid && (!VM || cf[k]);
            if (MODE == 1) {
                double rn = r + cq[k] * alpha_neg;
                if (vec_on) {
                    r = rn;
                    a.r[base + (long)t * kStrip] = rn;
                }
            }
            if (vec_on) rmax = smax(rmax, fabs(r));
            const double t1 = r - Pw;
            const double tt = t1 - Ps;
            const double zz = tt * cpc[k];
            double pw = ccx[k] * zz;
            double pn = ccy[k] * zz;
            if (PM) {
                pw = cf[k] ? pw : 0.0;
                pn = cf[k] ? pn : 0.0;
            }
            Pw = pw;
            Pn = pn;
            if (valid && (!PM || cf[k])) a.y[base + (long)t * kStrip] = zz;
            if (eout && l == kStrip - 1 && valid) st_relaxed(eout + x, as_bits(pn));
        }
    }
    rmax = wave_max(rmax);
    if (l == 0) atomicMax((unsigned long long*)&st->max_err, (unsigned long long)as_bits(rmax));
}

// ---------------------------------------------------------------- backward
// applyPreconditioner backward loop 7:789-803 fused with dotProduct(_z, _r)
// 7:880. Lane l runs x = W-1 - (t - (63-l)), so lane 63 leads. MODE 1 also
// stores z = A s (held in q) at non-FLUID cells, which is what the
// reference's _z holds there after matrixVectorProduct.
template <bool PM, bool VM, int MODE>
__global__ __launch_bounds__(64) void bwd_sweep_kernel(SweepArgs a) {
    const int j = strip_of_block(blockIdx.x, a.sk.nstrips);
    if (j < 0) return;
    PcgState* st = a.st;
    if ((MODE == 1 && st->done) || st->timeout) return;
    const int l = threadIdx.x;
    const int W = a.sk.W, H = a.sk.H, cols = a.sk.cols;
    const bool row_ok = j * kStrip + l < H;
    const long base = (long)j * a.sk.strip_sz + l;
    double* ein = j + 1 < a.sk.nstrips ? a.edge + (long)(j + 1) * W : nullptr;  // from the strip above
    double* eout = j > 0 ? a.edge + (long)j * W : nullptr;                     // to the strip below
    double zs = 0.0, acc = 0.0;

    double ny_[kU], ncx[kU], ncy[kU], npc[kU], nr[kU];
    uint8_t nf[kU];
    auto load = [&](int T) {
#pragma unroll
        for (int k = 0; k < kU; k++) {
            long c = cols - 1 - (T + k);
            long p = base + c * kStrip;
            ny_[k] = a.y[p];
            ncx[k] = a.cx[p];
            ncy[k] = a.cy[p];
            npc[k] = a.pc[p];
            nr[k] = a.r[p];
            if (PM || VM) nf[k] = a.flag[p];
        }
    };
    load(0);
    for (int T = 0; T < cols; T += kU) {
        double cy_[kU], ccx[kU], ccy[kU], cpc[kU], cr[kU];
        uint8_t cf[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            cy_[k] = ny_[k]; ccx[k] = ncx[k]; ccy[k] = ncy[k]; cpc[k] = npc[k]; cr[k] = nr[k];
            cf[k] = (PM || VM) ? nf[k] : 1;
        }
        if (T + kU < cols) load(T + kU);
        // lane 63 of this strip consumes x = cols-1-(T+k)-63 at step k
        double eblk = 0.0;
        if (ein) {
            int xe = cols - 1 - (T + l) - (kStrip - 1);
            eblk = edge_fetch(ein, xe, l < kU && xe >= 0 && xe < W, &st->timeout);
        }
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const long c = cols - 1 - (T + k);
            const int x = (int)c - l;
            const bool valid = row_ok && x >= 0 && x < W;
            const long p = base + c * kStrip;
            const double ev = ein ? lane_value(eblk, k) : 0.0;
            const double zn = wave_shl1(zs, ev);
            const double t1 = cy_[k] - ccx[k] * zs;
            const double tt = t1 - ccy[k] * zn;
            const double zz = tt * cpc[k];
            const bool fl = cf[k] != 0;
            zs = PM ? (fl ? zz : 0.0) : zz;
            double zdot = zz;
            if (valid) {
                if (!PM || fl) {
                    a.z[p] = zz;
                } else if (MODE == 1) {
                    zdot = a.q[p];
                    a.z[p] = zdot;
                } else {
                    zdot = a.z[p];  // the stale value the reference's _z holds there
                }
                if (!VM || fl) acc += zdot * cr[k];
            }
            if (eout && l == 0 && valid) st_relaxed(eout + x, as_bits(zs));
        }
    }
    acc = wave_sum(acc);
    if (l == 0) a.partial[j] = acc;
}

// ---------------------------------------------------------------- precon
// buildPreconditioner 7:739-768 (3:247-272 with every cell FLUID). A FLUID
// cell passes (px*px + tau*px*py) east and (py*py + tau*px*py) north, with
// px = aPlusX*precon and py = aPlusY*precon of its own; the receiver
// subtracts them in the reference's order. px and py are kept as the
// sweeps' coefficients (aPlusX[i]*precon[i] is the product the sweeps use).
struct PreconArgs {
    Skew sk;
    const double *adiag, *apx, *apy;
    double *pc, *cx, *cy;
    const uint8_t* flag;
    double* edge;
    int* timeout;
    double tau, sigma;
};

__global__ __launch_bounds__(64) void precon_kernel(PreconArgs a) {
    const int j = strip_of_block(blockIdx.x, a.sk.nstrips);
    if (j < 0) return;
    const int l = threadIdx.x;
    const int W = a.sk.W, H = a.sk.H, cols = a.sk.cols;
    const bool row_ok = j * kStrip + l < H;
    const long base = (long)j * a.sk.strip_sz + l;
    double* ein = j > 0 ? a.edge + (long)j * W : nullptr;
    double* eout = j + 1 < a.sk.nstrips ? a.edge + (long)(j + 1) * W : nullptr;
    const double tau = a.tau, sigma = a.sigma;
    double Ew = 0.0, En = 0.0;
    for (int T = 0; T < cols; T += kU) {
        double ad[kU], px_[kU], py_[kU];
        uint8_t f[kU];
#pragma unroll
        for (int k = 0; k < kU; k++) {
            long p = base + (long)(T + k) * kStrip;
            ad[k] = a.adiag[p];
            px_[k] = a.apx[p];
            py_[k] = a.apy[p];
            f[k] = a.flag[p];
        }
        double eblk = 0.0;
        if (ein) eblk = edge_fetch(ein, T + l, l < kU && T + l < W, a.timeout);
#pragma unroll
        for (int k = 0; k < kU; k++) {
            const int t = T + k;
            const int x = t - l;
            const bool valid = row_ok && x >= 0 && x < W;
            const double ev = ein ? lane_value(eblk, k) : 0.0;
            const double Ns = wave_shr1(En, ev);
            double e = ad[k];
            e = e - Ew;
            e = e - Ns;
            if (e < sigma * ad[k]) e = ad[k];
            const double pcv = 1.0 / sqrt(e);
            const double px = px_[k] * pcv;
            const double py = py_[k] * pcv;
            const double eo = px * px + tau * px * py;
            const double no = py * py + tau * px * py;
            const bool on = valid && f[k];
            Ew = on ? eo : 0.0;
            En = on ? no : 0.0;
            if (on) {
                long p = base + (long)t * kStrip;
                a.pc[p] = pcv;
                a.cx[p] = px;
                a.cy[p] = py;
            }
            if (eout && l == kStrip - 1 && valid) st_relaxed(eout + x, as_bits(En));
        }
    }
}

// ---------------------------------------------------------------- vector phases
struct VecArgs {
    Skew sk;
    double *p, *s, *z, *r, *q;
    const double *adiag, *apx, *apy;
    const uint8_t* flag;
    double* partial;
    PcgState* st;
    int vm;
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
__global__ void vec_update_kernel(VecArgs a, int mode, int check_done) {
    PcgState* st = a.st;
    if (check_done && st->done) return;
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x, y;
    if (!skew_cell(a.sk, i, x, y)) return;
    if (a.vm && !a.flag[i]) return;
    double s = a.s[i];
    if (mode != 1) a.p[i] = a.p[i] + s * st->alpha;
    if (mode != 2) a.s[i] = a.z[i] + s * st->beta;
}

// A2: matrixVectorProduct 7:816-833 (unmasked at every level; the four
// neighbour terms in the reference's order, skipped at the domain border)
// fused with the block partial of dotProduct(q, s).
template <int OUT_Q>
__global__ __launch_bounds__(256) void matvec_kernel(VecArgs a, int check_done) {
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
        const int W = sk.W, H = sk.H;
        const double* b = a.s;
        double t = a.adiag[i] * b[i];
        if (x > 0) t += a.apx[i - kStrip] * b[i - kStrip];
        if (y > 0) {
            long js = l > 0 ? i - kStrip - 1 : sk.pos(x, y - 1);
            t += a.apy[js] * b[js];
        }
        if (x < W - 1) t += a.apx[i] * b[i + kStrip];
        if (y < H - 1) {
            long jn = l < kStrip - 1 ? i + kStrip + 1 : sk.pos(x, y + 1);
            t += a.apy[i] * b[jn];
        }
        if (OUT_Q)
            a.q[i] = t;
        else
            a.z[i] = t;
        if (!a.vm || a.flag[i]) prod = t * b[i];
    }
    prod = wave_sum(prod);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Fixed-order sum of `n` partials by one workgroup of 256 threads.
IFL_DEV double block_sum_partials(const double* partial, int n) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) acc += partial[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    return red[0];
}

// dotProduct in the reference's left-to-right order (dot_order = SEQUENTIAL).
IFL_DEV double seq_dot(const Skew& sk, const double* a, const double* b, const uint8_t* flag, int vm) {
    double result = 0.0;
    for (int y = 0; y < sk.H; y++)
        for (int x = 0; x < sk.W; x++) {
            long p = sk.pos(x, y);
            if (!vm || flag[p]) result += a[p] * b[p];
        }
    return result;
}

// finish A2: alpha = sigma / dot(q, s)  (7:868)
__global__ __launch_bounds__(256) void fin_alpha_kernel(VecArgs a, int n, int seq) {
    PcgState* st = a.st;
    if (st->done) return;
    double d;
    if (seq) {
        if (threadIdx.x == 0) st->dot_tmp = seq_dot(a.sk, a.q, a.s, a.flag, a.vm);
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

// finish B: the convergence test of 7:872-876
__global__ void fin_check_kernel(PcgState* st, double tol) {
    if (st->done) return;
    if (st->max_err < tol) {
        st->done = 1;
        st->status = IFL_SOLVE_CONVERGED;
        st->printed_iter = st->iter;
    }
}

// finish C: sigmaNew = dot(z, r), beta = sigmaNew/sigma (7:880-882).
// init=1 is the entry of project(): sigma = dot(z, r) and the early-exit
// test of 7:858-862.
__global__ __launch_bounds__(256) void fin_sigma_kernel(VecArgs a, int n, int seq, int init, double tol, int limit) {
    PcgState* st = a.st;
    if (!init && st->done) return;
    double d;
    if (seq) {
        if (threadIdx.x == 0) st->dot_tmp = seq_dot(a.sk, a.z, a.r, a.flag, a.vm);
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

__global__ void state_reset_kernel(PcgState* st) {
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
        if (threadIdx.x == 0) *out = seq_dot(a.sk, u, v, a.flag, a.vm);
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
    a.vm = s->level >= 6;
    return a;
}

static SweepArgs sweep_args(Solver* s, int which_edge) {
    SweepArgs a;
    a.sk = s->sk;
    a.r = s->r; a.q = s->qv; a.y = s->yv; a.z = s->z;
    a.cx = s->cx; a.cy = s->cy; a.pc = s->precon;
    a.flag = s->dflag;
    a.edge = s->edge + (long)which_edge * s->sk.nstrips * s->W;
    a.partial = s->partial;
    a.st = s->st;
    a.mode = 0;
    return a;
}

template <int MODE>
static int launch_fwd(Solver* s) {
    SweepArgs a = sweep_args(s, 0);
    int nb = sweep_blocks(s->sk.nstrips);
    if (s->level == 3)
        fwd_sweep_kernel<false, false, MODE><<<nb, 64, 0, s->stream>>>(a);
    else if (s->level <= 5)
        fwd_sweep_kernel<true, false, MODE><<<nb, 64, 0, s->stream>>>(a);
    else
        fwd_sweep_kernel<true, true, MODE><<<nb, 64, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

template <int MODE>
static int launch_bwd(Solver* s) {
    SweepArgs a = sweep_args(s, 1);
    int nb = sweep_blocks(s->sk.nstrips);
    if (s->level == 3)
        bwd_sweep_kernel<false, false, MODE><<<nb, 64, 0, s->stream>>>(a);
    else if (s->level <= 5)
        bwd_sweep_kernel<true, false, MODE><<<nb, 64, 0, s->stream>>>(a);
    else
        bwd_sweep_kernel<true, true, MODE><<<nb, 64, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

static long skew_n(Solver* s) { return (long)s->sk.nstrips * s->sk.strip_sz; }

int k_build_precon(Solver* s) {
    PreconArgs a;
    a.sk = s->sk;
    a.adiag = s->adiag; a.apx = s->aplusx; a.apy = s->aplusy;
    a.pc = s->precon; a.cx = s->cx; a.cy = s->cy;
    a.flag = s->dflag;
    a.edge = s->edge + 2L * s->sk.nstrips * s->W;
    a.timeout = &s->st->timeout;
    a.tau = s->cfg.tau;
    a.sigma = s->cfg.sigma;
    precon_kernel<<<sweep_blocks(s->sk.nstrips), 64, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
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

// One iteration's kernels; each returns at once when the solve is done.
static int launch_iteration(Solver* s, int first, int limit) {
    VecArgs a = vec_args(s);
    long n = skew_n(s);
    int nbv = grid_blocks(n, 256);
    int seq = s->cfg.dot_order == IFL_DOT_SEQUENTIAL;
    if (!first) {
        vec_update_kernel<<<nbv, 256, 0, s->stream>>>(a, 0, 1);
        IFL_LAUNCH_CHECK();
    }
    matvec_kernel<1><<<nbv, 256, 0, s->stream>>>(a, 1);
    IFL_LAUNCH_CHECK();
    fin_alpha_kernel<<<1, 256, 0, s->stream>>>(a, nbv, seq);
    IFL_LAUNCH_CHECK();
    IFL_TRY(launch_fwd<1>(s));
    fin_check_kernel<<<1, 1, 0, s->stream>>>(s->st, s->cfg.tol);
    IFL_LAUNCH_CHECK();
    IFL_TRY(launch_bwd<1>(s));
    fin_sigma_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, seq, 0, s->cfg.tol, limit);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// project(limit) 7:853-886
int k_project(Solver* s, int limit, ifl_solve_stats* stats) {
    VecArgs a = vec_args(s);
    long n = skew_n(s);
    int seq = s->cfg.dot_order == IFL_DOT_SEQUENTIAL;
    if (s->profiling) IFL_HIP(hipEventRecord(s->ev0, s->stream));
    state_reset_kernel<<<1, 1, 0, s->stream>>>(s->st);
    IFL_LAUNCH_CHECK();
    IFL_HIP(hipMemsetAsync(s->p, 0, n * sizeof(double), s->stream));  // 7:854
    IFL_TRY(launch_fwd<0>(s));                                         // applyPreconditioner(_z, _r) 7:855
    IFL_TRY(launch_bwd<0>(s));
    IFL_HIP(hipMemcpyAsync(s->s, s->z, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));  // 7:856
    fin_sigma_kernel<<<1, 256, 0, s->stream>>>(a, s->sk.nstrips, seq, 1, s->cfg.tol, limit);  // 7:858-864
    IFL_LAUNCH_CHECK();
    int launched = 0;
    int batch = 4;
    bool first = true;
    while (true) {
        IFL_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, s->stream));
        IFL_HIP(hipStreamSynchronize(s->stream));
        if (s->st_host->done || s->st_host->timeout || launched >= limit) break;
        int nb = std::min(batch, limit - launched);
        for (int k = 0; k < nb; k++) {
            IFL_TRY(launch_iteration(s, first, limit));
            first = false;
        }
        launched += nb;
        batch = std::min(batch * 2, 64);
    }
    PcgState h = *s->st_host;
    if (h.timeout) {
        s->err = "PCG sweep hand-off timed out";
        return IFL_ERR_DEVICE;
    }
    // deferred tails of the last iteration
    if (h.status == IFL_SOLVE_CONVERGED) {
        vec_update_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a, 2, 0);  // p += alpha*s (7:869)
        IFL_LAUNCH_CHECK();
        IFL_HIP(hipMemcpyAsync(s->z, s->qv, n * sizeof(double), hipMemcpyDeviceToDevice, s->stream));  // _z = A s
    } else if (h.status == IFL_SOLVE_BUDGET && limit > 0) {
        vec_update_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a, 0, 0);
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
    } else {
        IFL_HIP(hipStreamSynchronize(s->stream));
    }
    if (stats) {
        stats->status = h.status;
        stats->iterations = h.printed_iter;
        stats->max_error = h.max_err;
        stats->outer = 0;
        stats->inner_total = 0;
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

}  // namespace ifl
