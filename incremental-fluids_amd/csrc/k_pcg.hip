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

constexpr uint64_t kSentinel = 0x7FF4DEADBEEF1234ull;  // signalling NaN

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


// ---------------------------------------------------------------- streamed sweeps
// A sweep step needs ~2.5 KB of fresh data per strip while one wave can keep
// only a few hundred nanoseconds of loads in flight, far less than the HBM
// latency. Each strip therefore runs as a workgroup of
//   wave 0        the recurrence, reading its operands from an LDS ring;
//   wave 1        the edge poller, fetching 64 upcoming inter-strip values
//                 per round trip into a small LDS ring;
//   waves 2..6    loaders, each streaming every kLoaders-th block of 8 steps
//                 (4 KB contiguous per array in the skewed layout) from HBM
//                 into the LDS ring, up to two blocks in flight each.
// Ring slots are recycled through counters in LDS (acquire/release at
// workgroup scope). Every wait is bounded; a timeout raises `abort`, which
// every wave checks, so the workgroup always drains.
constexpr int kBlk = 8;
constexpr int kLoaders = 5;
constexpr int kWaves = 2 + kLoaders;
constexpr int kERing = 512;  // edge ring, in steps
constexpr int kMaxArrays = 6;
constexpr long kLdsSpin = 1L << 22;

enum { SW_FWD = 0, SW_BWD = 1, SW_PRECON = 2 };

struct RingHdr {
    int consumed;            // blocks the compute wave has finished with
    int ready[kLoaders];     // per loader: 1 + last block published
    int edge_ready;          // steps whose edge value is in the ring
    int abort;
    int pad[8];
};

struct StreamArgs {
    Skew sk;
    const double* in[kMaxArrays];
    int na;
    const uint8_t* flag;
    double* edge;
    double* out0;  // fwd: y   bwd: z   precon: precon
    double* out1;  // fwd: r   precon: cx
    double* out2;  // precon: cy
    double* partial;
    PcgState* st;
    int* timeout;
    double tau, sigma;
    int nb;          // ring depth in blocks
    int slot_bytes;  // na*4096 + 512
    long long* trace;  // diagnostics: per strip and block {wait begin, wait end, published}, or null
};

IFL_DEV int lds_acq(int* p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP); }
IFL_DEV void lds_rel(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Spin until *p > bound (or abort); false when aborted.
IFL_DEV bool lds_wait_gt(int* p, int bound, RingHdr* hdr, int* gtimeout) {
    long spins = 0;
    while (lds_acq(p) <= bound) {
        if (lds_acq(&hdr->abort)) return false;
        if (++spins > kLdsSpin) {
            lds_rel(&hdr->abort, 1);
            atomicOr(gtimeout, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// first skewed column of block b (8 columns, ascending in memory)
template <int KIND>
IFL_DEV int block_col0(int b, int cols) {
    return KIND == SW_BWD ? cols - kBlk - b * kBlk : b * kBlk;
}
// the column lane-0 (fwd/precon) or lane-63 (bwd) edge value of step t belongs to
template <int KIND>
IFL_DEV int edge_x(int t, int cols) {
    return KIND == SW_BWD ? cols - kStrip - t : t;
}

// s_waitcnt with only vmcnt constrained (gfx9 encoding: vmcnt in bits 3:0
// and 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their maxima).
#define IFL_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 0xF) | (0x7 << 4) | (0xF << 8) | (((n) >> 4) << 14))

typedef __attribute__((address_space(3))) char lds_char;

// LDS counter access for the loader waves, written as inline asm: the
// compiler would otherwise wait for every in-flight LDS-DMA load (vmcnt(0))
// before any LDS instruction, serialising the load pipeline.
IFL_DEV unsigned lds_off(const void* p) { return (unsigned)(size_t)(const __attribute__((address_space(3))) char*)p; }
IFL_DEV int lds_ld_raw(const int* p) {
    int v;
    asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(lds_off(p)) : "memory");
    return v;
}
IFL_DEV void lds_st_raw(int* p, int v) {
    asm volatile("ds_write_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : : "v"(lds_off(p)), "v"(v) : "memory");
}
IFL_DEV bool lds_wait_gt_raw(const int* p, int bound, RingHdr* hdr, int* gtimeout) {
    long spins = 0;
    while (lds_ld_raw(p) <= bound) {
        if (lds_ld_raw(&hdr->abort)) return false;
        if (++spins > kLdsSpin) {
            lds_st_raw(&hdr->abort, 1);
            atomicOr(gtimeout, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

template <int KIND, int NA>
IFL_DEV void loader_wave(const StreamArgs& a, char* ring, RingHdr* hdr, int k, int lane, long sbase, int nblk) {
    const int cols = a.sk.cols;
    constexpr int kLoadsPerBlock = NA * 4 + 2;
    // Loads land in LDS directly (global_load_lds_dwordx4: lane i fills bytes
    // [16 i, 16 i + 16) of the 1 KB run at the wave-uniform LDS address).
    auto issue = [&](int b) {
        long off = sbase + (long)block_col0<KIND>(b, cols) * kStrip;
        lds_char* slot = (lds_char*)(ring + (long)(b % a.nb) * a.slot_bytes);
#pragma unroll
        for (int ai = 0; ai < NA; ai++)
#pragma unroll
            for (int m = 0; m < 4; m++)
                __builtin_amdgcn_global_load_lds((const void*)(a.in[ai] + off + m * 128 + lane * 2),
                                                 (__attribute__((address_space(3))) void*)(slot + ai * 4096 + m * 1024), 16, 0, 0);
#pragma unroll
        for (int m = 0; m < 2; m++)
            __builtin_amdgcn_global_load_lds((const void*)(a.flag + off + m * 256 + lane * 4),
                                             (__attribute__((address_space(3))) void*)(slot + NA * 4096 + m * 256), 4, 0, 0);
    };
    const int nmine = nblk > k ? (nblk - k + kLoaders - 1) / kLoaders : 0;
    if (nmine == 0) return;
    if (!lds_wait_gt_raw(&hdr->consumed, k - a.nb, hdr, a.timeout)) return;
    issue(k);
    for (int i = 0; i < nmine; i++) {
        const int b = k + i * kLoaders;
        const int bn = b + kLoaders;
        const bool has_next = i + 1 < nmine;
        // Put the next block in flight only if its slot is already free:
        // waiting for it here could wait on the compute wave, which may be
        // waiting for block b.
        bool issued = false;
        if (has_next && lds_ld_raw(&hdr->consumed) > bn - a.nb) {
            issue(bn);
            issued = true;
            IFL_WAIT_VMCNT(kLoadsPerBlock);  // block b has landed, bn may still be in flight
        } else {
            IFL_WAIT_VMCNT(0);
        }
        lds_st_raw(&hdr->ready[k], b + 1);
        if (a.trace && lane == 0) a.trace[((long)blockIdx.x * nblk + b) * 3 + 2] = __builtin_amdgcn_s_memrealtime();
        if (has_next && !issued) {
            if (!lds_wait_gt_raw(&hdr->consumed, bn - a.nb, hdr, a.timeout)) return;
            issue(bn);
        }
    }
    IFL_WAIT_VMCNT(0);
}

template <int KIND>
IFL_DEV void poller_wave(const StreamArgs& a, double* ering, RingHdr* hdr, int lane) {
    const int cols = a.sk.cols, W = a.sk.W;
    int t0 = 0;
    long spins = 0;
    while (t0 < cols) {
        if (lds_acq(&hdr->abort)) return;
        // ring capacity: steps below consumed*kBlk are free
        int cap = lds_acq(&hdr->consumed) * kBlk + kERing;
        int t = t0 + lane;
        int x = edge_x<KIND>(t, cols);
        bool inwin = t < cols && t < cap;
        bool need = inwin && x >= 0 && x < W;
        uint64_t v = need ? ld_relaxed(a.edge + x) : 0;
        bool ok = inwin && (!need || v != kSentinel);
        uint64_t m = __ballot(ok);
        int n = m == ~0ull ? 64 : (int)__builtin_ctzll(~m);  // ready prefix length
        if (lane < n) {
            ering[t % kERing] = as_double(v);
            if (need) st_relaxed(a.edge + x, kSentinel);
        }
        if (n > 0) {
            t0 += n;
            lds_rel(&hdr->edge_ready, t0);
            spins = 0;
        } else {
            if (++spins > kLdsSpin) {
                lds_rel(&hdr->abort, 1);
                atomicOr(a.timeout, 1);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

template <int KIND, bool PM, bool VM, int MODE, int NA>
__global__ __launch_bounds__(64 * kWaves) void stream_sweep_kernel(StreamArgs a) {
    extern __shared__ __align__(16) char smem[];
    const int j = strip_of_block(blockIdx.x, a.sk.nstrips);
    if (j < 0) return;
    PcgState* st = a.st;
    if (KIND != SW_PRECON && ((MODE == 1 && st->done) || st->timeout)) return;
    RingHdr* hdr = (RingHdr*)smem;
    double* ering = (double*)(smem + sizeof(RingHdr));
    char* ring = smem + sizeof(RingHdr) + kERing * sizeof(double);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (threadIdx.x < (int)(sizeof(RingHdr) / sizeof(int))) ((int*)hdr)[threadIdx.x] = 0;
    __syncthreads();

    const int W = a.sk.W, H = a.sk.H, cols = a.sk.cols;
    const int nblk = cols / kBlk;
    const long sbase = (long)j * a.sk.strip_sz;
    // edge source: the strip below (fwd, precon) or above (bwd)
    const bool has_in = KIND == SW_BWD ? j + 1 < a.sk.nstrips : j > 0;
    const bool has_out = KIND == SW_BWD ? j > 0 : j + 1 < a.sk.nstrips;
    double* ein = has_in ? a.edge + (long)(KIND == SW_BWD ? j + 1 : j) * W : nullptr;
    double* eout = has_out ? a.edge + (long)(KIND == SW_BWD ? j : j + 1) * W : nullptr;

    if (wave >= 2) {
        loader_wave<KIND, NA>(a, ring, hdr, wave - 2, lane, sbase, nblk);
        return;
    }
    if (wave == 1) {
        if (ein) {
            StreamArgs b = a;
            b.edge = ein;
            poller_wave<KIND>(b, ering, hdr, lane);
        }
        return;
    }

    // ---- wave 0: the recurrence ----
    const bool row_ok = j * kStrip + lane < H;
    const long gbase = sbase + lane;
    constexpr int na = NA;
    const double alpha_neg = (KIND == SW_FWD && MODE == 1) ? -st->alpha : 0.0;
    const double tau = a.tau, sigma = a.sigma;
    // 4-step rolling window of operands read from the ring (3 steps ahead)
    double w0[4], w1[4], w2[4], w3[4], w4[4], w5[4], wev[4];
    uint8_t wf[4];
    auto rd = [&](int t, int wi) {
        const int b = t >> 3, k = t & 7;
        const char* slot = ring + (long)(b % a.nb) * a.slot_bytes;
        const double* s0 = (const double*)slot;
        const int e = (KIND == SW_BWD ? (kBlk - 1 - k) : k) * kStrip + lane;
        w0[wi] = s0[e];
        w1[wi] = s0[512 + e];
        w2[wi] = s0[1024 + e];
        if (na > 3) w3[wi] = s0[1536 + e];
        if (na > 4) w4[wi] = s0[2048 + e];
        if (na > 5) w5[wi] = s0[2560 + e];
        wf[wi] = ((const uint8_t*)(slot + na * 4096))[e];
        wev[wi] = ein ? ering[t % kERing] : 0.0;
    };
    auto ready = [&](int b) -> bool {
        if (!lds_wait_gt(&hdr->ready[b % kLoaders], b, hdr, a.timeout)) return false;
        if (ein && !lds_wait_gt(&hdr->edge_ready, b * kBlk + kBlk - 1, hdr, a.timeout)) return false;
        return true;
    };

    double Pw = 0.0, Pn = 0.0, rmax = 0.0;  // fwd / precon carries (precon: Ew, En)
    double zs = 0.0, acc = 0.0;             // bwd carries
    if (!ready(0)) return;
    rd(0, 0);
    rd(1, 1);
    rd(2, 2);
    for (int b = 0; b < nblk; b++) {
        if (a.trace && lane == 0) a.trace[((long)blockIdx.x * nblk + b) * 3 + 0] = __builtin_amdgcn_s_memrealtime();
        if (b + 1 < nblk && !ready(b + 1)) return;
        if (a.trace && lane == 0) a.trace[((long)blockIdx.x * nblk + b) * 3 + 1] = __builtin_amdgcn_s_memrealtime();
        const int T = b * kBlk;
#pragma unroll
        for (int k = 0; k < kBlk; k++) {
            const int t = T + k;
            if (t + 3 < cols) rd(t + 3, (k + 3) & 3);
            const int wi = k & 3;
            const bool fl = wf[wi] != 0;
            if (KIND == SW_FWD) {
                // applyPreconditioner forward loop 7:773-787 (+ scaledAdd / infinityNorm in MODE 1)
                const int x = t - lane;
                const bool valid = row_ok && x >= 0 && x < W;
                const long p = gbase + (long)t * kStrip;
                const double Ps = wave_shr1(Pn, wev[wi]);
                double r = w0[wi];
                const bool vec_on = valid && (!VM || fl);
                if (MODE == 1) {
                    double rn = r + w4[wi] * alpha_neg;
                    if (vec_on) {
                        r = rn;
                        a.out1[p] = rn;
                    }
                }
                if (vec_on) rmax = smax(rmax, fabs(r));
                const double t1 = r - Pw;
                const double tt = t1 - Ps;
                const double zz = tt * w3[wi];
                double pw = w1[wi] * zz;
                double pn = w2[wi] * zz;
                if (PM) {
                    pw = fl ? pw : 0.0;
                    pn = fl ? pn : 0.0;
                }
                Pw = pw;
                Pn = pn;
                if (valid && (!PM || fl)) a.out0[p] = zz;
                if (eout && lane == kStrip - 1 && valid) st_relaxed(eout + x, as_bits(pn));
            } else if (KIND == SW_BWD) {
                // applyPreconditioner backward loop 7:789-803 + dotProduct(_z, _r)
                const int c = cols - 1 - t;
                const int x = c - lane;
                const bool valid = row_ok && x >= 0 && x < W;
                const long p = gbase + (long)c * kStrip;
                const double zn = wave_shl1(zs, wev[wi]);
                const double t1 = w0[wi] - w1[wi] * zs;
                const double tt = t1 - w2[wi] * zn;
                const double zz = tt * w3[wi];
                zs = PM ? (fl ? zz : 0.0) : zz;
                if (valid) {
                    double zdot = zz;
                    if (!PM || fl)
                        a.out0[p] = zz;
                    else if (PM && !VM)
                        zdot = w5[wi];  // the reference's _z there: A s (MODE 1) or the stale value (MODE 0)
                    if (!VM || fl) acc += zdot * w4[wi];
                }
                if (eout && lane == 0 && valid) st_relaxed(eout + x, as_bits(zs));
            } else {
                // buildPreconditioner 7:744-767
                const int x = t - lane;
                const bool valid = row_ok && x >= 0 && x < W;
                const double Ns = wave_shr1(Pn, wev[wi]);
                const double ad = w0[wi];
                double e = ad;
                e = e - Pw;
                e = e - Ns;
                if (e < sigma * ad) e = ad;
                const double pcv = 1.0 / sqrt(e);
                const double px = w1[wi] * pcv;
                const double py = w2[wi] * pcv;
                const double eo = px * px + tau * px * py;
                const double no = py * py + tau * px * py;
                const bool on = valid && fl;
                Pw = on ? eo : 0.0;
                Pn = on ? no : 0.0;
                if (on) {
                    const long p = gbase + (long)t * kStrip;
                    a.out0[p] = pcv;
                    a.out1[p] = px;
                    a.out2[p] = py;
                }
                if (eout && lane == kStrip - 1 && valid) st_relaxed(eout + x, as_bits(Pn));
            }
        }
        // every read of slot b has been consumed by now: hand it back
        __hip_atomic_store(&hdr->consumed, b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (KIND == SW_FWD) {
        rmax = wave_max(rmax);
        if (lane == 0) atomicMax((unsigned long long*)&st->max_err, (unsigned long long)as_bits(rmax));
    } else if (KIND == SW_BWD) {
        acc = wave_sum(acc);
        if (lane == 0) a.partial[j] = acc;
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
    int pm;  // levels 4+: the preconditioner writes _z at FLUID cells only
    int vm;  // levels 6+: vector operations are masked to FLUID cells
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
__global__ void vec_update_kernel(VecArgs a, int mode, int check_done) {
    PcgState* st = a.st;
    if (check_done && st->done) return;
    long n = (long)a.sk.nstrips * a.sk.strip_sz;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x, y;
    if (!skew_cell(a.sk, i, x, y)) return;
    const bool fl = a.flag[i] != 0;
    if (a.vm && !fl) return;
    double s = a.s[i];
    if (mode != 1) a.p[i] = a.p[i] + s * st->alpha;
    if (mode != 2) a.s[i] = (a.pm && !fl ? a.q[i] : a.z[i]) + s * st->beta;
}

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
    a.pm = s->level >= 4;
    a.vm = s->level >= 6;
    return a;
}

static StreamArgs stream_args(Solver* s, int kind) {
    StreamArgs a;
    memset(&a, 0, sizeof a);
    a.trace = nullptr;
    a.sk = s->sk;
    a.flag = s->dflag;
    a.edge = s->edge + (long)kind * s->sk.nstrips * s->W;
    a.partial = s->partial;
    a.st = s->st;
    a.timeout = &s->st->timeout;
    a.tau = s->cfg.tau;
    a.sigma = s->cfg.sigma;
    return a;
}

template <int KIND, bool PM, bool VM, int MODE, int NA>
static int launch_stream(Solver* s, StreamArgs a) {
    a.na = NA;
    a.slot_bytes = NA * 4096 + 512;
    size_t fixed = sizeof(RingHdr) + kERing * sizeof(double);
    int nb = (int)((s->lds_bytes - fixed) / a.slot_bytes);
    if (nb > 12) nb = 12;
    if (nb < 2) {
        s->err = "not enough LDS for the sweep ring";
        return IFL_ERR_UNSUPPORTED;
    }
    a.nb = nb;
    size_t shmem = fixed + (size_t)nb * a.slot_bytes;
    static bool attr_set = false;
    if (!attr_set) {
        IFL_HIP(hipFuncSetAttribute((const void*)stream_sweep_kernel<KIND, PM, VM, MODE, NA>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
        attr_set = true;
    }
    stream_sweep_kernel<KIND, PM, VM, MODE, NA><<<sweep_blocks(s->sk.nstrips), 64 * kWaves, shmem, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

template <int MODE>
static int launch_fwd(Solver* s) {
    StreamArgs a = stream_args(s, SW_FWD);
    a.in[0] = s->r; a.in[1] = s->cx; a.in[2] = s->cy; a.in[3] = s->precon; a.in[4] = s->qv;
    a.out0 = s->yv;
    a.out1 = s->r;
    constexpr int NA = MODE == 1 ? 5 : 4;
    if (s->level == 3) return launch_stream<SW_FWD, false, false, MODE, NA>(s, a);
    if (s->level <= 5) return launch_stream<SW_FWD, true, false, MODE, NA>(s, a);
    return launch_stream<SW_FWD, true, true, MODE, NA>(s, a);
}

template <int MODE>
static int launch_bwd(Solver* s) {
    StreamArgs a = stream_args(s, SW_BWD);
    a.in[0] = s->yv; a.in[1] = s->cx; a.in[2] = s->cy; a.in[3] = s->precon; a.in[4] = s->r;
    a.in[5] = MODE == 1 ? s->qv : s->z;  // the reference's _z at non-FLUID cells (levels 4-5 dot)
    a.out0 = s->z;
    if (s->level == 3) return launch_stream<SW_BWD, false, false, MODE, 5>(s, a);
    if (s->level <= 5) return launch_stream<SW_BWD, true, false, MODE, 6>(s, a);
    return launch_stream<SW_BWD, true, true, MODE, 5>(s, a);
}

static long skew_n(Solver* s) { return (long)s->sk.nstrips * s->sk.strip_sz; }

int k_build_precon(Solver* s) {
    StreamArgs a = stream_args(s, SW_PRECON);
    a.in[0] = s->adiag; a.in[1] = s->aplusx; a.in[2] = s->aplusy;
    a.out0 = s->precon;
    a.out1 = s->cx;
    a.out2 = s->cy;
    return launch_stream<SW_PRECON, true, false, 0, 3>(s, a);
}

// Diagnostics: one sweep of `kind` on the current state with the compute
// wave's waits and the loaders' publications time-stamped (100 MHz clock).
int k_trace_sweep(Solver* s, int kind, long long* dev_trace) {
    StreamArgs a;
    if (kind == SW_PRECON) {
        a = stream_args(s, SW_PRECON);
        a.in[0] = s->adiag; a.in[1] = s->aplusx; a.in[2] = s->aplusy;
        a.out0 = s->precon; a.out1 = s->cx; a.out2 = s->cy;
        a.trace = dev_trace;
        return launch_stream<SW_PRECON, true, false, 0, 3>(s, a);
    }
    if (kind == SW_FWD) {
        a = stream_args(s, SW_FWD);
        a.in[0] = s->r; a.in[1] = s->cx; a.in[2] = s->cy; a.in[3] = s->precon; a.in[4] = s->qv;
        a.out0 = s->yv; a.out1 = s->r;
        a.trace = dev_trace;
        return s->level >= 6 ? launch_stream<SW_FWD, true, true, 0, 4>(s, a) : launch_stream<SW_FWD, true, false, 0, 4>(s, a);
    }
    a = stream_args(s, SW_BWD);
    a.in[0] = s->yv; a.in[1] = s->cx; a.in[2] = s->cy; a.in[3] = s->precon; a.in[4] = s->r; a.in[5] = s->z;
    a.out0 = s->z;
    a.trace = dev_trace;
    return s->level >= 6 ? launch_stream<SW_BWD, true, true, 0, 5>(s, a) : launch_stream<SW_BWD, true, false, 0, 6>(s, a);
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
        if (a.pm) {
            z_nonfluid_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
            IFL_LAUNCH_CHECK();
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
