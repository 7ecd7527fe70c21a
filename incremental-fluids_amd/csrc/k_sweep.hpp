// The streamed MIC(0) sweeps: ring layout, the six waves' roles and the
// per-strip body (see k_pcg.hip for the PCG around them). Included by
// k_pcg.hip, whose kernels deal strips grouped by XCD, and by k_sweep_up.hip,
// whose kernels deal them upstream first.
#pragma once
#include <type_traits>

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

// Two ways to deal a launch's strips to its workgroups (blocks are dealt
// round robin over the 8 XCDs and dispatched in block order):
//  order 0: consecutive strips share an XCD (strip (b mod 8) * per + b / 8),
//           the edge hand-offs staying in one L2. A strip's upstream
//           neighbour may then have a higher block id, so every strip of the
//           launch must be resident at once;
//  order 1: upstream first (block b is strip b of a forward sweep, strip
//           n - 1 - b of a backward one), so the lowest unfinished strip can
//           always run and no strip needs the others resident.
// Order 1 is used when the strips outnumber the CUs or when another solver
// is stepping on the device at the same time; alone, order 0 measured
// faster at 2048^2 (forward / backward 384 / 356 us, against 388 / 400 us).
IFL_DEV int strip_of_block(int b, int n, bool reverse, int order) {
    if (order) return b < n ? (reverse ? n - 1 - b : b) : -1;
    int per = (n + 7) >> 3;
    int s = (b & 7) * per + (b >> 3);
    return ((b >> 3) < per && s < n) ? s : -1;
}
static int sweep_blocks(int n, int order) { return order ? n : 8 * ((n + 7) / 8); }

// ---------------------------------------------------------------- streamed sweeps
// One strip runs as a workgroup of six waves. Waves w and w + 4 of a
// workgroup run on the same SIMD (the dispatcher deals a workgroup's waves
// round robin over the CU's four SIMDs), and a wave issues in order, so the
// recurrence gets a SIMD of its own and the light roles share:
//   wave 0        the recurrence. It only does what the dependency chain
//                 needs: operands come from an LDS ring into registers half a
//                 block ahead, FLUID masks come as 64-lane masks through scalar
//                 loads, and results go to an LDS output ring. A single wave
//                 issues one instruction at a time, so everything that is not
//                 on the chain lives in the other waves;
//   wave 1        the writer: stores the results of each finished block to
//                 HBM, hands the edge row to the neighbouring strip, and
//                 forms the reductions (max|r|, the z.r partial);
//   waves 2, 3    loaders: each streams every kLoaders-th block of 8 steps
//                 (4 KB contiguous per array in the skewed layout) from HBM
//                 straight into the LDS ring;
//   wave 4        idle (it would share wave 0's SIMD): returns at once;
//   wave 5        the edge poller (sharing the writer's SIMD): fetches
//                 upcoming inter-strip values from global memory, 64 steps
//                 per round trip, into an LDS ring.
// Measured at 2048^2 against eight waves (five loaders, the poller on wave
// 1): forward sweep 551 -> 498 us, backward 541 -> 508 us.
// Slot life cycle: loader fills -> wave 0 computes -> writer drains -> free.
// Counters in LDS order the hand-offs (LDS instructions of one workgroup are
// executed in order). Every wait is bounded; a timeout raises `abort`, which
// every wave checks, so the workgroup always drains.
//
// The recurrence masks each cell's result to +0 at non-FLUID cells (zm) and
// the factor sweep stores cx = cy = +0 there, so a neighbour the reference
// skips contributes exactly cx*(+0) = +0, the same as a skipped term.
constexpr int kBlk = 8;
constexpr int kLoaders = 2;
constexpr int kWaves = 6;
enum { kWaveCompute = 0, kWaveWriter = 1, kWaveLoader0 = 2, kWaveIdle = 4, kWavePoller = 5 };
constexpr int kERing = 512;  // edge ring, in steps
constexpr int kMaxIn = 6;
constexpr int kMaxOut = 6;
constexpr long kLdsSpin = 1L << 22;

enum { SW_FWD = 0, SW_BWD = 1, SW_PRECON = 2 };

// The first eight counters are what wave 0 polls; they are read together
// (two 16-byte LDS reads, one wait).
struct __align__(16) RingHdr {
    int ready[kLoaders];   // per loader: 1 + last block published
    int edge_ready;        // steps whose incoming edge value is in the edge ring
    int abort;
    int consumed;          // blocks whose slot is free again (writer)
    int computed;          // blocks wave 0 has finished (wave 0)
    int pad[12 - kLoaders];
};
static_assert(sizeof(RingHdr) == 64, "RingHdr layout");

struct StreamArgs {
    Skew sk;
    const void* in[kMaxIn];
    const uint64_t* fmask;
    double* edge;           // this sweep's edge buffers (nstrips * W)
    void* out[kMaxOut];     // fwd: y, r   bwd: z   precon: precon, cx, cy (+ their single-precision copies)
    double* partial;
    PcgState* st;
    int* timeout;
    double tau, sigma;      // the factor build's tau and sigma (7:743-744); the PCG
                            // loop's backward sweep: sigma = the convergence tolerance
    int strip0, nlaunch;    // the strips of this launch
    int pad;                // unused: keeps the fields below where the sweeps were
                            // measured fastest (without it 388 / 363 us per forward /
                            // backward sweep at 2048^2, with it 384 / 357 us, the
                            // same box and code: the layout moves the kernels'
                            // argument loads and their schedule)
    int nb;                 // ring depth in blocks
    int slot_bytes;         // bytes of one ring slot
    long long* trace;       // diagnostics: kTraceStamps stamps per strip and block, or null
};

// s_waitcnt with only vmcnt constrained (gfx9 encoding: vmcnt in bits 3:0
// and 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their maxima).
#define IFL_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 0xF) | (0x7 << 4) | (0xF << 8) | (((n) >> 4) << 14))

typedef __attribute__((address_space(3))) char lds_char;

// LDS counters are read and written through inline asm: the compiler would
// otherwise wait for every in-flight LDS-DMA load (vmcnt(0)) before any LDS
// instruction in the loaders, and it must not move ring accesses across a
// counter access in any wave ("memory" clobber).
IFL_DEV unsigned lds_off(const void* p) { return (unsigned)(size_t)(const __attribute__((address_space(3))) char*)p; }
IFL_DEV int lds_ld(const int* p) {
    int v;
    asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(lds_off(p)) : "memory");
    return __builtin_amdgcn_readfirstlane(v);
}
// ready[0..4], edge_ready, abort, consumed of `h`, wave-uniform
IFL_DEV void lds_ld8(const RingHdr* h, int (&c)[8]) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    v4i lo, hi;
    asm volatile("ds_read_b128 %0, %2\n\tds_read_b128 %1, %2 offset:16\n\ts_waitcnt lgkmcnt(0)"
                 : "=v"(lo), "=v"(hi)
                 : "v"(lds_off(h))
                 : "memory");
    c[0] = __builtin_amdgcn_readfirstlane(lo.x);
    c[1] = __builtin_amdgcn_readfirstlane(lo.y);
    c[2] = __builtin_amdgcn_readfirstlane(lo.z);
    c[3] = __builtin_amdgcn_readfirstlane(lo.w);
    c[4] = __builtin_amdgcn_readfirstlane(hi.x);
    c[5] = __builtin_amdgcn_readfirstlane(hi.y);
    c[6] = __builtin_amdgcn_readfirstlane(hi.z);
    c[7] = __builtin_amdgcn_readfirstlane(hi.w);
}
IFL_DEV void lds_st(int* p, int v) {
    asm volatile("ds_write_b32 %0, %1" : : "v"(lds_off(p)), "v"(v) : "memory");
}
// Spin until *p > bound; false on abort or timeout (which raises abort).
IFL_DEV bool lds_wait_gt(const int* p, int bound, RingHdr* hdr, int* gtimeout) {
    long spins = 0;
    while (lds_ld(p) <= bound) {
        if (lds_ld(&hdr->abort)) return false;
        if (++spins > kLdsSpin) {
            lds_st(&hdr->abort, 1);
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
// skewed column processed at step t, and its index inside its block
template <int KIND>
IFL_DEV int step_col(int t, int cols) {
    return KIND == SW_BWD ? cols - 1 - t : t;
}
template <int KIND>
IFL_DEV int step_in_block(int k) {
    return KIND == SW_BWD ? kBlk - 1 - k : k;
}
// the column the incoming edge value of step t belongs to (lane 0 in the
// forward direction, lane 63 backward)
template <int KIND>
IFL_DEV int edge_x(int t, int cols) {
    return KIND == SW_BWD ? cols - kStrip - t : t;
}

typedef double vec2d __attribute__((ext_vector_type(2)));
typedef float vec2f __attribute__((ext_vector_type(2)));
typedef int vec4i __attribute__((ext_vector_type(4)));
template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef vec2d type; };
template <> struct Vec2<float> { typedef vec2f type; };

IFL_DEV double lane_shr1(double v, double edge) { return wave_shr1(v, edge); }
IFL_DEV float lane_shr1(float v, float edge) { return wave_shr1f(v, edge); }
IFL_DEV double lane_shl1(double v, double edge) { return wave_shl1(v, edge); }
IFL_DEV float lane_shl1(float v, float edge) { return wave_shl1f(v, edge); }

// Layout of one sweep. T is the precision of the triangular solves: double
// in the drop-in, float in the mixed form (the factor build is always
// double). NG arrays are streamed from HBM; NS of them are kept in the LDS
// slot (the forward update folds q into r on the way). MODE 0 is the stage
// form and MODE 1 the PCG loop, where the residual and the vectors the writer
// reduces with stay double and the factor (cx, cy, precon) and the forward
// result y are T; MODE 2 is the mixed form's inner PCG loop, where every
// vector is T too. Each array occupies 512*es bytes of a slot, followed by
// wave 0's output (512 elements).
template <int KIND, int MODE, int NG, typename T>
struct SweepShape {
    static constexpr int NS = NG - (KIND == SW_FWD && MODE >= 1 ? 1 : 0);
    static constexpr int es(int ai) {
        return (MODE == 2 && KIND != SW_PRECON) ? (int)sizeof(T)
               : KIND == SW_FWD ? ((ai == 0 || ai == 4) ? 8 : (int)sizeof(T))
                                : (KIND == SW_BWD ? (ai <= 3 ? (int)sizeof(T) : 8) : 8);
    }
    static constexpr int rb(int ai) { return 512 * es(ai); }
    static constexpr int ro(int ai) { return ai == 0 ? 0 : ro(ai - 1) + rb(ai - 1); }
    static constexpr int lo(int ai) { return ai == 0 ? 0 : lo(ai - 1) + rb(ai - 1) / 1024; }
    static constexpr int NLD = lo(NG);  // 1 KB loads per block
    static constexpr int OUT = ro(NS);  // output region
    static constexpr int SLOT = OUT + 512 * (KIND == SW_PRECON ? 8 : (int)sizeof(T));
    static constexpr int R = NLD <= 16 ? 3 : 2;  // blocks in flight per loader, 4 VGPRs per load
    // arrays wave 0 reads in precision T, and where they are in the slot
    static constexpr int NT = KIND == SW_BWD ? 4 : 3;
    static constexpr int tarr(int i) { return KIND == SW_FWD ? i + 1 : i; }
};

IFL_DEV bool lane_on(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// The masks do not change during a sweep: reading them through the constant
// address space lets wave-uniform reads become scalar loads.
typedef const __attribute__((address_space(4))) uint64_t const_u64;
IFL_DEV const const_u64* strip_masks(const StreamArgs& a, int j) {
    return (const const_u64*)(uintptr_t)(a.fmask + (long)j * a.sk.cols);
}

// Each loader streams blocks k, k + kLoaders, ... with non-temporal loads
// (every byte is read once per sweep; measured at 2048^2 with the writer's
// non-temporal block stores: -40 us per backward sweep). Up to R blocks are in
// flight in its registers (global_load_dwordx4: lane i holds bytes
// [16 i, 16 i + 16) of each 1 KB run of an array's block), so the prefetch
// depth does not depend on free LDS slots; a block is written into its slot
// once the slot is free. In the forward sweep with the residual update
// (MODE 1) the loader forms r + q*(-alpha) (scaledAdd 7:870; at FLUID cells
// only from level 6 on) and keeps only the result.
template <int KIND, int MODE, int LV, int NG, typename T>
IFL_DEV void loader_wave(const StreamArgs& a, char* ring, RingHdr* hdr, int k, int lane, long sbase, int nblk, int j) {
    typedef SweepShape<KIND, MODE, NG, T> S;
    constexpr int R = S::R, NL = S::NLD;
    constexpr bool UPD = KIND == SW_FWD && MODE >= 1;
    const int cols = a.sk.cols;
    const int nmine = nblk > k ? (nblk - k + kLoaders - 1) / kLoaders : 0;
    const double alpha_neg = UPD ? -a.st->alpha : 0.0;
    const float alpha_neg_f = (float)alpha_neg;  // the mixed form's inner loop
    const const_u64* fm = strip_masks(a, j);
    vec4i b0[NL], b1[NL], b2[NL];
    // the FLUID masks the residual update needs (levels 6+), loaded with the
    // block so that their latency is hidden like the block's
    constexpr bool MASKED = UPD && LV >= 6;
    uint64_t w0[4], w1[4], w2[4];
    auto issue = [&](int i, vec4i* B, uint64_t* WM) __attribute__((always_inline)) {
        const int col0 = block_col0<KIND>(k + i * kLoaders, cols);
        const long e0 = sbase + (long)col0 * kStrip;
#pragma unroll
        for (int ai = 0; ai < NG; ai++) {
            const char* p = (const char*)a.in[ai] + e0 * S::es(ai) + lane * 16;
#pragma unroll
            for (int m = 0; m < S::rb(ai) / 1024; m++) B[S::lo(ai) + m] = __builtin_nontemporal_load((const vec4i*)(p + m * 1024));
        }
        if (MASKED && S::es(0) == 8) {
            // element 2*lane + c of run m is step 2m + lane/32, row 2*(lane%32) + c
#pragma unroll
            for (int m = 0; m < 4; m++) WM[m] = fm[col0 + 2 * m + (lane >> 5)];
        } else if (MASKED) {
            // single precision: element 4*lane + c of run m is step 4m + lane/16,
            // row 4*(lane%16) + c
#pragma unroll
            for (int m = 0; m < 2; m++) WM[m] = fm[col0 + 4 * m + (lane >> 4)];
        }
    };
    // land block i (from B) in its slot, then reuse B for block i + R
    auto stage = [&](int i, vec4i* B, uint64_t* WM) __attribute__((always_inline)) -> bool {
        if (i >= nmine) return true;
        const int b = k + i * kLoaders;
        if (UPD && S::es(0) == 8) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const vec2d r = __builtin_bit_cast(vec2d, B[m]);
                vec2d rn = r + __builtin_bit_cast(vec2d, B[S::lo(4) + m]) * alpha_neg;
                if (MASKED) {
                    const unsigned bits = (unsigned)(WM[m] >> (2 * (lane & 31)));
                    rn.x = (bits & 1u) ? rn.x : r.x;
                    rn.y = (bits & 2u) ? rn.y : r.y;
                }
                B[m] = __builtin_bit_cast(vec4i, rn);
            }
        } else if (UPD) {
            typedef float vec4f __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const vec4f r = __builtin_bit_cast(vec4f, B[m]);
                vec4f rn = r + __builtin_bit_cast(vec4f, B[S::lo(4) + m]) * alpha_neg_f;
                if (MASKED) {
                    const unsigned bits = (unsigned)(WM[m] >> (4 * (lane & 15)));
                    rn.x = (bits & 1u) ? rn.x : r.x;
                    rn.y = (bits & 2u) ? rn.y : r.y;
                    rn.z = (bits & 4u) ? rn.z : r.z;
                    rn.w = (bits & 8u) ? rn.w : r.w;
                }
                B[m] = __builtin_bit_cast(vec4i, rn);
            }
        }
        if (!lds_wait_gt(&hdr->consumed, b - a.nb, hdr, a.timeout)) return false;
        char* slot = ring + (long)(b % a.nb) * a.slot_bytes;
#pragma unroll
        for (int ai = 0; ai < S::NS; ai++)
#pragma unroll
            for (int m = 0; m < S::rb(ai) / 1024; m++)
                *(vec4i*)(slot + S::ro(ai) + m * 1024 + lane * 16) = B[S::lo(ai) + m];
        lds_st(&hdr->ready[k], b + 1);
        if (i + R < nmine) issue(i + R, B, WM);
        return true;
    };
    if (nmine > 0) issue(0, b0, w0);
    if (R > 1 && nmine > 1) issue(1, b1, w1);
    if (R > 2 && nmine > 2) issue(2, b2, w2);
    for (int i = 0; i < nmine; i += R) {
        if (!stage(i, b0, w0)) return;
        if (R > 1 && !stage(i + 1, b1, w1)) return;
        if (R > 2 && !stage(i + 2, b2, w2)) return;
    }
}

// Strips without a neighbour on the input side get +0 edges.
template <int KIND>
IFL_DEV void poller_wave(const StreamArgs& a, const double* ein, double* ering, RingHdr* hdr, int lane) {
    const int cols = a.sk.cols, W = a.sk.W;
    int t0 = 0;
    long spins = 0;
    // ring capacity: the edges of finished blocks are no longer read
    int cap = lds_ld(&hdr->computed) * kBlk + kERing;
    while (t0 < cols) {
        const int t = t0 + lane;
        const int x = edge_x<KIND>(t, cols);
        const bool inwin = t < cols && t < cap;
        const bool need = ein && inwin && x >= 0 && x < W;
        uint64_t v = need ? ld_relaxed(ein + x) : 0;
        // while the load is in flight: the abort flag, and the capacity for
        // the next round (read before that round's load: never too large)
        const int ab = lds_ld(&hdr->abort);
        const int cap_next = lds_ld(&hdr->computed) * kBlk + kERing;
        if (ab) return;
        const bool ok = inwin && (!need || v != kSentinel);
        const uint64_t m = __ballot(ok);
        const int n = m == ~0ull ? 64 : (int)__builtin_ctzll(~m);  // ready prefix length
        if (lane < n) ering[t % kERing] = as_double(v);
        cap = cap_next;
        if (n > 0) {
            t0 += n;
            lds_st(&hdr->edge_ready, t0);
            spins = 0;
        } else {
            // no sleep: each poll is a memory round trip already, and the
            // sooner a value is seen the sooner this strip moves (measured at
            // 2048^2: -12 us per forward sweep, -20 us per backward sweep)
            if (++spins > kLdsSpin) {
                lds_st(&hdr->abort, 1);
                atomicOr(a.timeout, 1);
                return;
            }
        }
    }
    // Every incoming edge of this sweep has been read: back to the sentinel
    // for the next sweep of this kind (here rather than inside the poll loop,
    // where the next poll would wait for each store).
    if (ein) {
        double* e = const_cast<double*>(ein);
        for (int x = lane; x < W; x += 64) e[x] = as_double(kSentinel);
    }
}

// Wave 0. Operand registers are double-buffered by block (8 steps): a block
// first issues the LDS reads of the next block, then runs its steps (against
// half-block buffers: 382 / 371 us instead of 408 / 416 us per forward /
// backward sweep at 2048^2, with the trace stamps compiled out).
// TR: the diagnostic build with per-block time stamps (ifl_debug_trace_sweep);
// the stamps' branches cost ~6 % of a sweep, so the solver's own sweeps are
// built without them.
template <int KIND, int MODE, int NG, typename T, bool TR>
IFL_DEV void compute_wave(const StreamArgs& a, char* ring, RingHdr* hdr, const double* ering, int j, int lane) {
    typedef SweepShape<KIND, MODE, NG, T> S;
    typedef typename Vec2<T>::type V2;
    constexpr int NT = S::NT;
    const int cols = a.sk.cols, nblk = cols / kBlk;
    const const_u64* fm = strip_masks(a, j);
    const double tau = a.tau, sigma = a.sigma;
    long long* tr = TR ? a.trace + (long)j * nblk * kTraceStamps : nullptr;
    auto stamp = [&](int b, int i) {
        if (TR && lane == 0) tr[b * kTraceStamps + i] = __builtin_amdgcn_s_memrealtime();
    };

    // Readiness is cached: the counters are re-read only when the cached
    // limit does not cover the request. Loader k has published blocks
    // k, k + kLoaders, ..., ready[k] - 1, so every block below the smallest
    // "next block" over the loaders is in the ring.
    int block_limit = 0, edge_limit = 0;
    // true once block b is in the ring and the edges of steps < tend are in
    // the edge ring
    auto ready = [&](int b, int tend) -> bool {
        if (b < block_limit && tend <= edge_limit) return true;
        long spins = 0;
        while (true) {
            int c[8];
            lds_ld8(hdr, c);
            int lim = 1 << 30;
#pragma unroll
            for (int k = 0; k < kLoaders; k++) lim = min(lim, c[k] == 0 ? k : c[k] - 1 + kLoaders);
            block_limit = lim;
            edge_limit = c[kLoaders];
            if (b < block_limit && tend <= edge_limit) return true;
            if (c[kLoaders + 1]) return false;  // abort
            if (++spins > kLdsSpin) {
                lds_st(&hdr->abort, 1);
                atomicOr(a.timeout, 1);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    };

    // Operands of a block, shaped as the loads deliver them so that the
    // loop-carried registers are the load destinations (no copies, which
    // would wait for the loads): X[i][p] holds steps 2p and 2p+1 of the i-th
    // T array (one ds_read2st64), Q[p] those of the updated residual (forward
    // sweep), E[p] the two edge values of those steps, M0/M1 the eight lane
    // masks in column order (scalar loads). The reads of block b+1 go out at
    // the start of block b.
    typedef uint64_t vec4u __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(4))) vec4u const_u64x4;
    constexpr int NP = kBlk / 2;
    V2 XA[NT][NP], XB[NT][NP];
    // the updated residual: double, or T in the mixed form's inner loop
    typedef typename std::conditional<MODE == 2, T, double>::type QE;
    typedef typename std::conditional<MODE == 2, V2, vec2d>::type QV;
    QV QA[NP], QB[NP];
    vec2d EA[NP], EB[NP];
    vec4u MA0, MA1, MB0, MB1;
    auto load_blk = [&](const char* slot, int b, V2 (&X)[NT][NP], QV (&Q)[NP], vec2d (&E)[NP], vec4u& M0, vec4u& M1) {
        const int bm = b < nblk ? b : nblk - 1;
        // a block's 8 edges are contiguous in the ring (8 divides its size)
        const double* eb = ering + ((unsigned)(bm * kBlk) & (unsigned)(kERing - 1));
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const int e0 = step_in_block<KIND>(2 * p) * kStrip + lane;
            const int e1 = step_in_block<KIND>(2 * p + 1) * kStrip + lane;
#pragma unroll
            for (int i = 0; i < NT; i++) {
                const T* arr = (const T*)(slot + S::ro(S::tarr(i)));
                X[i][p].x = arr[e0];
                X[i][p].y = arr[e1];
            }
            if (KIND == SW_FWD) {
                const QE* arr = (const QE*)slot;
                Q[p].x = arr[e0];
                Q[p].y = arr[e1];
            }
            E[p] = *(const vec2d*)(eb + 2 * p);
        }
        const int c0 = block_col0<KIND>(bm, cols);  // the block's 8 columns, lowest first
        M0 = *(const const_u64x4*)(fm + c0);
        M1 = *(const const_u64x4*)(fm + c0 + 4);
    };
    auto mask_of = [&](const vec4u& M0, const vec4u& M1, int k) -> uint64_t {
        const int i = KIND == SW_BWD ? kBlk - 1 - k : k;
        const vec4u& M = i < 4 ? M0 : M1;
        const int c = i & 3;
        return c == 0 ? M.x : c == 1 ? M.y : c == 2 ? M.z : M.w;
    };
    T Pw = 0, Pn = 0, zs = 0;
    auto run_blk = [&](const char* slot, int b, const V2 (&X)[NT][NP], const QV (&Q)[NP], const vec2d (&E)[NP],
                       const vec4u& M0, const vec4u& M1, const char* nslot, V2 (&Y)[NT][NP], QV (&P)[NP],
                       vec2d (&F)[NP], vec4u& N0, vec4u& N1) {
        T* out = (T*)(slot + S::OUT);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        load_blk(nslot, b + 1, Y, P, F, N0, N1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < kBlk; k++) {
            const int e = step_in_block<KIND>(k) * kStrip + lane;
            const bool f = lane_on(mask_of(M0, M1, k));
            const int p = k >> 1;
            const bool hi = k & 1;
            auto x = [&](int i) -> T { return hi ? X[i][p].y : X[i][p].x; };
            const T ek = (T)(hi ? E[p].y : E[p].x);
            if (KIND == SW_FWD) {
                const T Ps = lane_shr1(Pn, ek);
                const T t1 = (T)(hi ? Q[p].y : Q[p].x) - Pw;
                const T tt = t1 - Ps;
                const T zz = tt * x(2);
                const T zm = f ? zz : (T)0;
                Pw = x(0) * zm;
                Pn = x(1) * zm;
                out[e] = zm;
            } else if (KIND == SW_BWD) {
                const T zn = lane_shl1(zs, ek);
                const T t1 = x(0) - x(1) * zs;
                const T tt = t1 - x(2) * zn;
                const T zz = tt * x(3);
                zs = f ? zz : (T)0;
                out[e] = zs;
            } else {
                const double Ns = lane_shr1((double)Pn, (double)ek);
                const double ad = x(0);
                double ev = ad;
                ev = ev - Pw;
                ev = ev - Ns;
                if (ev < sigma * ad) ev = ad;
                const double pcv = 1.0 / sqrt(ev);
                const double px = x(1) * pcv;
                const double py = x(2) * pcv;
                const double eo = px * px + tau * px * py;
                const double no = py * py + tau * px * py;
                Pw = f ? eo : 0.0;
                Pn = f ? no : 0.0;
                out[e] = pcv;
            }
        }
    };
    {
        const int c0 = block_col0<KIND>(0, cols);
        const vec4u m = *(const const_u64x4*)(fm + c0);
        asm volatile("" : : "s"(m.x));
    }
    if (!ready(0, kBlk)) return;
    const char* slot = ring;  // block 0 is in slot 0
    load_blk(slot, 0, XA, QA, EA, MA0, MA1);
    int sidx = 0;
    auto next_slot = [&]() {
        sidx = sidx + 1 == a.nb ? 0 : sidx + 1;
        return ring + (long)sidx * a.slot_bytes;
    };
    for (int b = 0; b < nblk; b += 2) {
        // block b from A; block b+1 goes to B
        stamp(b, 0);
        if (b + 1 < nblk && !ready(b + 1, (b + 2) * kBlk)) return;
        stamp(b, 1);
        const char* nslot = next_slot();
        run_blk(slot, b, XA, QA, EA, MA0, MA1, nslot, XB, QB, EB, MB0, MB1);
        lds_st(&hdr->computed, b + 1);
        stamp(b, 4);
        slot = nslot;
        if (b + 1 >= nblk) break;
        stamp(b + 1, 0);
        if (b + 2 < nblk && !ready(b + 2, (b + 3) * kBlk)) return;
        stamp(b + 1, 1);
        nslot = next_slot();
        run_blk(slot, b + 1, XB, QB, EB, MB0, MB1, nslot, XA, QA, EA, MA0, MA1);
        lds_st(&hdr->computed, b + 2);
        stamp(b + 1, 4);
        slot = nslot;
    }
}

// The writer wave: drains finished blocks.
template <int KIND, int MODE, int LV, int NG, typename T, bool TR>
IFL_DEV void writer_wave(const StreamArgs& a, char* ring, RingHdr* hdr, int j, int lane, double* eout) {
    typedef SweepShape<KIND, MODE, NG, T> S;
    constexpr bool VM = LV >= 6;   // vector operations masked to FLUID cells
    constexpr bool AUX = LV == 4;  // levels 4-5: the dot reads the reference's _z at non-FLUID cells
    const int W = a.sk.W, H = a.sk.H, cols = a.sk.cols, nblk = cols / kBlk;
    const const_u64* fm = strip_masks(a, j);
    const long sbase = (long)j * a.sk.strip_sz;
    const bool row_ok = j * kStrip + lane < H;
    const double tau = a.tau;
    // the mixed form's pressure solves keep r less its mean over FLUID cells;
    // the convergence test reads r itself
    const double r0 = KIND == SW_FWD && sizeof(T) == 4 ? a.st->rmean : 0.0;
    double rmax = 0.0, acc = 0.0, zsum = 0.0;
    for (int b = 0; b < nblk; b++) {
        if (!lds_wait_gt(&hdr->computed, b, hdr, a.timeout)) return;
        char* slot = ring + (long)(b % a.nb) * a.slot_bytes;
        auto in = [&](int ai) { return slot + S::ro(ai); };  // region of streamed array ai
        const T* out = (const T*)(slot + S::OUT);
        const int col0 = block_col0<KIND>(b, cols);
        const long gblk = sbase + (long)col0 * kStrip;  // first element of the block in HBM
        // the block's 8 masks in one scalar load; the sweeps' boundary rows
        // need none, so there they are loaded after it (scalar loads return
        // out of order: the row's LDS reads would wait for them)
        uint64_t M[kBlk];
        if (KIND == SW_PRECON) {
#pragma unroll
            for (int k = 0; k < kBlk; k++) M[k] = fm[col0 + k];
        }
        // first the boundary row for the neighbouring strip (it waits for it),
        // one step per lane: the same operations on the same operands as
        // wave 0 performed
        if (eout) {
            // lane i < 8 takes step i of the block (the factor build: its
            // column's mask)
            const int t = b * kBlk + (lane & (kBlk - 1));
            const int c = step_col<KIND>(t, cols);
            uint64_t mc = 0;
            if (KIND == SW_PRECON) {
                mc = M[0];
#pragma unroll
                for (int k = 1; k < kBlk; k++) mc = c - col0 == k ? M[k] : mc;
            }
            const int e = step_in_block<KIND>(lane & (kBlk - 1)) * kStrip + (KIND == SW_BWD ? 0 : kStrip - 1);
            const int x = KIND == SW_BWD ? c : c - (kStrip - 1);
            if (lane < kBlk && x >= 0 && x < W) {
                double v;
                if (KIND == SW_FWD) {
                    const T pn = ((const T*)in(2))[e] * out[e];  // Pn = cy * zm of lane 63
                    v = (double)pn;
                } else if (KIND == SW_BWD) {
                    v = (double)out[e];  // zs of lane 0 (+0 at a non-FLUID cell)
                } else {
                    const double px = ((const double*)in(1))[e] * (double)out[e];
                    const double py = ((const double*)in(2))[e] * (double)out[e];
                    v = ((mc >> 63) & 1ull) ? py * py + tau * px * py : 0.0;
                }
                st_relaxed(eout + x, as_bits(v));
            }
        }
        if (KIND != SW_PRECON) {
#pragma unroll
            for (int k = 0; k < kBlk; k++) M[k] = fm[col0 + k];
        }
        // Per step: loads first and selects instead of branches, so that the
        // loop is straight-line code with no masked stores (wave 0's
        // recurrence, on another SIMD of the CU, measurably slows down while
        // this wave runs branchy code or masked stores).
#pragma unroll
        for (int k = 0; k < kBlk; k++) {
            const int t = b * kBlk + k;
            const int c = step_col<KIND>(t, cols);
            const int e = step_in_block<KIND>(k) * kStrip + lane;
            const int x = c - lane;
            const bool valid = row_ok && x >= 0 && x < W;
            const bool f = lane_on(M[c - col0]);
            const bool take = VM ? f : valid;
            if (KIND == SW_FWD) {
                // infinityNorm(_r) 7:871 (masked in 6+) of the residual the sweep used
                const double rv = MODE == 2 ? (double)((const T*)in(0))[e] : ((const double*)in(0))[e];
                const double m = smax(rmax, fabs(sizeof(T) == 4 ? rv + r0 : rv));
                rmax = take ? m : rmax;
            } else if (KIND == SW_BWD) {
                // _z (wave 0 left +0 at non-FLUID cells); dotProduct(_z, _r)
                // 7:880 (unmasked in 3-5, where the reference's _z is A s at
                // non-FLUID cells)
                const double zz = (double)out[e];
                if (MODE == 0) {
                    // the stage form keeps _z at non-FLUID cells as it was
                    if (f) ((double*)a.out[0])[gblk + e] = zz;
                } else if (MODE == 1 && sizeof(T) != 8) {
                    // inside the PCG loop _z at non-FLUID cells is not read
                    // (levels 4-5 take A s from q; the exit restores it), so
                    // every cell is stored; the mixed form widens on the way
                    // (one store per step: widening whole blocks after the
                    // loop measured slower, 430 against 408 us at 2048^2,
                    // from the second read of the block from LDS)
                    ((double*)a.out[0])[gblk + e] = zz;
                }
                const double rv = MODE == 2 ? (double)((const T*)in(4))[e] : ((const double*)in(4))[e];
                const double aux = !AUX ? 0.0 : MODE == 2 ? (double)((const T*)in(5))[e] : ((const double*)in(5))[e];
                const double zdot = f ? zz : aux;
                const double na = acc + zdot * rv;
                acc = take ? na : acc;
                if (MODE == 2 || sizeof(T) == 4)
                    zsum = f ? zsum + zz : zsum;  // the mean taken off z (deflation)
            } else {
                // the factor at FLUID cells; cx = aPlusX*precon, cy likewise, +0 elsewhere
                const double pcv = (double)out[e];
                double* apx = (double*)in(1);
                double* apy = (double*)in(2);
                if (f) ((double*)a.out[0])[gblk + e] = pcv;
                const double px = apx[e] * pcv, py = apy[e] * pcv;
                apx[e] = f ? px : 0.0;
                apy[e] = f ? py : 0.0;
                if (a.out[3]) {  // the mixed form's single-precision copies
                    ((float*)a.out[3])[gblk + e] = f ? (float)pcv : 0.0f;
                    ((float*)a.out[4])[gblk + e] = (float)apx[e];
                    ((float*)a.out[5])[gblk + e] = (float)apy[e];
                }
            }
        }
        // whole-block stores of the unmasked outputs, 1 KB per instruction
        auto store_block = [&](void* dst, const char* src, int es) {
            char* d = (char*)dst + gblk * es + lane * 16;
#pragma unroll
            for (int m = 0; m < 4; m++)
                if (m * 1024 < 512 * es) __builtin_nontemporal_store(*(const vec4i*)(src + m * 1024 + lane * 16), (vec4i*)(d + m * 1024));
        };
        if (KIND == SW_FWD) {
            store_block(a.out[0], (const char*)out, sizeof(T));
            if (MODE >= 1) store_block(a.out[1], in(0), S::es(0));
        } else if (KIND == SW_BWD) {
            if ((MODE == 1 && sizeof(T) == 8) || MODE == 2) store_block(a.out[0], (const char*)out, sizeof(T));
        } else if (KIND == SW_PRECON) {
            store_block(a.out[1], in(1), 8);
            store_block(a.out[2], in(2), 8);
        }
        if (TR && lane == 0) a.trace[((long)j * nblk + b) * kTraceStamps + 5] = __builtin_amdgcn_s_memrealtime();
        lds_st(&hdr->consumed, b + 1);
    }
    if (KIND == SW_FWD) {
        rmax = wave_max(rmax);
        if (lane == 0) atomicMax((unsigned long long*)&a.st->max_err, (unsigned long long)as_bits(rmax));
    } else if (KIND == SW_BWD) {
        acc = wave_sum(acc);
        if (lane == 0) a.partial[j] = acc;
        if (MODE == 2 || sizeof(T) == 4) {
            zsum = wave_sum(zsum);
            if (lane == 0) a.partial[a.sk.nstrips + j] = zsum;
        }
    }
}

template <int KIND, int MODE, int LV, int NG, typename T, bool TR>
IFL_DEV void sweep_body(const StreamArgs& a, int j) {
    extern __shared__ __align__(16) char smem[];
    PcgState* st = a.st;
    if (KIND != SW_PRECON && ((MODE >= 1 && st->done) || st->timeout)) return;
    // inside the PCG loop the backward sweep follows the convergence test
    // (7:872-876): it does nothing once max|r| < tol (a.sigma holds tol)
    if (KIND == SW_BWD && MODE == 1 && st->max_err < a.sigma) return;
    RingHdr* hdr = (RingHdr*)smem;
    double* ering = (double*)(smem + sizeof(RingHdr));
    char* ring = smem + sizeof(RingHdr) + kERing * sizeof(double);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (threadIdx.x < (int)(sizeof(RingHdr) / sizeof(int))) ((int*)hdr)[threadIdx.x] = 0;
    __syncthreads();

    const int W = a.sk.W, nblk = a.sk.cols / kBlk;
    const long sbase = (long)j * a.sk.strip_sz;
    // edge source: the strip below (fwd, precon) or above (bwd)
    const bool has_in = KIND == SW_BWD ? j + 1 < a.sk.nstrips : j > 0;
    const bool has_out = KIND == SW_BWD ? j > 0 : j + 1 < a.sk.nstrips;
    const double* ein = has_in ? a.edge + (long)(KIND == SW_BWD ? j + 1 : j) * W : nullptr;
    double* eout = has_out ? a.edge + (long)(KIND == SW_BWD ? j : j + 1) * W : nullptr;

    if (wave == kWaveCompute)
        compute_wave<KIND, MODE, NG, T, TR>(a, ring, hdr, ering, j, lane);
    else if (wave == kWaveWriter)
        writer_wave<KIND, MODE, LV, NG, T, TR>(a, ring, hdr, j, lane, eout);
    else if (wave == kWavePoller)
        poller_wave<KIND>(a, ein, ering, hdr, lane);
    else if (wave >= kWaveLoader0 && wave < kWaveLoader0 + kLoaders)
        loader_wave<KIND, MODE, LV, NG, T>(a, ring, hdr, wave - kWaveLoader0, lane, sbase, nblk, j);
}

// The sweep kernels that deal a launch's strips upstream first
// (k_sweep_up.hip): one launch of `a.nlaunch` strips with `shmem` bytes of
// LDS on the solver's stream.
template <int KIND, int MODE, int LV, int NG, typename T>
int launch_sweep_up(Solver* s, const StreamArgs& a, size_t shmem);

}  // namespace ifl
