// Particle side of solver 8 (8-flip/Fluid.cpp): particle creation and
// bookkeeping (ParticleQuantities 8:692-940), the particle-to-grid and
// grid-to-particle transfers (fromParticles 8:663-688, gridToParticles
// 8:904-911, copy/diff/undiff 8:374-388), forward particle advection
// (8:816-862, 8:931-939) and the extrapolation that also fills grid cells
// left without particles (8:478-651).
//
// Every result is bit-identical to the reference's sequential loops:
//  * particle-to-grid: each grid node sums its samples in particle-index
//    order. Particles are grouped by the cell whose four corners they feed
//    (a stable radix sort keyed by that cell, values = particle indices), and
//    a node merges the four groups that touch it by particle index.
//  * initParticles / seedParticles: frand() is a linear congruential
//    generator, so any draw follows from the start state by jump-ahead; prefix
//    sums of attempts and acceptances per cell give each attempt its draws and
//    its particle slot.
//  * pruneParticles (swap-with-last): replayed exactly by one thread over the
//    particles of crowded cells only, which are rare.
//  * extrapolate: level-synchronous when no two interior cells without
//    particles touch and each has a fluid neighbour (then the stack order of
//    8:617-646 cannot change any value); otherwise the reference's stack walk
//    is replayed by one thread.
#include <hipcub/hipcub.hpp>

#include "ifl_solver.hpp"

namespace ifl {

constexpr int kMaxPerCell = 12;  // 8:693
constexpr int kMinPerCell = 3;   // 8:695
constexpr int kAvgPerCell = 4;   // 8:697

// ------------------------------------------------------------ frand 8:38-46
// x86 (int) of a double: NaN and out-of-range values give INT_MIN, which the
// reference's range checks then reject (countParticles 8:758, prune 8:771).
IFL_DEV int x86_trunc(double v) {
    return (v == v && v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000u;
}

IFL_DEV float frand_value(uint32_t s) { return __uint_as_float((s >> 8) | 0x3F800000u) - 1.0f; }

// x + frand() with int x: the sum is formed in single precision (8:743, 8:801)
IFL_DEV double jitter(int x, uint32_t s) { return (double)((float)x + frand_value(s)); }

// pointInBody 8:726-732
IFL_DEV bool point_in_body(const Body* __restrict__ bodies, int nb, double hx, double x, double y) {
    for (int i = 0; i < nb; i++)
        if (body_distance(bodies[i], x * hx, y * hx) < 0.0) return true;
    return false;
}

// scalars shared by the particle kernels of one step
struct FlipScalars {
    int count;      // particle count after the last kernel that changes it
    int cand;       // prune: particles in crowded cells
    int moves;      // prune: swaps recorded
    int hard;       // extrapolate: the stack order matters, replay it
    int ready;      // extrapolate: cells initially on the stack
    int executed;   // seed: attempts that drew from frand
    uint32_t rng;   // frand state after the last kernel that draws
    int overflow;   // a bounded work list overflowed (never expected)
};

namespace {

// ------------------------------------------------------------ initParticles 8:735-751
__global__ void init_count_kernel(int W, int H, uint32_t s0, const Body* __restrict__ bodies, int nb, double hx,
                                  int* __restrict__ cnt, uint8_t* __restrict__ bits) {
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)W * H) return;
    int x = (int)(c % W), y = (int)(c / W);
    uint32_t s = lcg_jump(s0, 2ull * kAvgPerCell * c);
    int n = 0;
    uint8_t b = 0;
    for (int i = 0; i < kAvgPerCell; i++) {
        s = lcg_next(s);
        double px = jitter(x, s);
        s = lcg_next(s);
        double py = jitter(y, s);
        if (!point_in_body(bodies, nb, hx, px, py)) {
            b |= (uint8_t)(1u << i);
            n++;
        }
    }
    cnt[c] = n;
    bits[c] = b;
}

// P: the particle state's element type (double, or float in the
// reduced-precision forms; a jittered position is a float sum, exact in both)
template <typename P>
__global__ void init_write_kernel(int W, int H, uint32_t s0, const int* __restrict__ cnt, const int* __restrict__ off,
                                  const uint8_t* __restrict__ bits, P* __restrict__ posx, P* __restrict__ posy,
                                  FlipScalars* sc) {
    long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)W * H;
    if (c >= n) return;
    int x = (int)(c % W), y = (int)(c / W);
    uint32_t s = lcg_jump(s0, 2ull * kAvgPerCell * c);
    int j = off[c];
    uint8_t b = bits[c];
    for (int i = 0; i < kAvgPerCell; i++) {
        s = lcg_next(s);
        double px = jitter(x, s);
        s = lcg_next(s);
        double py = jitter(y, s);
        bool last = (c == n - 1) && (i == kAvgPerCell - 1);
        if (b & (1u << i)) {
            posx[j] = (P)px;
            posy[j] = (P)py;
            j++;
        } else if (last) {
            // a rejected draw stays in the slot past the last particle (8:743-748)
            posx[j] = (P)px;
            posy[j] = (P)py;
        }
    }
    if (c == n - 1) {
        sc->count = off[c] + cnt[c];
        sc->rng = lcg_jump(s0, 2ull * kAvgPerCell * n);
    }
}

// ------------------------------------------------------------ gridToParticles 8:904-911
template <typename P>
struct G2PArgs {
    QtyView q[4];
    const P* posx;
    const P* posy;
    P* prop[4];
    int n;
    int cubic;  // the variant's Catmull-Rom samples (qcerp) instead of lerp
    double alpha;
};

template <typename P>
__global__ void g2p_kernel(G2PArgs<P> a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double x = a.posx[i], y = a.posy[i];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const QtyView& q = a.q[t];
        double p = a.prop[t][i];
        p *= 1.0 - a.alpha;
        p += a.cubic ? qcerp(q.src, q.w, q.h, q.ox, q.oy, x, y) : qlerp(q.src, q.w, q.h, q.ox, q.oy, x, y);
        a.prop[t][i] = (P)p;
    }
}

// ------------------------------------------------------------ copy / diff / undiff 8:374-388
__global__ void copy_old_kernel(const double* __restrict__ src, double* __restrict__ old, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) old[i] = src[i];
}

__global__ void diff_kernel(double* __restrict__ src, const double* __restrict__ old, long n, double alpha, int undo) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = (1.0 - alpha) * old[i];
    src[i] = undo ? src[i] + d : src[i] - d;
}

// ------------------------------------------------------------ countParticles 8:754-763
template <typename P>
__global__ void count_kernel(const P* __restrict__ posx, const P* __restrict__ posy, int n, int W, int H,
                             int* __restrict__ counts) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ix = x86_trunc(posx[i]);
    int iy = x86_trunc(posy[i]);
    if (ix >= 0 && iy >= 0 && ix < W && iy < H) atomicAdd(&counts[ix + iy * W], 1);
}

// ------------------------------------------------------------ pruneParticles 8:766-786
// A particle can only be removed while its cell holds more than 12, and
// counts only fall, so the candidates are the particles of cells that start
// above 12.
template <typename P>
__global__ void prune_flag_kernel(const P* __restrict__ posx, const P* __restrict__ posy, int n, int W, int H,
                                  const int* __restrict__ counts, uint8_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ix = x86_trunc(posx[i]);
    int iy = x86_trunc(posy[i]);
    bool c = ix >= 0 && iy >= 0 && ix < W && iy < H && counts[ix + iy * W] > kMaxPerCell;
    flag[i] = c ? 1 : 0;
}

// The reference loop, replayed over the candidates. Slots at or past the
// shrinking end are only ever read (each is the "last" particle once), so the
// particle found at a slot is the original one until the slot is visited;
// the swaps are recorded and applied afterwards in parallel.
template <typename P>
__global__ void prune_walk_kernel(const P* __restrict__ posx, const P* __restrict__ posy, int n, int W, int H,
                                  int* __restrict__ counts, const int* __restrict__ cand, const int* __restrict__ ncand,
                                  int* __restrict__ mv_dst, int* __restrict__ mv_src, FlipScalars* sc) {
    if (threadIdx.x != 0) return;
    int K = *ncand;
    int end = n, nm = 0;
    bool stop = false;
    for (int k = 0; k < K && !stop; k++) {
        int i = cand[k];
        if (i >= end) break;
        int cur = i;
        for (;;) {
            int ix = x86_trunc(posx[cur]);
            int iy = x86_trunc(posy[cur]);
            if (ix < 0 || iy < 0 || ix >= W || iy >= H) break;
            int c = ix + iy * W;
            if (counts[c] <= kMaxPerCell) break;
            int j = --end;
            counts[c]--;
            if (j == i) {  // the particle at i was the last one: dropped, loop ends
                stop = true;
                break;
            }
            mv_dst[nm] = i;
            mv_src[nm] = j;
            nm++;
            cur = j;
        }
    }
    sc->count = end;
    sc->moves = nm;
}

template <typename P>
struct ParticleArrays {
    P* a[6];  // posx, posy, prop d, t, u, v
};

// Sources lie at or past the final end and destinations below it (or at it,
// for a particle moved in and then dropped), so the copies are independent;
// repeated moves into one slot are consecutive and only the last one counts.
template <typename P>
__global__ void prune_apply_kernel(ParticleArrays<P> pa, const int* __restrict__ mv_dst, const int* __restrict__ mv_src,
                                   const FlipScalars* sc) {
    int nm = sc->moves;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nm; k += gridDim.x * blockDim.x) {
        int d = mv_dst[k];
        if (k + 1 < nm && mv_dst[k + 1] == d) continue;
        int sidx = mv_src[k];
#pragma unroll
        for (int t = 0; t < 6; t++) pa.a[t][d] = pa.a[t][sidx];
    }
}

// ------------------------------------------------------------ seedParticles 8:789-813
__global__ void seed_attempts_kernel(const int* __restrict__ counts, long n, int* __restrict__ att) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = kMinPerCell - counts[i];
    att[i] = a > 0 ? a : 0;
}

// The rejection test looks at particle `idx` (the cell index, 8:802), not at
// the new particle. When every cell index is below the count at the start of
// the pass, that particle is never written by the pass, so a cell accepts or
// rejects all of its attempts at once.
template <typename P>
__global__ void seed_accept_kernel(const P* __restrict__ posx, const P* __restrict__ posy,
                                   const int* __restrict__ att, long n, const Body* __restrict__ bodies, int nb,
                                   double hx, const FlipScalars* sc, int* __restrict__ acc) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = att[i];
    bool ok = a > 0 && (long)sc->count >= n && !point_in_body(bodies, nb, hx, posx[i], posy[i]);
    acc[i] = ok ? a : 0;
}

// attempts that draw: all of them, unless the count reaches the maximum
// (8:795-796), after which the pass returns before drawing again
__global__ void seed_plan_kernel(const int* __restrict__ att, const int* __restrict__ aoff, const int* __restrict__ acc,
                                 const int* __restrict__ joff, long n, int max_particles, FlipScalars* sc) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((long)sc->count < n) return;  // the sequential pass handles this case
    int total_att = aoff[n - 1] + att[n - 1];
    int total_acc = joff[n - 1] + acc[n - 1];
    int mfree = max_particles - sc->count;
    if (i == 0 && (mfree == 0 || total_acc < mfree)) sc->executed = mfree == 0 ? 0 : total_att;
    int a = acc[i];
    if (mfree > 0 && a > 0 && joff[i] < mfree && mfree <= joff[i] + a)
        sc->executed = aoff[i] + (mfree - 1 - joff[i]) + 1;
}

template <typename P>
struct SeedArgs {
    QtyView q[4];
    ParticleArrays<P> pa;
    const int* att;
    const int* aoff;
    const int* acc;
    const int* joff;
    long n;
    int W;
    int max_particles;
    uint32_t s0;
    FlipScalars* sc;
};

template <typename P>
IFL_DEV void seed_props(const SeedArgs<P>& a, int j, double px, double py) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const QtyView& q = a.q[t];
        a.pa.a[2 + t][j] = (P)qlerp(q.src, q.w, q.h, q.ox, q.oy, px, py);
    }
}

template <typename P>
__global__ void seed_write_kernel(SeedArgs<P> a) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    int p0 = a.sc->count;
    if ((long)p0 < a.n) return;
    int nat = a.att[i];
    if (nat == 0) return;
    int executed = a.sc->executed;
    int x = (int)(i % a.W), y = (int)(i / a.W);
    bool accept = a.acc[i] > 0;
    int a0 = a.aoff[i];
    for (int r = 0; r < nat; r++) {
        int at = a0 + r;
        if (at >= executed) break;
        uint32_t s = lcg_jump(a.s0, 2ull * at);
        s = lcg_next(s);
        double px = jitter(x, s);
        s = lcg_next(s);
        double py = jitter(y, s);
        if (accept) {
            int j = p0 + a.joff[i] + r;
            a.pa.a[0][j] = (P)px;
            a.pa.a[1][j] = (P)py;
            seed_props(a, j, px, py);
        } else if (at == executed - 1) {
            // the last draw was rejected: it stays in the slot past the last
            // particle (8:800-803)
            int j = p0 + a.joff[a.n - 1] + a.acc[a.n - 1];
            a.pa.a[0][j] = (P)px;
            a.pa.a[1][j] = (P)py;
        }
    }
}

__global__ void seed_finish_kernel(const int* __restrict__ acc, const int* __restrict__ joff, long n, int max_particles,
                                   uint32_t s0, FlipScalars* sc) {
    if (threadIdx.x != 0 || (long)sc->count < n) return;
    int total_acc = joff[n - 1] + acc[n - 1];
    int mfree = max_particles - sc->count;
    sc->rng = lcg_jump(s0, 2ull * sc->executed);
    sc->count += total_acc < mfree ? total_acc : mfree;
}

// Fewer particles than cells: particle `idx` may be one this pass writes, so
// the pass is replayed in order by one thread.
template <typename P>
__global__ void seed_sequential_kernel(SeedArgs<P> a, const int* __restrict__ counts, const Body* __restrict__ bodies,
                                       int nb, double hx) {
    if (threadIdx.x != 0) return;
    int cnt = a.sc->count;
    if ((long)cnt >= a.n) return;
    uint32_t s = a.s0;
    for (long idx = 0; idx < a.n; idx++) {
        int x = (int)(idx % a.W), y = (int)(idx / a.W);
        for (int i = 0; i < kMinPerCell - counts[idx]; i++) {
            if (cnt == a.max_particles) goto done;
            int j = cnt;
            s = lcg_next(s);
            a.pa.a[0][j] = (P)jitter(x, s);
            s = lcg_next(s);
            a.pa.a[1][j] = (P)jitter(y, s);
            if (point_in_body(bodies, nb, hx, a.pa.a[0][idx], a.pa.a[1][idx])) continue;
            seed_props(a, j, a.pa.a[0][j], a.pa.a[1][j]);
            cnt++;
        }
    }
done:
    a.sc->count = cnt;
    a.sc->rng = s;
}

// ------------------------------------------------------------ fromParticles 8:663-688
// The cell (ix, iy) whose corners a particle feeds, with the reference's clamp.
IFL_DEV void p2g_cell(const QtyView& q, double px, double py, double& x, double& y, int& ix, int& iy) {
    x = px - q.ox;
    y = py - q.oy;
    x = smax(0.5, smin(q.w - 1.5, x));
    y = smax(0.5, smin(q.h - 1.5, y));
    ix = (int)x;
    iy = (int)y;
}

__global__ void p2g_key_kernel(QtyView q, const double* __restrict__ posx, const double* __restrict__ posy, int n,
                               uint32_t* __restrict__ key, int* __restrict__ val, int* __restrict__ hist) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x, y;
    int ix, iy;
    p2g_cell(q, posx[i], posy[i], x, y, ix, iy);
    uint32_t k = (uint32_t)(ix + iy * q.w);
    key[i] = k;
    val[i] = i;
    atomicAdd(&hist[k], 1);
}

struct P2GArgs {
    QtyView q;
    uint8_t* cell2;  // second quantity on the same grid (t with d), or null
    double* src2;
    const double* posx;
    const double* posy;
    const double* prop;
    const double* prop2;
    const int* start;  // group start per cell, n_cells + 1 entries
    const int* order;  // particle indices grouped by cell, ascending within a group
    double* weight;
};

// addSample 8:296-303 for the node (nx, ny), summed over particles in index
// order, then the division / EMPTY marking of 8:682-687.
__global__ void p2g_gather_kernel(P2GArgs a) {
    const QtyView& q = a.q;
    long nid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nid >= (long)q.w * q.h) return;
    int nx = (int)(nid % q.w), ny = (int)(nid / q.w);
    int p[4], e[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int bx = nx - 1 + (k & 1), by = ny - 1 + (k >> 1);
        if (bx >= 0 && by >= 0 && bx <= q.w - 2 && by <= q.h - 2) {
            int b = bx + by * q.w;
            p[k] = a.start[b];
            e[k] = a.start[b + 1];
        } else {
            p[k] = e[k] = 0;
        }
    }
    int hd[4];
#pragma unroll
    for (int k = 0; k < 4; k++) hd[k] = p[k] < e[k] ? a.order[p[k]] : INT_MAX;
    double wsum = 0.0, s1 = 0.0, s2 = 0.0;
    const bool two = a.src2 != nullptr;
    for (;;) {
        int kmin = 0;
#pragma unroll
        for (int k = 1; k < 4; k++)
            if (hd[k] < hd[kmin]) kmin = k;
        int i = hd[kmin];
        if (i == INT_MAX) break;
        double x, y;
        int ix, iy;
        p2g_cell(q, a.posx[i], a.posy[i], x, y, ix, iy);
        double kk = (1.0 - fabs(nx - x)) * (1.0 - fabs(ny - y));
        wsum += kk;
        s1 += kk * a.prop[i];
        if (two) s2 += kk * a.prop2[i];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k == kmin) {
                p[k]++;
                hd[k] = p[k] < e[k] ? a.order[p[k]] : INT_MAX;
            }
        }
    }
    a.weight[nid] = wsum;
    if (wsum != 0.0) {
        q.src[nid] = s1 / wsum;
        if (two) a.src2[nid] = s2 / wsum;
    } else {
        q.src[nid] = s1;
        if (q.cell[nid] == IFL_CELL_FLUID) q.cell[nid] = IFL_CELL_EMPTY;
        if (two) {
            a.src2[nid] = s2;
            if (a.cell2[nid] == IFL_CELL_FLUID) a.cell2[nid] = IFL_CELL_EMPTY;
        }
    }
}

// ------------------------------------------------------------ fromParticles, single precision
// The reduced-precision forms replace the ordered gather by a scatter into
// LDS. particlesToGrid's three transfers (d with t, u, v) run together:
// every particle is binned, for each grid, by the 16x16-cell tile of the
// cell whose corners it feeds (count, scan, scatter of one 16-byte record per
// grid: the cell in the tile, the two fractions, the property); one
// workgroup per tile orders its records by cell in LDS and gathers, for
// each of its 17x17 nodes, the hat weights and weighted values of the
// particles of the four cells around it in single precision, then flushes
// them once: nodes inside the tile belong to it alone (plain
// stores), nodes on its edge are shared with up to three neighbouring tiles
// (global atomics). The order of the additions is not fixed, so the sums
// vary in their last bits from run to run; a node's weight is zero exactly
// when the double weight is (see p2g32_scatter_kernel).
constexpr int kP2GTile = 16;
constexpr int kP2GNodes = kP2GTile + 1;

struct P2G32Args {
    int n;                   // particles
    int nq;                  // grids transferred (1, or 3 for particlesToGrid)
    QtyView q[3];
    const float* prop[3];
    const float* prop2;      // temperature with density (grid 0), or null
    uint8_t* cell2;
    double* src2;
    int ntx[3];              // tiles per row of each grid
    int tile_base[4];        // grid g's tiles are [tile_base[g], tile_base[g + 1])
    const float* posx;
    const float* posy;
    int* tile_cnt;           // all tiles
    int* tile_start;         // all tiles + 1
    int* tile_fill;          // all tiles
    float4* rec;             // per record: cell in the tile (bits), fx, fy, property
    float* rec2;             // temperature of grid 0's records
    float* acc;              // per grid: weight, value, value 2 (3 x `stride` each)
    long stride;
    double* weight;          // the last grid's weights (the reference's _weight after the transfers)
};

// cnt[key] += 1 for every lane with `on`, one atomic per distinct key in the
// wave (neighbouring particles mostly share a tile: per-lane atomics on one
// address serialise); returns the lane's slot, the counter's old value plus
// its rank among the lanes of its key. Every lane of the wave must call it.
IFL_DEV int wave_count(int* cnt, int key, bool on) {
    const int lane = __lane_id();
    uint64_t pending = __ballot(on);
    int slot = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lk = __shfl(key, leader);
        const uint64_t same = __ballot(on && key == lk) & pending;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt[lk], __popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) slot = base + __popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    return slot;
}

IFL_DEV int p2g32_key(const P2G32Args& a, int g, int ix, int iy) {
    return a.tile_base[g] + (iy / kP2GTile) * a.ntx[g] + ix / kP2GTile;
}

__global__ void p2g32_count_kernel(P2G32Args a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < a.n;
    const double px = on ? a.posx[i] : 0.0, py = on ? a.posy[i] : 0.0;
    for (int g = 0; g < a.nq; g++) {
        double x, y;
        int ix, iy;
        p2g_cell(a.q[g], px, py, x, y, ix, iy);
        (void)wave_count(a.tile_cnt, p2g32_key(a, g, ix, iy), on);
    }
}

// The fractions are taken in double (x - ix is exact) and rounded once; a
// fraction that would round up to 1 is held at the largest float below it,
// so every hat factor (1 - f or f) is zero exactly when it is in double.
__global__ void p2g32_scatter_kernel(P2G32Args a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < a.n;
    const double px = on ? a.posx[i] : 0.0, py = on ? a.posy[i] : 0.0;
    const float below1 = 0x1.fffffep-1f;
    for (int g = 0; g < a.nq; g++) {
        double x, y;
        int ix, iy;
        p2g_cell(a.q[g], px, py, x, y, ix, iy);
        const int key = p2g32_key(a, g, ix, iy);
        const int slot = wave_count(a.tile_fill, key, on);
        if (!on) continue;
        const int k = a.tile_start[key] + slot;
        const float fx = fminf((float)(x - ix), below1), fy = fminf((float)(y - iy), below1);
        const int local = (ix % kP2GTile) | ((iy % kP2GTile) << 8);
        a.rec[k] = make_float4(__int_as_float(local), fx, fy, a.prop[g][i]);
        if (g == 0 && a.prop2) a.rec2[k] = a.prop2[i];
    }
}

// One workgroup per tile. The tile's records are taken in chunks of
// kP2GChunk: each thread loads up to kP2GChunk / 256 of them into registers
// and counts them per cell (integer LDS atomics give each its slot in its
// cell), and after a scan of the counts writes them into LDS ordered by
// cell; then each thread gathers, for its nodes, the records of the four
// cells around the node and adds their hat weights and weighted values in
// registers (no floating-point atomics). Measured at 1024^2 (the three
// grids, ~4.1M particles): 138 us with chunks of 1024 records (most tiles
// take two), 150 us with 512 and 196 us with 2048 (fewer workgroups per CU),
// against 460 us with every particle's weights added to the nodes by LDS
// float atomics.
constexpr int kP2GChunk = 1024;

__global__ __launch_bounds__(256) void p2g32_tile_kernel(P2G32Args a) {
    constexpr int NN = kP2GNodes, NC = kP2GTile * kP2GTile;
    constexpr int R = kP2GChunk / 256;
    __shared__ float4 srec[kP2GChunk];
    __shared__ float srec2[kP2GChunk];
    __shared__ int hist[NC + 1], cstart[NC + 1];
    const int t = blockIdx.x;
    const int g = t >= a.tile_base[2] && a.nq > 2 ? 2 : t >= a.tile_base[1] && a.nq > 1 ? 1 : 0;
    const int lt = t - a.tile_base[g];
    const int x0 = (lt % a.ntx[g]) * kP2GTile, y0 = (lt / a.ntx[g]) * kP2GTile;
    const bool two = g == 0 && a.prop2 != nullptr;
    const int b = a.tile_start[t], cnt = a.tile_cnt[t];
    // this thread's nodes: tid and tid + 256 (289 nodes)
    float w[2] = {0.0f, 0.0f}, v[2] = {0.0f, 0.0f}, v2[2] = {0.0f, 0.0f};
    for (int c0 = 0; c0 < cnt; c0 += kP2GChunk) {
        const int m = min(kP2GChunk, cnt - c0);
        for (int k = threadIdx.x; k <= NC; k += blockDim.x) hist[k] = 0;
        __syncthreads();
        float4 rec[R];
        float p2[R];
        int cell[R], slot[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int k = threadIdx.x + r * 256;
            cell[r] = -1;
            slot[r] = 0;
            rec[r] = k < m ? a.rec[b + c0 + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            p2[r] = two && k < m ? a.rec2[b + c0 + k] : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int k = threadIdx.x + r * 256;
            if (k < m) {
                const int local = __float_as_int(rec[r].x);
                cell[r] = (local & 0xFF) + (local >> 8) * kP2GTile;
                slot[r] = atomicAdd(&hist[cell[r]], 1);
            }
        }
        __syncthreads();
        if (threadIdx.x < 64) {  // exclusive scan of the 256 counts by one wave
            const int lane = threadIdx.x;
            int v4[4], sum = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                v4[i] = hist[lane * 4 + i];
                sum += v4[i];
            }
            int incl = sum;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int o = __shfl_up(incl, off);
                if (lane >= off) incl += o;
            }
            int run = incl - sum;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                cstart[lane * 4 + i] = run;
                run += v4[i];
            }
            if (lane == 63) cstart[NC] = run;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++)
            if (cell[r] >= 0) {
                const int at = cstart[cell[r]] + slot[r];
                srec[at] = rec[r];
                if (two) srec2[at] = p2[r];
            }
        __syncthreads();
        // gather: node (lx, ly) takes the cells (lx - 1 .. lx, ly - 1 .. ly)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int node = threadIdx.x + j * 256;
            if (node >= NN * NN) continue;
            const int lx = node % NN, ly = node / NN;
#pragma unroll
            for (int dy = -1; dy <= 0; dy++)
#pragma unroll
                for (int dx = -1; dx <= 0; dx++) {
                    const int cx = lx + dx, cy = ly + dy;
                    if (cx < 0 || cy < 0 || cx >= kP2GTile || cy >= kP2GTile) continue;
                    const int c = cx + cy * kP2GTile;
                    const int q1 = cstart[c + 1];
#pragma unroll 4
                    for (int q = cstart[c]; q < q1; q++) {
                        const float4 r = srec[q];
                        const float hx = dx == 0 ? 1.0f - r.y : r.y;  // node ix or ix + 1
                        const float hy = dy == 0 ? 1.0f - r.z : r.z;
                        const float kk = hx * hy;
                        w[j] += kk;
                        v[j] += kk * r.w;
                        if (two) v2[j] += kk * srec2[q];
                    }
                }
        }
        __syncthreads();
    }
    const QtyView& q = a.q[g];
    float* aw = a.acc + 3 * g * a.stride;
    float* a1 = aw + a.stride;
    float* a2 = aw + 2 * a.stride;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int node = threadIdx.x + j * 256;
        if (node >= NN * NN) continue;
        const int lx = node % NN, ly = node / NN;
        const int gx = x0 + lx, gy = y0 + ly;
        if (gx >= q.w || gy >= q.h) continue;
        const long n = gx + (long)gy * q.w;
        if (lx > 0 && ly > 0 && lx < kP2GTile && ly < kP2GTile) {
            aw[n] = w[j];  // only this tile's particles reach these nodes
            a1[n] = v[j];
            if (two) a2[n] = v2[j];
        } else if (w[j] != 0.0f) {
            atomicAdd(&aw[n], w[j]);
            atomicAdd(&a1[n], v[j]);
            if (two) atomicAdd(&a2[n], v2[j]);
        }
    }
}

// the division or EMPTY marking of 8:682-687, grid blockIdx.y
__global__ void p2g32_finish_kernel(P2G32Args a) {
    const int g = blockIdx.y;
    const QtyView& q = a.q[g];
    const long nodes = (long)q.w * q.h;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodes) return;
    const float* aw = a.acc + 3 * g * a.stride;
    const float w = aw[i], v1 = aw[a.stride + i];
    if (g == a.nq - 1) a.weight[i] = (double)w;
    const bool two = g == 0 && a.src2 != nullptr;
    if (w != 0.0f) {
        q.src[i] = (double)(v1 / w);
        if (two) a.src2[i] = (double)(aw[2 * a.stride + i] / w);
    } else {
        q.src[i] = (double)v1;
        if (q.cell[i] == IFL_CELL_FLUID) q.cell[i] = IFL_CELL_EMPTY;
        if (two) {
            a.src2[i] = (double)aw[2 * a.stride + i];
            if (a.cell2[i] == IFL_CELL_FLUID) a.cell2[i] = IFL_CELL_EMPTY;
        }
    }
}

__global__ void fill_i32_kernel(int* p, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void fill_f32_kernel(float* p, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ------------------------------------------------------------ extrapolate 8:478-651
// fillSolidMask 8:478-513. ready[i] = 1 for the cells the reference pushes
// first (8:614-620); `hard` is raised when an EMPTY cell has no fluid
// neighbour or touches another interior EMPTY cell.
__global__ void flip_mask_kernel(QtyView q, uint8_t* __restrict__ ready, FlipScalars* sc) {
    int w = q.w, h = q.h;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)w * h) return;
    int x = (int)(i % w), y = (int)(i / w);
    if (x == 0 || y == 0 || x == w - 1 || y == h - 1) {
        q.mask[i] = 0xFF;
        ready[i] = 0;
        return;
    }
    uint8_t c = q.cell[i];
    uint8_t m = 0;
    if (c == IFL_CELL_SOLID) {
        double nx = q.nx[i], ny = q.ny[i];
        if (nx != 0.0 && q.cell[i + sgn(nx)] != IFL_CELL_FLUID) m |= 1;
        if (ny != 0.0 && q.cell[i + (long)sgn(ny) * w] != IFL_CELL_FLUID) m |= 2;
    } else if (c == IFL_CELL_EMPTY) {
        uint8_t l = q.cell[i - 1], r = q.cell[i + 1], d = q.cell[i - w], u = q.cell[i + w];
        m = (l != IFL_CELL_FLUID && r != IFL_CELL_FLUID && d != IFL_CELL_FLUID && u != IFL_CELL_FLUID) ? 1 : 0;
        bool touch = (x > 1 && l == IFL_CELL_EMPTY) || (x < w - 2 && r == IFL_CELL_EMPTY) ||
                     (y > 1 && d == IFL_CELL_EMPTY) || (y < h - 2 && u == IFL_CELL_EMPTY);
        if (m == 1 || touch) atomicOr(&sc->hard, 1);
    }
    q.mask[i] = m;
    ready[i] = (c != IFL_CELL_FLUID && m == 0) ? 1 : 0;
}

// extrapolateNormal 8:515-524
IFL_DEV double flip_normal(const QtyView& q, long i) {
    double nx = q.nx[i], ny = q.ny[i];
    double sx = q.src[i + sgn(nx)];
    double sy = q.src[i + (long)sgn(ny) * q.w];
    return (fabs(nx) * sx + fabs(ny) * sy) / (fabs(nx) + fabs(ny));
}

// extrapolateAverage 8:528-545; 0/0 gives the x86 default NaN
IFL_DEV double flip_average(const QtyView& q, long i) {
    double value = 0.0;
    int count = 0;
    long nb[4] = {i - 1, i + 1, i - q.w, i + q.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (q.cell[nb[k]] == IFL_CELL_FLUID) {
            value += q.src[nb[k]];
            count++;
        }
    }
    return count ? value / count : __longlong_as_double((long long)0xFFF8000000000000ull);
}

// neighbour k of i (-1, +1, -w, +w) lists i as its upstream cell
IFL_DEV bool points_at(const QtyView& q, long j, int k) {
    return k == 0 ? q.nx[j] > 0.0 : k == 1 ? q.nx[j] < 0.0 : k == 2 ? q.ny[j] > 0.0 : q.ny[j] < 0.0;
}

// The order-free case: every EMPTY cell is ready at once and reads only fluid
// cells; SOLID cells wait on their upstream cells exactly as in 7:573-587.
__global__ __launch_bounds__(1024) void flip_levels_kernel(QtyView q, int* __restrict__ lists, long cap,
                                                           FlipScalars* sc) {
    if (sc->hard) return;
    __shared__ int s_next;
    int w = q.w;
    int* cur = lists;
    int* nxt = lists + cap;
    int ncur = sc->ready;
    unsigned* mask32 = (unsigned*)q.mask;
    long max_levels = 4L * (q.w + q.h) + 16;
    for (long lvl = 0; ncur > 0 && lvl < max_levels; lvl++) {
        for (int k = threadIdx.x; k < ncur; k += blockDim.x) {
            int c = cur[k];
            q.src[c] = q.cell[c] == IFL_CELL_EMPTY ? flip_average(q, c) : flip_normal(q, c);
        }
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < ncur; k += blockDim.x) {
            int c = cur[k];
            if (q.cell[c] == IFL_CELL_EMPTY) q.cell[c] = IFL_CELL_FLUID;  // 8:631
            long nb[4] = {(long)c - 1, (long)c + 1, (long)c - w, (long)c + w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                long j = nb[e];
                if (!points_at(q, j, e) || q.cell[j] != IFL_CELL_SOLID) continue;  // freeSolidNeighbour 8:547-553
                unsigned bit = (e < 2) ? 1u : 2u;
                int sh = (int)(j & 3) * 8;
                unsigned old = atomicAnd(&mask32[j >> 2], ~(bit << sh));
                if ((((old >> sh) & 0xFFu) & ~bit) == 0) {
                    int slot = atomicAdd(&s_next, 1);
                    if (slot < cap) nxt[slot] = (int)j;
                }
            }
        }
        __syncthreads();
        ncur = s_next < cap ? s_next : (int)cap;
        int* t = cur;
        cur = nxt;
        nxt = t;
        __syncthreads();
    }
}

// The reference's stack walk (8:622-646), one thread, from the initial stack
// in raster order. Every cell enters the stack at most once.
__global__ void flip_stack_kernel(QtyView q, int* __restrict__ stack, long cap, FlipScalars* sc) {
    if (threadIdx.x != 0 || !sc->hard) return;
    int w = q.w;
    long n = sc->ready;
    while (n > 0) {
        long idx = stack[--n];
        if (q.cell[idx] == IFL_CELL_EMPTY) {
            q.src[idx] = flip_average(q, idx);
            q.cell[idx] = IFL_CELL_FLUID;
        } else {
            q.src[idx] = flip_normal(q, idx);
        }
        long nb[4] = {idx - 1, idx + 1, idx - w, idx + w};
        for (int k = 0; k < 4; k++) {
            long j = nb[k];
            if (points_at(q, j, k) && q.cell[j] == IFL_CELL_SOLID) {
                q.mask[j] &= (uint8_t)~(k < 2 ? 1 : 2);
                if (q.mask[j] == 0) {
                    if (n >= cap) {
                        sc->overflow = 1;
                        return;
                    }
                    stack[n++] = (int)j;
                }
            }
        }
        for (int k = 0; k < 4; k++) {
            long j = nb[k];
            if (q.cell[j] == IFL_CELL_EMPTY && q.mask[j] == 1) {  // freeEmptyNeighbour 8:558-565
                q.mask[j] = 0;
                if (n >= cap) {
                    sc->overflow = 1;
                    return;
                }
                stack[n++] = (int)j;
            }
        }
    }
}

// extrapolateEmptyBorders 8:570-609: edges from their inner neighbours, then
// corners from the edges, then every EMPTY cell becomes FLUID.
__global__ __launch_bounds__(256) void flip_border_kernel(QtyView q) {
    int w = q.w, h = q.h;
    double* s = q.src;
    const uint8_t* c = q.cell;
    for (int x = 1 + threadIdx.x; x < w - 1; x += blockDim.x) {
        long it = x, ib = x + (long)(h - 1) * w;
        if (c[it] == IFL_CELL_EMPTY) s[it] = s[it + w];
        if (c[ib] == IFL_CELL_EMPTY) s[ib] = s[ib - w];
    }
    for (int y = 1 + threadIdx.x; y < h - 1; y += blockDim.x) {
        long il = (long)y * w, ir = (long)y * w + w - 1;
        if (c[il] == IFL_CELL_EMPTY) s[il] = s[il + 1];
        if (c[ir] == IFL_CELL_EMPTY) s[ir] = s[ir - 1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long tl = 0, tr = w - 1, bl = (long)(h - 1) * w, br = (long)h * w - 1;
        if (c[tl] == IFL_CELL_EMPTY) s[tl] = 0.5 * (s[tl + 1] + s[tl + w]);
        if (c[tr] == IFL_CELL_EMPTY) s[tr] = 0.5 * (s[tr - 1] + s[tr + w]);
        if (c[bl] == IFL_CELL_EMPTY) s[bl] = 0.5 * (s[bl + 1] + s[bl - w]);
        if (c[br] == IFL_CELL_EMPTY) s[br] = 0.5 * (s[br - 1] + s[br - w]);
    }
}

__global__ void empty_to_fluid_kernel(uint8_t* __restrict__ cell, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cell[i] == IFL_CELL_EMPTY) cell[i] = IFL_CELL_FLUID;
}

// ------------------------------------------------------------ ParticleQuantities::advect 8:931-939
template <typename P>
struct AdvectPArgs {
    QtyView u, v;
    P* posx;
    P* posy;
    int n, W, H;
    double hx, dt;
    const Body* bodies;
    int nb;
};

// backProject 8:816-839
template <typename P>
IFL_DEV void particle_back_project(const AdvectPArgs<P>& a, double& x, double& y) {
    double d = 1e30;
    int closest = -1;
    for (int i = 0; i < a.nb; i++) {
        double id = body_distance(a.bodies[i], x * a.hx, y * a.hx);
        if (id < d) {
            d = id;
            closest = i;
        }
    }
    if (d < -1.0) {
        x *= a.hx;
        y *= a.hx;
        body_closest(a.bodies[closest], x, y);
        double nx, ny;
        body_normal(a.bodies[closest], nx, ny, x, y);
        x -= nx * a.hx;
        y -= ny * a.hx;
        x /= a.hx;
        y /= a.hx;
    }
}

// rungeKutta3 8:844-862, forward in time; the last stage is not divided by
// the cell size (8:857-858)
template <typename P>
__global__ void particle_advect_kernel(AdvectPArgs<P> a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const QtyView &u = a.u, &v = a.v;
    double x = a.posx[i], y = a.posy[i];
    double first_u = qlerp(u.src, u.w, u.h, u.ox, u.oy, x, y) / a.hx;
    double first_v = qlerp(v.src, v.w, v.h, v.ox, v.oy, x, y) / a.hx;
    double mid_x = x + 0.5 * a.dt * first_u;
    double mid_y = y + 0.5 * a.dt * first_v;
    double mid_u = qlerp(u.src, u.w, u.h, u.ox, u.oy, mid_x, mid_y) / a.hx;
    double mid_v = qlerp(v.src, v.w, v.h, v.ox, v.oy, mid_x, mid_y) / a.hx;
    double last_x = x + 0.75 * a.dt * mid_u;
    double last_y = y + 0.75 * a.dt * mid_v;
    double last_u = qlerp(u.src, u.w, u.h, u.ox, u.oy, last_x, last_y);
    double last_v = qlerp(v.src, v.w, v.h, v.ox, v.oy, last_x, last_y);
    x += a.dt * ((2.0 / 9.0) * first_u + (3.0 / 9.0) * mid_u + (4.0 / 9.0) * last_u);
    y += a.dt * ((2.0 / 9.0) * first_v + (3.0 / 9.0) * mid_v + (4.0 / 9.0) * last_v);
    particle_back_project(a, x, y);
    a.posx[i] = (P)smax(smin(x, a.W - 0.001), 0.0);
    a.posy[i] = (P)smax(smin(y, a.H - 0.001), 0.0);
}

int bits_for(long n) {
    int b = 1;
    while ((1L << b) < n) b++;
    return b;
}

// the particle state: double in the drop-in, float in the reduced-precision forms
bool fp32_particles(const Solver* s) { return s->cfg.precision != IFL_PREC_FP64; }
template <typename P> ParticleArrays<P> particle_arrays(Solver* s);
template <> ParticleArrays<double> particle_arrays<double>(Solver* s) {
    ParticleArrays<double> pa;
    pa.a[0] = s->posx;
    pa.a[1] = s->posy;
    for (int t = 0; t < 4; t++) pa.a[2 + t] = s->prop[t];
    return pa;
}
template <> ParticleArrays<float> particle_arrays<float>(Solver* s) {
    ParticleArrays<float> pa;
    pa.a[0] = s->fposx;
    pa.a[1] = s->fposy;
    for (int t = 0; t < 4; t++) pa.a[2 + t] = s->fprop[t];
    return pa;
}

// hipcub calls with the solver's preallocated temporary storage
int cub_exclusive_sum(Solver* s, const int* in, int* out, long n) {
    size_t bytes = s->cub_bytes;
    IFL_HIP(hipcub::DeviceScan::ExclusiveSum(s->cub_tmp, bytes, in, out, (int)n, s->stream));
    return IFL_OK;
}

int cub_select(Solver* s, const uint8_t* flags, int* out, int* nsel, long n) {
    size_t bytes = s->cub_bytes;
    hipcub::CountingInputIterator<int> it(0);
    IFL_HIP(hipcub::DeviceSelect::Flagged(s->cub_tmp, bytes, it, flags, out, nsel, (int)n, s->stream));
    return IFL_OK;
}

int read_scalars(Solver* s) {
    IFL_HIP(hipMemcpyAsync(s->flip_host, s->flip_sc, sizeof(FlipScalars), hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    return IFL_OK;
}

FlipScalars* host_sc(Solver* s) { return (FlipScalars*)s->flip_host; }
FlipScalars* dev_sc(Solver* s) { return (FlipScalars*)s->flip_sc; }

// set the device count and frand state from the host copies
int push_scalars(Solver* s) {
    FlipScalars* h = host_sc(s);
    memset(h, 0, sizeof *h);
    h->count = s->particle_count;
    h->rng = frand_state();
    IFL_HIP(hipMemcpyAsync(s->flip_sc, h, sizeof(FlipScalars), hipMemcpyHostToDevice, s->stream));
    return IFL_OK;
}

}  // namespace

// Temporary storage for the radix sort, scans and selections, sized for the
// largest calls this solver makes.
int flip_alloc(Solver* s) {
    long mp = s->max_particles;
    long cells = (long)(s->W + 1) * (s->H + 1) + 1;
    IFL_HIP(hipMalloc((void**)&s->pkey, mp * sizeof(uint32_t)));
    IFL_HIP(hipMalloc((void**)&s->pkey2, mp * sizeof(uint32_t)));
    IFL_HIP(hipMalloc((void**)&s->pidx, mp * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->pidx2, mp * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->pflag, mp));
    IFL_HIP(hipMalloc((void**)&s->cbuf0, cells * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->cbuf1, cells * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->cbuf2, cells * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->cbuf3, cells * sizeof(int)));
    IFL_HIP(hipMalloc((void**)&s->cflag, cells));
    IFL_HIP(hipMalloc((void**)&s->flip_sc, sizeof(FlipScalars)));
    IFL_HIP(hipMemsetAsync(s->flip_sc, 0, sizeof(FlipScalars), s->stream));
    IFL_HIP(hipHostMalloc((void**)&s->flip_host, sizeof(FlipScalars), hipHostMallocDefault));
    size_t need = 0, b = 0;
    int n = (int)mp;
    IFL_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, s->pkey, s->pkey2, s->pidx, s->pidx2, n, 0, 32, s->stream));
    need = std::max(need, b);
    long nscan = std::max(mp, cells);
    IFL_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, s->pidx, s->pidx2, (int)nscan, s->stream));
    need = std::max(need, b);
    hipcub::CountingInputIterator<int> it(0);
    IFL_HIP(hipcub::DeviceSelect::Flagged(nullptr, b, it, s->pflag, s->pidx, s->cbuf0, (int)nscan, s->stream));
    need = std::max(need, b);
    s->cub_bytes = need;
    IFL_HIP(hipMalloc(&s->cub_tmp, need));
    if (s->cfg.precision != IFL_PREC_FP64) {
        IFL_HIP(hipMalloc((void**)&s->rec32, 3 * mp * sizeof(float4)));
        IFL_HIP(hipMalloc((void**)&s->rec32b, 3 * mp * sizeof(float)));
        IFL_HIP(hipMalloc((void**)&s->acc32, 9 * cells * sizeof(float)));
    }
    return IFL_OK;
}

void flip_free(Solver* s) {
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    F(s->pkey); F(s->pkey2); F(s->pidx); F(s->pidx2); F(s->pflag);
    F(s->cbuf0); F(s->cbuf1); F(s->cbuf2); F(s->cbuf3); F(s->cflag);
    F(s->flip_sc); F(s->cub_tmp); F(s->rec32); F(s->rec32b); F(s->acc32);
    if (s->flip_host) (void)hipHostFree(s->flip_host);
}

// ParticleQuantities ctor 8:865-878 (initParticles) and the FluidSolver
// ctor's gridToParticles(1.0) (8:1314)
int flip_init_particles(Solver* s) {
    long n = (long)s->W * s->H;
    uint32_t s0 = frand_state();
    int* cnt = s->cbuf0;
    int* off = s->cbuf1;
    uint8_t* bits = s->cflag;
    init_count_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->W, s->H, s0, s->bodies_d, s->nbodies, s->hx, cnt,
                                                                   bits);
    IFL_LAUNCH_CHECK();
    IFL_TRY(cub_exclusive_sum(s, cnt, off, n));
    if (fp32_particles(s)) {
        init_write_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->W, s->H, s0, cnt, off, bits, s->fposx,
                                                                       s->fposy, dev_sc(s));
    } else {
        init_write_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->W, s->H, s0, cnt, off, bits, s->posx,
                                                                       s->posy, dev_sc(s));
    }
    IFL_LAUNCH_CHECK();
    IFL_TRY(read_scalars(s));
    s->particle_count = host_sc(s)->count;
    frand_set_state(host_sc(s)->rng);
    IFL_TRY(k_grid_to_particles(s, 1.0));
    s->particles_ready = true;
    return IFL_OK;
}

template <typename P>
static int grid_to_particles(Solver* s, double alpha, bool cubic) {
    int n = s->particle_count;
    if (n == 0) return IFL_OK;
    const ParticleArrays<P> pa = particle_arrays<P>(s);
    G2PArgs<P> a;
    for (int t = 0; t < 4; t++) {
        a.q[t] = s->q[t];
        a.prop[t] = pa.a[2 + t];
    }
    a.posx = pa.a[0];
    a.posy = pa.a[1];
    a.n = n;
    a.cubic = cubic ? 1 : 0;
    a.alpha = alpha;
    IFL_TRY(prof_begin(s, IFL_KERNEL_G2P, -1));
    g2p_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_end(s));
    return IFL_OK;
}

int k_grid_to_particles(Solver* s, double alpha, bool cubic) {
    return fp32_particles(s) ? grid_to_particles<float>(s, alpha, cubic) : grid_to_particles<double>(s, alpha, cubic);
}

int k_copy_old(Solver* s, int qi) {
    const QtyView& q = s->q[qi];
    long n = (long)q.w * q.h;
    copy_old_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(q.src, q.dst, n);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

int k_diff(Solver* s, int qi, double alpha, int undo) {
    const QtyView& q = s->q[qi];
    long n = (long)q.w * q.h;
    diff_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(q.src, q.dst, n, alpha, undo);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

int k_count_particles(Solver* s) {
    long cells = (long)s->W * s->H;
    IFL_HIP(hipMemsetAsync(s->counts, 0, cells * sizeof(int), s->stream));
    int n = s->particle_count;
    if (n == 0) return IFL_OK;
    if (fp32_particles(s))
        count_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->fposx, s->fposy, n, s->W, s->H, s->counts);
    else
        count_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->posx, s->posy, n, s->W, s->H, s->counts);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

template <typename P>
static int prune(Solver* s) {
    int n = s->particle_count;
    if (n == 0) return IFL_OK;
    IFL_TRY(push_scalars(s));
    const ParticleArrays<P> pa = particle_arrays<P>(s);
    prune_flag_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(pa.a[0], pa.a[1], n, s->W, s->H, s->counts,
                                                                   s->pflag);
    IFL_LAUNCH_CHECK();
    int* cand = s->pidx;
    int* ncand = s->cbuf0;
    IFL_TRY(cub_select(s, s->pflag, cand, ncand, n));
    prune_walk_kernel<<<1, 64, 0, s->stream>>>(pa.a[0], pa.a[1], n, s->W, s->H, s->counts, cand, ncand, s->pidx2,
                                               (int*)s->pkey, dev_sc(s));
    IFL_LAUNCH_CHECK();
    prune_apply_kernel<<<256, 256, 0, s->stream>>>(pa, s->pidx2, (const int*)s->pkey, dev_sc(s));
    IFL_LAUNCH_CHECK();
    s->prune_pending = true;
    return IFL_OK;
}

int k_prune_finish(Solver* s) {
    if (!s->prune_pending) return IFL_OK;
    s->prune_pending = false;
    IFL_TRY(read_scalars(s));
    s->particle_count = host_sc(s)->count;
    return IFL_OK;
}

int k_prune(Solver* s) {
    IFL_TRY(fp32_particles(s) ? prune<float>(s) : prune<double>(s));
    return k_prune_finish(s);
}

template <typename P>
static int seed(Solver* s) {
    long n = (long)s->W * s->H;
    const ParticleArrays<P> pa = particle_arrays<P>(s);
    IFL_TRY(push_scalars(s));
    uint32_t s0 = frand_state();
    int* att = s->cbuf0;
    int* aoff = s->cbuf1;
    int* acc = s->cbuf2;
    int* joff = s->cbuf3;
    seed_attempts_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->counts, n, att);
    IFL_LAUNCH_CHECK();
    IFL_TRY(cub_exclusive_sum(s, att, aoff, n));
    seed_accept_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(pa.a[0], pa.a[1], att, n, s->bodies_d, s->nbodies,
                                                                    s->hx, dev_sc(s), acc);
    IFL_LAUNCH_CHECK();
    IFL_TRY(cub_exclusive_sum(s, acc, joff, n));
    seed_plan_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(att, aoff, acc, joff, n, s->max_particles, dev_sc(s));
    IFL_LAUNCH_CHECK();
    SeedArgs<P> a;
    for (int t = 0; t < 4; t++) a.q[t] = s->q[t];
    a.pa = pa;
    a.att = att;
    a.aoff = aoff;
    a.acc = acc;
    a.joff = joff;
    a.n = n;
    a.W = s->W;
    a.max_particles = s->max_particles;
    a.s0 = s0;
    a.sc = dev_sc(s);
    seed_write_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    seed_finish_kernel<<<1, 64, 0, s->stream>>>(acc, joff, n, s->max_particles, s0, dev_sc(s));
    IFL_LAUNCH_CHECK();
    seed_sequential_kernel<<<1, 64, 0, s->stream>>>(a, s->counts, s->bodies_d, s->nbodies, s->hx);
    IFL_LAUNCH_CHECK();
    IFL_TRY(read_scalars(s));
    s->particle_count = host_sc(s)->count;
    frand_set_state(host_sc(s)->rng);
    return IFL_OK;
}

int k_seed(Solver* s) { return fp32_particles(s) ? seed<float>(s) : seed<double>(s); }

// fromParticles of the grids qis[0..nq) (with t on d's grid when with_t),
// from the float particle state
static int k_from_particles32(Solver* s, const int* qis, int nq, bool with_t) {
    P2G32Args a;
    a.n = s->particle_count;
    a.nq = nq;
    a.tile_base[0] = 0;
    long most = 0;
    for (int g = 0; g < 3; g++) {
        const int qi = g < nq ? qis[g] : qis[0];
        a.q[g] = s->q[qi];
        a.prop[g] = s->fprop[qi];
        a.ntx[g] = (a.q[g].w + kP2GTile - 1) / kP2GTile;
        const int nty = (a.q[g].h + kP2GTile - 1) / kP2GTile;
        a.tile_base[g + 1] = a.tile_base[g] + (g < nq ? a.ntx[g] * nty : 0);
        most = std::max(most, (long)a.q[g].w * a.q[g].h);
    }
    const int ntiles = a.tile_base[3];
    const bool t0 = with_t && qis[0] == IFL_Q_D;
    a.prop2 = t0 ? s->fprop[IFL_Q_T] : nullptr;
    a.cell2 = t0 ? s->q[IFL_Q_T].cell : nullptr;
    a.src2 = t0 ? s->q[IFL_Q_T].src : nullptr;
    a.posx = s->fposx;
    a.posy = s->fposy;
    a.tile_cnt = s->cbuf0;
    a.tile_start = s->cbuf1;
    a.tile_fill = s->cbuf2;
    a.rec = s->rec32;
    a.rec2 = s->rec32b;
    a.acc = s->acc32;
    a.stride = (long)(s->W + 1) * (s->H + 1) + 1;
    a.weight = s->weight;
    IFL_TRY(prof_begin(s, IFL_KERNEL_P2G, -1));
    fill_i32_kernel<<<grid_blocks(ntiles + 1, 256), 256, 0, s->stream>>>(s->cbuf0, ntiles + 1);
    IFL_LAUNCH_CHECK();
    fill_i32_kernel<<<grid_blocks(ntiles, 256), 256, 0, s->stream>>>(s->cbuf2, ntiles);
    IFL_LAUNCH_CHECK();
    fill_f32_kernel<<<grid_blocks(3L * nq * a.stride, 256), 256, 0, s->stream>>>(s->acc32, 3L * nq * a.stride);
    IFL_LAUNCH_CHECK();
    if (a.n > 0) {
        p2g32_count_kernel<<<grid_blocks(a.n, 256), 256, 0, s->stream>>>(a);
        IFL_LAUNCH_CHECK();
        IFL_TRY(cub_exclusive_sum(s, s->cbuf0, s->cbuf1, ntiles + 1));
        p2g32_scatter_kernel<<<grid_blocks(a.n, 256), 256, 0, s->stream>>>(a);
        IFL_LAUNCH_CHECK();
        p2g32_tile_kernel<<<ntiles, 256, 0, s->stream>>>(a);
        IFL_LAUNCH_CHECK();
    }
    p2g32_finish_kernel<<<dim3(grid_blocks(most, 256), nq), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_end(s));
    return IFL_OK;
}

// fromParticles for quantity qi; with qi == IFL_Q_D the temperature, which
// lives on the same grid, is transferred in the same pass (the weights of a
// node depend only on the positions).
int k_from_particles(Solver* s, int qi, bool with_t) {
    if (fp32_particles(s)) return k_from_particles32(s, &qi, 1, with_t);
    const QtyView& q = s->q[qi];
    long cells = (long)q.w * q.h;
    int n = s->particle_count;
    int* hist = s->cbuf0;
    int* start = s->cbuf1;
    IFL_HIP(hipMemsetAsync(hist, 0, (cells + 1) * sizeof(int), s->stream));
    if (n > 0) {
        p2g_key_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(q, s->posx, s->posy, n, s->pkey, s->pidx, hist);
        IFL_LAUNCH_CHECK();
        size_t bytes = s->cub_bytes;
        IFL_HIP(hipcub::DeviceRadixSort::SortPairs(s->cub_tmp, bytes, s->pkey, s->pkey2, s->pidx, s->pidx2, n, 0,
                                                   bits_for(cells), s->stream));
    }
    IFL_TRY(cub_exclusive_sum(s, hist, start, cells + 1));
    P2GArgs a;
    a.q = q;
    a.cell2 = with_t ? s->q[IFL_Q_T].cell : nullptr;
    a.src2 = with_t ? s->q[IFL_Q_T].src : nullptr;
    a.posx = s->posx;
    a.posy = s->posy;
    a.prop = s->prop[qi];
    a.prop2 = with_t ? s->prop[IFL_Q_T] : nullptr;
    a.start = start;
    a.order = s->pidx2;
    a.weight = s->weight;
    IFL_TRY(prof_begin(s, IFL_KERNEL_P2G, -1));
    p2g_gather_kernel<<<grid_blocks(cells, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_end(s));
    return IFL_OK;
}

int k_extrapolate_flip(Solver* s, int qi) {
    const QtyView& q = s->q[qi];
    long cells = (long)q.w * q.h;
    long cap = (long)(s->W + 1) * (s->H + 1);
    FlipScalars* sc = dev_sc(s);
    IFL_HIP(hipMemsetAsync(&sc->hard, 0, sizeof(int), s->stream));
    flip_mask_kernel<<<grid_blocks(cells, 256), 256, 0, s->stream>>>(q, s->cflag, sc);
    IFL_LAUNCH_CHECK();
    IFL_TRY(cub_select(s, s->cflag, s->front, &sc->ready, cells));
    flip_levels_kernel<<<1, 1024, 0, s->stream>>>(q, s->front, cap, sc);
    IFL_LAUNCH_CHECK();
    flip_stack_kernel<<<1, 64, 0, s->stream>>>(q, s->front, 2 * cap, sc);
    IFL_LAUNCH_CHECK();
    flip_border_kernel<<<1, 256, 0, s->stream>>>(q);
    IFL_LAUNCH_CHECK();
    empty_to_fluid_kernel<<<grid_blocks(cells, 256), 256, 0, s->stream>>>(q.cell, cells);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// particlesToGrid 8:916-927 up to the prune's kernels
int k_particles_to_grid_begin(Solver* s) {
    if (s->cfg.precision != IFL_PREC_FP64) {
        // the three transfers together (each extrapolation reads and writes
        // only its own grid, and no transfer reads a grid), then the
        // extrapolations in the reference's order
        const int qis[3] = {IFL_Q_D, IFL_Q_U, IFL_Q_V};
        IFL_TRY(k_from_particles32(s, qis, 3, true));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_D));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_T));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_U));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_V));
    } else {
        IFL_TRY(k_from_particles(s, IFL_Q_D, true));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_D));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_T));
        IFL_TRY(k_from_particles(s, IFL_Q_U, false));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_U));
        IFL_TRY(k_from_particles(s, IFL_Q_V, false));
        IFL_TRY(k_extrapolate_flip(s, IFL_Q_V));
    }
    IFL_TRY(k_count_particles(s));
    return fp32_particles(s) ? prune<float>(s) : prune<double>(s);
}

// particlesToGrid 8:916-927
int k_particles_to_grid(Solver* s) {
    IFL_TRY(k_particles_to_grid_begin(s));
    IFL_TRY(k_prune_finish(s));
    return k_seed(s);
}

template <typename P>
static int particle_advect(Solver* s, double dt) {
    int n = s->particle_count;
    if (n == 0) return IFL_OK;
    const ParticleArrays<P> pa = particle_arrays<P>(s);
    AdvectPArgs<P> a;
    a.u = s->q[IFL_Q_U];
    a.v = s->q[IFL_Q_V];
    a.posx = pa.a[0];
    a.posy = pa.a[1];
    a.n = n;
    a.W = s->W;
    a.H = s->H;
    a.hx = s->hx;
    a.dt = dt;
    a.bodies = s->bodies_d;
    a.nb = s->nbodies;
    IFL_TRY(prof_begin(s, IFL_KERNEL_PARTICLE_ADVECT, -1));
    particle_advect_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    IFL_TRY(prof_end(s));
    return IFL_OK;
}

int k_particle_advect(Solver* s, double dt) {
    return fp32_particles(s) ? particle_advect<float>(s, dt) : particle_advect<double>(s, dt);
}

int flip_check_overflow(Solver* s) {
    IFL_TRY(read_scalars(s));
    if (host_sc(s)->overflow) {
        s->err = "extrapolation work list overflowed";
        return IFL_ERR_HIP;
    }
    return IFL_OK;
}

}  // namespace ifl
