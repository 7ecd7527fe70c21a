// Grid-field kernels of FluidSolver::update(): solid fields, extrapolation,
// advection, inflow, right-hand side, densities, matrix assembly, pressure
// update, buoyancy and boundary conditions.
//
// The reference writes most of these as scatter loops (a cell adds to its
// four faces). Here every output element gathers its contributions in the
// order the reference's row-major loop would have added them, so the sums
// are bit-identical (accumulation orders cited per kernel).
#include <algorithm>

#include "ifl_solver.hpp"

namespace ifl {

// ------------------------------------------------------------ fillSolidFields
// phi at the (w+1)(h+1) corners, 7:479-488
IFL_DEV void fill_phi_body(const QtyView& q, const Body* __restrict__ bodies, int nb, long gi) {
    int w1 = q.w + 1, h1 = q.h + 1;
    if (gi >= (long)w1 * h1) return;
    int i = (int)gi;
    int ix = i % w1, iy = i / w1;
    double x = (ix + q.ox - 0.5) * q.hx;
    double y = (iy + q.oy - 0.5) * q.hx;
    double phi = body_distance(bodies[0], x, y);
    for (int k = 1; k < nb; k++) phi = smin(phi, body_distance(bodies[k], x, y));
    q.phi[i] = phi;
}
__global__ void fill_phi_kernel(QtyView q, const Body* __restrict__ bodies, int nb) {
    fill_phi_body(q, bodies, nb, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

// body, volume, normal, cell per cell: 4:453-479 (level 4), 7:490-521 (5+)
IFL_DEV void fill_cells_body(const QtyView& q, const Body* __restrict__ bodies, int nb, int level, long gi) {
    if (gi >= (long)q.w * q.h) return;
    int i = (int)gi;
    int ix = i % q.w, iy = i / q.w;
    double x = (ix + q.ox) * q.hx;
    double y = (iy + q.oy) * q.hx;
    int body = 0;
    double d = body_distance(bodies[0], x, y);
    for (int k = 1; k < nb; k++) {
        double id = body_distance(bodies[k], x, y);
        if (id < d) {
            body = k;
            d = id;
        }
    }
    q.body[i] = (uint8_t)body;
    double nx, ny;
    if (level == 4) {
        q.cell[i] = d < 0.0 ? IFL_CELL_SOLID : IFL_CELL_FLUID;
        body_normal(bodies[body], nx, ny, x, y);
        q.nx[i] = nx;
        q.ny[i] = ny;
        return;
    }
    int ip = ix + iy * (q.w + 1);
    double vol = 1.0 - occupancy(q.phi[ip], q.phi[ip + 1], q.phi[ip + q.w + 1], q.phi[ip + q.w + 2]);
    if (vol < 0.01) vol = 0.0;
    q.volume[i] = vol;
    body_normal(bodies[body], nx, ny, x, y);
    q.nx[i] = nx;
    q.ny[i] = ny;
    q.cell[i] = (vol == 0.0) ? IFL_CELL_SOLID : IFL_CELL_FLUID;
}
__global__ void fill_cells_kernel(QtyView q, const Body* __restrict__ bodies, int nb, int level) {
    fill_cells_body(q, bodies, nb, level, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

// fillSolidFields with host-evaluated bodies (7:475-522): phi, the closest
// body and its normal come from the host callbacks in the reference's order;
// volume and cell follow from them exactly as on the device.
static int fill_solid_host(Solver* s, int qi) {
    const QtyView q = s->q[qi];
    const ifl_host_body_ops& ops = s->hops;
    const int w = q.w, h = q.h, nb = s->nbodies;
    const long n = (long)w * h, nc = (long)(w + 1) * (h + 1);
    std::vector<double> xs, ys, tmp, phi;
    if (s->level >= 5) {
        xs.resize(nc);
        ys.resize(nc);
        for (long i = 0; i < nc; i++) {
            xs[i] = ((int)(i % (w + 1)) + q.ox - 0.5) * q.hx;
            ys[i] = ((int)(i / (w + 1)) + q.oy - 0.5) * q.hx;
        }
        phi.resize(nc);
        tmp.resize(nc);
        ops.distance(ops.user, 0, xs.data(), ys.data(), phi.data(), (size_t)nc);
        for (int k = 1; k < nb; k++) {
            ops.distance(ops.user, k, xs.data(), ys.data(), tmp.data(), (size_t)nc);
            for (long i = 0; i < nc; i++) phi[i] = smin(phi[i], tmp[i]);
        }
    }
    xs.resize(n);
    ys.resize(n);
    for (long i = 0; i < n; i++) {
        xs[i] = ((int)(i % w) + q.ox) * q.hx;
        ys[i] = ((int)(i / w) + q.oy) * q.hx;
    }
    std::vector<double> d(n), nx(n), ny(n), vol;
    std::vector<uint8_t> body(n, 0), cell(n);
    tmp.resize(n);
    ops.distance(ops.user, 0, xs.data(), ys.data(), d.data(), (size_t)n);
    for (int k = 1; k < nb; k++) {
        ops.distance(ops.user, k, xs.data(), ys.data(), tmp.data(), (size_t)n);
        for (long i = 0; i < n; i++)
            if (tmp[i] < d[i]) {
                body[i] = (uint8_t)k;
                d[i] = tmp[i];
            }
    }
    // distanceNormal of each cell's closest body, one batch per body
    std::vector<long> sel;
    std::vector<double> bx, by, bnx, bny;
    for (int k = 0; k < nb; k++) {
        sel.clear();
        for (long i = 0; i < n; i++)
            if (body[i] == k) sel.push_back(i);
        if (sel.empty()) continue;
        bx.resize(sel.size());
        by.resize(sel.size());
        bnx.resize(sel.size());
        bny.resize(sel.size());
        for (size_t j = 0; j < sel.size(); j++) {
            bx[j] = xs[sel[j]];
            by[j] = ys[sel[j]];
        }
        ops.normal(ops.user, k, bx.data(), by.data(), bnx.data(), bny.data(), sel.size());
        for (size_t j = 0; j < sel.size(); j++) {
            nx[sel[j]] = bnx[j];
            ny[sel[j]] = bny[j];
        }
    }
    if (s->level == 4) {
        for (long i = 0; i < n; i++) cell[i] = d[i] < 0.0 ? IFL_CELL_SOLID : IFL_CELL_FLUID;  // 4:466-470
    } else {
        vol.resize(n);
        for (long i = 0; i < n; i++) {
            const long ip = (i % w) + (i / w) * (w + 1);
            double v = 1.0 - occupancy(phi[ip], phi[ip + 1], phi[ip + w + 1], phi[ip + w + 2]);
            if (v < 0.01) v = 0.0;
            vol[i] = v;
            cell[i] = v == 0.0 ? IFL_CELL_SOLID : IFL_CELL_FLUID;
        }
        IFL_HIP(hipMemcpyAsync(q.phi, phi.data(), nc * sizeof(double), hipMemcpyHostToDevice, s->stream));
        IFL_HIP(hipMemcpyAsync(q.volume, vol.data(), n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    }
    IFL_HIP(hipMemcpyAsync(q.body, body.data(), n, hipMemcpyHostToDevice, s->stream));
    IFL_HIP(hipMemcpyAsync(q.cell, cell.data(), n, hipMemcpyHostToDevice, s->stream));
    IFL_HIP(hipMemcpyAsync(q.nx, nx.data(), n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    IFL_HIP(hipMemcpyAsync(q.ny, ny.data(), n * sizeof(double), hipMemcpyHostToDevice, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));  // the host vectors go out of scope
    return IFL_OK;
}

int k_fill_solid(Solver* s, int qi) {
    if (s->nbodies == 0) return IFL_OK;  // 7:476-477
    if (s->host_geom) return fill_solid_host(s, qi);
    QtyView q = s->q[qi];
    if (s->level >= 5) {
        long n = (long)(q.w + 1) * (q.h + 1);
        fill_phi_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(q, s->bodies_d, s->nbodies);
        IFL_LAUNCH_CHECK();
    }
    fill_cells_kernel<<<grid_blocks((long)q.w * q.h, 256), 256, 0, s->stream>>>(q, s->bodies_d, s->nbodies, s->level);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ extrapolate
// fillSolidMask 7:524-542: interior non-fluid cells record which upstream
// neighbours (along the normal) are still unknown.
IFL_DEV void solid_mask_body(const QtyView& q, int* __restrict__ list, int* __restrict__ cnt, int i) {
    int w = q.w, h = q.h;
    if (i >= w * h) return;
    int x = i % w, y = i / w;
    if (x < 1 || y < 1 || x >= w - 1 || y >= h - 1) return;
    if (q.cell[i] == IFL_CELL_FLUID) return;
    double nx = q.nx[i], ny = q.ny[i];
    uint8_t m = 0;
    if (nx != 0.0 && q.cell[i + sgn(nx)] != IFL_CELL_FLUID) m |= 1;
    if (ny != 0.0 && q.cell[i + sgn(ny) * w] != IFL_CELL_FLUID) m |= 2;
    q.mask[i] = m;
    if (m == 0) list[atomicAdd(cnt, 1)] = i;  // initial ready set, 7:563-571
}
__global__ void solid_mask_kernel(QtyView q, int* __restrict__ list, int* __restrict__ cnt) {
    solid_mask_body(q, list, cnt, blockIdx.x * blockDim.x + threadIdx.x);
}

IFL_DEV double src_safe(const QtyView& q, long i) { return (i >= 0 && i < (long)q.w * q.h) ? q.src[i] : 0.0; }

// extrapolateNormal 7:544-552
IFL_DEV double extrap_normal(const QtyView& q, int i) {
    double nx = q.nx[i], ny = q.ny[i];
    double sx = src_safe(q, (long)i + sgn(nx));
    double sy = src_safe(q, (long)i + (long)sgn(ny) * q.w);
    return (fabs(nx) * sx + fabs(ny) * sy) / (fabs(nx) + fabs(ny));
}

// The reference drains a LIFO work list (7:573-587). Each interior cell
// becomes ready exactly once, when its last upstream neighbour is final, and
// its value reads only upstream values, so processing ready cells level by
// level gives the same values in any order. One workgroup walks the levels.
IFL_DEV void extrapolate_levels_body(const QtyView& q, int* __restrict__ lists, int* __restrict__ cnt, long cap,
                                     int& s_next) {
    int w = q.w;
    long n = (long)q.w * q.h;
    int* cur = lists;
    int* nxt = lists + cap;
    int ncur = cnt[0];
    unsigned* mask32 = (unsigned*)q.mask;
    // Every legitimate level retires at least one cell and the dependency
    // depth along normals is bounded by the grid size; the cap only guards a
    // cyclic configuration on the domain border (see the oracle's header).
    long max_levels = 4L * (q.w + q.h) + 16;
    for (long lvl = 0; ncur > 0 && lvl < max_levels; lvl++) {
        for (int k = threadIdx.x; k < ncur; k += blockDim.x) {
            int c = cur[k];
            q.src[c] = extrap_normal(q, c);
        }
        if (threadIdx.x == 0) s_next = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < ncur; k += blockDim.x) {
            int c = cur[k];
            // freeNeighbour on the four neighbours whose normal points at c, 7:579-586
            long nb[4] = {(long)c - 1, (long)c + 1, (long)c - w, (long)c + w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                long j = nb[e];
                if (j < 0 || j >= n) continue;
                bool cond = e == 0 ? q.nx[j] > 0.0 : e == 1 ? q.nx[j] < 0.0 : e == 2 ? q.ny[j] > 0.0 : q.ny[j] < 0.0;
                if (!cond) continue;
                unsigned bit = (e < 2) ? 1u : 2u;
                int sh = (int)(j & 3) * 8;
                unsigned old = atomicAnd(&mask32[j >> 2], ~(bit << sh));
                unsigned nb8 = ((old >> sh) & 0xFFu) & ~bit;
                if (nb8 == 0 && q.cell[j] != IFL_CELL_FLUID) {
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
__global__ __launch_bounds__(1024) void extrapolate_levels_kernel(QtyView q, int* __restrict__ lists, int* __restrict__ cnt, long cap) {
    __shared__ int s_next;
    extrapolate_levels_body(q, lists, cnt, cap, s_next);
}

IFL_DEV void border_solid_body(const QtyView& q, int* flag, int i) {
    int w = q.w, h = q.h;
    int perim = 2 * w + 2 * h;
    if (i >= perim) return;
    int x, y;
    if (i < w) { x = i; y = 0; }
    else if (i < 2 * w) { x = i - w; y = h - 1; }
    else if (i < 2 * w + h) { x = 0; y = i - 2 * w; }
    else { x = w - 1; y = i - 2 * w - h; }
    if (q.cell[x + y * w] != IFL_CELL_FLUID) atomicOr(flag, 1);
}
__global__ void border_solid_kernel(QtyView q, int* flag) {
    border_solid_body(q, flag, blockIdx.x * blockDim.x + threadIdx.x);
}

int k_extrapolate(Solver* s, int qi) {
    if (s->level < 4) return IFL_OK;
    if (s->level == 8) return IFL_ERR_UNSUPPORTED;  // handled by the particle path
    QtyView q = s->q[qi];
    long n = (long)q.w * q.h;
    IFL_HIP(hipMemsetAsync(s->front_cnt, 0, 2 * sizeof(int), s->stream));
    solid_mask_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(q, s->front, s->front_cnt);
    IFL_LAUNCH_CHECK();
    border_solid_kernel<<<grid_blocks(2L * q.w + 2L * q.h, 256), 256, 0, s->stream>>>(q, s->front_cnt + 1);
    IFL_LAUNCH_CHECK();
    long cap = (long)(s->W + 1) * (s->H + 1);
    extrapolate_levels_kernel<<<1, 1024, 0, s->stream>>>(q, s->front, s->front_cnt, cap);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ batched extrapolation
// The batched form's extrapolations of d, u and v: for each quantity, one
// launch per phase (counters, fillSolidMask, the border flag, the level walk
// with one workgroup per solver) over a table of every solver's quantity,
// the single-solver kernels' bodies.
struct ExtrapInst {
    QtyView q;
    int* front;
    int* front_cnt;
    long cap;
};

__global__ void batch_extrap_zero_kernel(const ExtrapInst* t, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i].front_cnt[0] = 0;
    t[i].front_cnt[1] = 0;
}
__global__ __launch_bounds__(256) void batch_solid_mask_kernel(const ExtrapInst* t, const int2* map) {
    const int2 m = map[blockIdx.x];
    solid_mask_body(t[m.x].q, t[m.x].front, t[m.x].front_cnt, m.y * 256 + threadIdx.x);
}
__global__ __launch_bounds__(256) void batch_border_solid_kernel(const ExtrapInst* t, const int2* map) {
    const int2 m = map[blockIdx.x];
    border_solid_body(t[m.x].q, t[m.x].front_cnt + 1, m.y * 256 + threadIdx.x);
}
__global__ __launch_bounds__(1024) void batch_extrapolate_levels_kernel(const ExtrapInst* t) {
    __shared__ int s_next;
    const ExtrapInst& e = t[blockIdx.x];
    extrapolate_levels_body(e.q, e.front, e.front_cnt, e.cap, s_next);
}

struct ExtrapDev {
    std::vector<Solver*> solvers;
    // one table per quantity (indexed by IFL_Q_*), built on first use
    ExtrapInst* inst_h[4] = {nullptr, nullptr, nullptr, nullptr};  // pinned
    ExtrapInst* inst_d[4] = {nullptr, nullptr, nullptr, nullptr};
    int2* mask_map[4] = {nullptr, nullptr, nullptr, nullptr};
    int2* border_map[4] = {nullptr, nullptr, nullptr, nullptr};
    int nmask[4] = {0, 0, 0, 0}, nborder[4] = {0, 0, 0, 0};
    std::vector<hipEvent_t> ready;
    hipEvent_t joined = nullptr;
};

void extrap_dev_free(ExtrapDev* d) {
    if (!d) return;
    for (int k = 0; k < 4; k++) {
        if (d->inst_h[k]) (void)hipHostFree(d->inst_h[k]);
        if (d->inst_d[k]) (void)hipFree(d->inst_d[k]);
        if (d->mask_map[k]) (void)hipFree(d->mask_map[k]);
        if (d->border_map[k]) (void)hipFree(d->border_map[k]);
    }
    for (hipEvent_t e : d->ready) (void)hipEventDestroy(e);
    if (d->joined) (void)hipEventDestroy(d->joined);
    delete d;
}

int k_extrapolate_batch(Solver* const* ss, int n, hipStream_t stream, ExtrapDev** dev, int qmask, bool join) {
    Solver* s = ss[0];  // error reporting
    ExtrapDev* d = *dev;
    if (!d || d->solvers.size() != (size_t)n || !std::equal(ss, ss + n, d->solvers.begin())) {
        extrap_dev_free(d);
        d = *dev = new ExtrapDev();
        d->solvers.assign(ss, ss + n);
        d->ready.assign(n, nullptr);
        for (int i = 0; i < n; i++) IFL_HIP(hipEventCreateWithFlags(&d->ready[i], hipEventDisableTiming));
        IFL_HIP(hipEventCreateWithFlags(&d->joined, hipEventDisableTiming));
    }
    for (int k = 0; k < 4; k++) {
        if (!(qmask & (1 << k)) || d->inst_d[k]) continue;
        std::vector<int2> mm, bm;
        for (int i = 0; i < n; i++) {
            const QtyView& q = ss[i]->q[k];
            const long cells = (long)q.w * q.h, perim = 2L * q.w + 2L * q.h;
            for (long b = 0; b < (cells + 255) / 256; b++) mm.push_back(make_int2(i, (int)b));
            for (long b = 0; b < (perim + 255) / 256; b++) bm.push_back(make_int2(i, (int)b));
        }
        d->nmask[k] = (int)mm.size();
        d->nborder[k] = (int)bm.size();
        IFL_HIP(hipMalloc((void**)&d->mask_map[k], mm.size() * sizeof(int2)));
        IFL_HIP(hipMemcpy(d->mask_map[k], mm.data(), mm.size() * sizeof(int2), hipMemcpyHostToDevice));
        IFL_HIP(hipMalloc((void**)&d->border_map[k], bm.size() * sizeof(int2)));
        IFL_HIP(hipMemcpy(d->border_map[k], bm.data(), bm.size() * sizeof(int2), hipMemcpyHostToDevice));
        IFL_HIP(hipHostMalloc((void**)&d->inst_h[k], n * sizeof(ExtrapInst), hipHostMallocDefault));
        IFL_HIP(hipMalloc((void**)&d->inst_d[k], n * sizeof(ExtrapInst)));
    }
    // the solvers' src arrays swap every step, so the tables are written
    // afresh (the previous step's copies completed before it returned)
    for (int k = 0; k < 4; k++) {
        if (!(qmask & (1 << k))) continue;
        for (int i = 0; i < n; i++) {
            ExtrapInst& e = d->inst_h[k][i];
            e.q = ss[i]->q[k];
            e.front = ss[i]->front;
            e.front_cnt = ss[i]->front_cnt;
            e.cap = (long)(ss[i]->W + 1) * (ss[i]->H + 1);
        }
    }
    if (join)
        for (int i = 0; i < n; i++) {
            IFL_HIP(hipEventRecord(d->ready[i], ss[i]->stream));
            IFL_HIP(hipStreamWaitEvent(stream, d->ready[i], 0));
        }
    for (int k = 0; k < 4; k++) {
        if (!(qmask & (1 << k))) continue;
        IFL_HIP(hipMemcpyAsync(d->inst_d[k], d->inst_h[k], n * sizeof(ExtrapInst), hipMemcpyHostToDevice, stream));
        batch_extrap_zero_kernel<<<grid_blocks(n, 256), 256, 0, stream>>>(d->inst_d[k], n);
        IFL_LAUNCH_CHECK();
        batch_solid_mask_kernel<<<d->nmask[k], 256, 0, stream>>>(d->inst_d[k], d->mask_map[k]);
        IFL_LAUNCH_CHECK();
        batch_border_solid_kernel<<<d->nborder[k], 256, 0, stream>>>(d->inst_d[k], d->border_map[k]);
        IFL_LAUNCH_CHECK();
        batch_extrapolate_levels_kernel<<<n, 1024, 0, stream>>>(d->inst_d[k]);
        IFL_LAUNCH_CHECK();
    }
    if (join) {
        IFL_HIP(hipEventRecord(d->joined, stream));
        for (int i = 0; i < n; i++) IFL_HIP(hipStreamWaitEvent(ss[i]->stream, d->joined, 0));
    }
    return IFL_OK;
}

// ------------------------------------------------------------ non-finite check
// Raises *flag when any of the n values is NaN or infinite.
__global__ __launch_bounds__(256) void nonfinite_kernel(const double* __restrict__ a, long n, int* flag) {
    bool bad = false;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        bad |= !isfinite(a[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

int k_nonfinite(Solver* s, int* dev_flag) {
    IFL_HIP(hipMemsetAsync(dev_flag, 0, sizeof(int), s->stream));
    for (int qi = 0; qi < 4; qi++) {
        if (qi == IFL_Q_T && s->level < 6) continue;
        const QtyView& q = s->q[qi];
        long n = (long)q.w * q.h;
        int blocks = (int)std::min<long>(grid_blocks(n, 256), 4L * s->num_cus);
        nonfinite_kernel<<<blocks, 256, 0, s->stream>>>(q.src, n, dev_flag);
        IFL_LAUNCH_CHECK();
    }
    return IFL_OK;
}

// ------------------------------------------------------------ maxTimestep
// Solver 1's CFL bound (1:310-328): the largest speed at the cell centres.
// A speed is sqrt of a sum of squares, so it is +0 or positive (or NaN, which
// std::max never keeps); the maximum of such values does not depend on the
// order, and their bit patterns order like the values, so one atomicMax on
// the bits combines the waves exactly.
__global__ __launch_bounds__(256) void max_speed_kernel(QtyView u, QtyView v, int W, int H,
                                                         unsigned long long* __restrict__ out) {
    long n = (long)W * H;
    double m = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        int x = (int)(c % W), y = (int)(c / W);
        double uu = qlerp(u.src, u.w, u.h, u.ox, u.oy, x + 0.5, y + 0.5);
        double vv = qlerp(v.src, v.w, v.h, v.ox, v.oy, x + 0.5, y + 0.5);
        double velocity = sqrt(uu * uu + vv * vv);
        m = smax(m, velocity);
    }
    for (int off = 32; off > 0; off >>= 1) m = smax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

int k_max_timestep(Solver* s, double* dt_out) {
    unsigned long long* d = (unsigned long long*)s->front_cnt;  // 8 bytes of scratch
    IFL_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s->stream));
    long n = (long)s->W * s->H;
    int blocks = (int)std::min<long>(grid_blocks(n, 256), 4L * s->num_cus);
    max_speed_kernel<<<blocks, 256, 0, s->stream>>>(s->q[IFL_Q_U], s->q[IFL_Q_V], s->W, s->H, d);
    IFL_LAUNCH_CHECK();
    unsigned long long bits = 0;
    IFL_HIP(hipMemcpyAsync(&bits, d, sizeof bits, hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    double max_velocity;
    memcpy(&max_velocity, &bits, sizeof max_velocity);
    double max_timestep = 2.0 * s->hx / max_velocity;  // 1:324
    *dt_out = (1.0 < max_timestep) ? 1.0 : max_timestep;  // std::min(maxTimestep, 1.0), 1:327
    return IFL_OK;
}

// ------------------------------------------------------------ toImage
// 3:455-465 / 4:903-916 / 5:961-972 / 6:1021-1055 / 7:1047-1081 /
// 8:1426-1467: one RGBA pixel (two with the heat panel) per cell. (int) of a
// double follows x86 (NaN and out-of-range give INT_MIN); a byte store keeps
// the low 8 bits of the int.
IFL_DEV int x86_int(double v) {
    return (v == v && v > -2147483649.0 && v < 2147483648.0) ? (int)v : (int)0x80000000u;
}

struct ImageArgs {
    const double* d;
    const double* volume;
    const uint8_t* cell;
    const double* t;
    double t_amb;
    int W, H, level, heat;
    uint8_t* rgba;  // the image rows from y0 on
    int y0, rows;   // the rows this launch renders
};

__global__ void to_image_kernel(ImageArgs a) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)a.W * a.rows) return;
    const long i = (long)a.y0 * a.W + k;
    const int x = (int)(i % a.W), y = (int)(i / a.W), lv = a.level;
    // a.rgba holds the image from row y0 on
    uint8_t* rgba = a.rgba;
    const int yl = y - a.y0;
    if (lv <= 5) {
        int shade = lv == 5 ? x86_int((1.0 - a.d[i]) * a.volume[i] * 255.0) : x86_int((1.0 - a.d[i]) * 255.0);
        shade = imax(imin(shade, 255), 0);
        if (lv == 4 && a.cell[i] == IFL_CELL_SOLID) shade = 0;  // 4:908-909
        rgba[k * 4 + 0] = rgba[k * 4 + 1] = rgba[k * 4 + 2] = (uint8_t)shade;
        rgba[k * 4 + 3] = 0xFF;
        return;
    }
    long idxl = 0, idxr;
    if (a.heat) {
        idxl = 4 * ((long)x + (long)yl * a.W * 2);
        idxr = 4 * ((long)x + (long)yl * a.W * 2 + a.W);
    } else {
        idxr = 4 * k;
    }
    const double volume = a.volume[i];
    double shade = (1.0 - a.d[i]) * volume;
    shade = smin(smax(shade, 0.0), 1.0);
    const uint8_t g = (uint8_t)x86_int(shade * 255.0);
    rgba[idxr + 0] = rgba[idxr + 1] = rgba[idxr + 2] = g;
    rgba[idxr + 3] = 0xFF;
    if (lv == 8 && a.cell[i] == IFL_CELL_EMPTY) {  // 8:1445-1448
        rgba[idxr] = 0xFF;
        rgba[idxr + 1] = rgba[idxr + 2] = 0;
    }
    if (a.heat) {
        double t = lv == 8 ? fabs(a.t[i] - a.t_amb) / 70.0 : (a.t[i] - a.t_amb) / 700.0;  // 8:1451, 7:1067
        t = smin(smax(t, 0.0), 1.0);
        const double r = 1.0 + volume * (smin(t * 4.0, 1.0) - 1.0);
        const double gg = 1.0 + volume * (smin(t * 2.0, 1.0) - 1.0);
        const double b = 1.0 + volume * (smax(smin(t * 4.0 - 3.0, 1.0), 0.0) - 1.0);
        rgba[idxl + 0] = (uint8_t)x86_int(r * 255.0);
        rgba[idxl + 1] = (uint8_t)x86_int(gg * 255.0);
        rgba[idxl + 2] = (uint8_t)x86_int(b * 255.0);
        rgba[idxl + 3] = 0xFF;
    }
}

int k_to_image(Solver* s, uint8_t* dev_rgba, int heat, int y0, int rows) {
    ImageArgs a;
    a.d = s->q[IFL_Q_D].src;
    a.volume = s->q[IFL_Q_D].volume;
    a.cell = s->q[IFL_Q_D].cell;
    a.t = s->level >= 6 ? s->q[IFL_Q_T].src : nullptr;
    a.t_amb = s->cfg.t_amb;
    a.W = s->W;
    a.H = s->H;
    a.level = s->level;
    a.heat = heat && s->level >= 6;
    a.rgba = dev_rgba;
    a.y0 = y0;
    a.rows = rows;
    long n = (long)s->W * rows;
    to_image_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ advect
struct AdvectArgs {
    QtyView q[4];
    int nq;
    const double* u;
    const double* v;
    int uw, uh, vw, vh;
    double hx, dt;
    int level;
    const Body* bodies;
    // host-evaluated bodies: back-projections are queued here instead
    HostReq* req;   // hreq_cap requests per quantity
    int* req_cnt;   // per quantity
    long req_cap;
};

// rungeKutta3 7:297-315 + backProject 7:424-435 + cerp 7:404-422, as called
// from advect 7:437-454 (level 3: 3:154-165, all cells, no back-projection)
IFL_DEV void advect_body(const AdvectArgs& a, int qy, long gi) {
    const QtyView& q = a.q[qy];
    if (gi >= (long)q.w * q.h) return;
    int i = (int)gi;
    if (a.level >= 4 && q.cell[i] != IFL_CELL_FLUID) return;
    int ix = i % q.w, iy = i / q.w;
    double x = ix + q.ox;
    double y = iy + q.oy;
    const double hx = a.hx, dt = a.dt;
    double first_u = qlerp(a.u, a.uw, a.uh, 0.0, 0.5, x, y) / hx;
    double first_v = qlerp(a.v, a.vw, a.vh, 0.5, 0.0, x, y) / hx;
    double mid_x = x - 0.5 * dt * first_u;
    double mid_y = y - 0.5 * dt * first_v;
    double mid_u = qlerp(a.u, a.uw, a.uh, 0.0, 0.5, mid_x, mid_y) / hx;
    double mid_v = qlerp(a.v, a.vw, a.vh, 0.5, 0.0, mid_x, mid_y) / hx;
    double last_x = x - 0.75 * dt * mid_u;
    double last_y = y - 0.75 * dt * mid_v;
    double last_u = qlerp(a.u, a.uw, a.uh, 0.0, 0.5, last_x, last_y);
    double last_v = qlerp(a.v, a.vw, a.vh, 0.5, 0.0, last_x, last_y);
    x -= dt * ((2.0 / 9.0) * first_u + (3.0 / 9.0) * mid_u + (4.0 / 9.0) * last_u);
    y -= dt * ((2.0 / 9.0) * first_v + (3.0 / 9.0) * mid_v + (4.0 / 9.0) * last_v);
    if (a.level >= 4) {
        int rx = imin(imax(trunc_idx(x - q.ox), 0), q.w - 1);
        int ry = imin(imax(trunc_idx(y - q.oy), 0), q.h - 1);
        int ri = rx + ry * q.w;
        if (q.cell[ri] != IFL_CELL_FLUID) {
            x = (x - q.ox) * hx;
            y = (y - q.oy) * hx;
            if (a.req) {
                // the host evaluates closestSurfacePoint; advect_finish_kernel samples
                const int k = atomicAdd(&a.req_cnt[qy], 1);
                HostReq* r = a.req + qy * a.req_cap + k;
                r->idx = i;
                r->body = q.body[ri];
                r->x = x;
                r->y = y;
                return;
            }
            body_closest(a.bodies[q.body[ri]], x, y);
            x = x / hx + q.ox;
            y = y / hx + q.oy;
        }
    }
    q.dst[i] = qcerp(q.src, q.w, q.h, q.ox, q.oy, x, y);
}
__global__ void advect_kernel(AdvectArgs a) { advect_body(a, blockIdx.y, blockIdx.x * (long)blockDim.x + threadIdx.x); }

// The mixed form's advection: the same trace (rungeKutta3 with its undivided
// last stage, backProject, cerp) with every velocity sample and the final
// cerp interpolated in single precision. The trace positions stay double: a
// single-precision position 2048 cells from the origin has an ulp of 1.2e-4
// cells, and an end point moved across a cell edge flips the backProject
// test next to a body (a 1e-3 relative error in u at 512x512).
__global__ void advect32_kernel(AdvectArgs a) {
    const QtyView& q = a.q[blockIdx.y];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.w * q.h) return;
    if (a.level >= 4 && q.cell[i] != IFL_CELL_FLUID) return;
    int ix = i % q.w, iy = i / q.w;
    const double hx = a.hx, dt = a.dt;
    double x = ix + q.ox;
    double y = iy + q.oy;
    const float hxf = (float)hx;
    float first_u = qlerpf(a.u, a.uw, a.uh, 0.0, 0.5, x, y) / hxf;
    float first_v = qlerpf(a.v, a.vw, a.vh, 0.5, 0.0, x, y) / hxf;
    double mid_x = x - 0.5 * dt * first_u;
    double mid_y = y - 0.5 * dt * first_v;
    float mid_u = qlerpf(a.u, a.uw, a.uh, 0.0, 0.5, mid_x, mid_y) / hxf;
    float mid_v = qlerpf(a.v, a.vw, a.vh, 0.5, 0.0, mid_x, mid_y) / hxf;
    double last_x = x - 0.75 * dt * mid_u;
    double last_y = y - 0.75 * dt * mid_v;
    float last_u = qlerpf(a.u, a.uw, a.uh, 0.0, 0.5, last_x, last_y);
    float last_v = qlerpf(a.v, a.vw, a.vh, 0.5, 0.0, last_x, last_y);
    x -= dt * ((2.0 / 9.0) * first_u + (3.0 / 9.0) * mid_u + (4.0 / 9.0) * last_u);
    y -= dt * ((2.0 / 9.0) * first_v + (3.0 / 9.0) * mid_v + (4.0 / 9.0) * last_v);
    if (a.level >= 4) {
        int rx = imin(imax(trunc_idx(x - q.ox), 0), q.w - 1);
        int ry = imin(imax(trunc_idx(y - q.oy), 0), q.h - 1);
        int ri = rx + ry * q.w;
        if (q.cell[ri] != IFL_CELL_FLUID) {
            double xd = (x - q.ox) * hx;
            double yd = (y - q.oy) * hx;
            body_closest(a.bodies[q.body[ri]], xd, yd);
            x = xd / hx + q.ox;
            y = yd / hx + q.oy;
        }
    }
    q.dst[i] = (double)qcerpf(q.src, q.w, q.h, q.ox, q.oy, x, y);
}

// the cerp of advect (7:450) at the back-projected points the host returned
__global__ void advect_finish_kernel(AdvectArgs a) {
    const QtyView& q = a.q[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.req_cnt[blockIdx.y]) return;
    const HostReq r = a.req[blockIdx.y * a.req_cap + k];
    const double x = r.x / a.hx + q.ox;
    const double y = r.y / a.hx + q.oy;
    q.dst[r.idx] = qcerp(q.src, q.w, q.h, q.ox, q.oy, x, y);
}

// backProject with host-evaluated bodies: the trace end points that land in a
// solid cell go to the host's closestSurfacePoint, one batch per body
static int advect_host_finish(Solver* s, AdvectArgs& a, long maxn) {
    int cnt[4] = {0, 0, 0, 0};
    IFL_HIP(hipMemcpyAsync(cnt, a.req_cnt, a.nq * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    IFL_HIP(hipStreamSynchronize(s->stream));
    int most = 0;
    std::vector<HostReq> req;
    std::vector<double> xs, ys;
    std::vector<long> sel;
    for (int qi = 0; qi < a.nq; qi++) {
        if (cnt[qi] == 0) continue;
        most = std::max(most, cnt[qi]);
        req.resize(cnt[qi]);
        HostReq* dev = a.req + qi * a.req_cap;
        IFL_HIP(hipMemcpy(req.data(), dev, cnt[qi] * sizeof(HostReq), hipMemcpyDeviceToHost));
        for (int b = 0; b < s->nbodies; b++) {
            sel.clear();
            for (int k = 0; k < cnt[qi]; k++)
                if (req[k].body == b) sel.push_back(k);
            if (sel.empty()) continue;
            xs.resize(sel.size());
            ys.resize(sel.size());
            for (size_t j = 0; j < sel.size(); j++) {
                xs[j] = req[sel[j]].x;
                ys[j] = req[sel[j]].y;
            }
            s->hops.closest(s->hops.user, b, xs.data(), ys.data(), sel.size());
            for (size_t j = 0; j < sel.size(); j++) {
                req[sel[j]].x = xs[j];
                req[sel[j]].y = ys[j];
            }
        }
        IFL_HIP(hipMemcpy(dev, req.data(), cnt[qi] * sizeof(HostReq), hipMemcpyHostToDevice));
    }
    (void)maxn;
    if (most > 0) {
        dim3 grid(grid_blocks(most, 256), a.nq);
        advect_finish_kernel<<<grid, 256, 0, s->stream>>>(a);
        IFL_LAUNCH_CHECK();
    }
    return IFL_OK;
}

int k_advect(Solver* s, int qmask, double dt) {
    AdvectArgs a;
    a.nq = 0;
    long maxn = 0;
    for (int qi = 0; qi < 4; qi++)
        if (qmask & (1 << qi)) {
            a.q[a.nq++] = s->q[qi];
            maxn = std::max(maxn, (long)s->q[qi].w * s->q[qi].h);
        }
    if (a.nq == 0) return IFL_OK;
    const QtyView& u = s->q[IFL_Q_U];
    const QtyView& v = s->q[IFL_Q_V];
    a.u = u.src;
    a.v = v.src;
    a.uw = u.w; a.uh = u.h; a.vw = v.w; a.vh = v.h;
    a.hx = s->hx;
    a.dt = dt;
    a.level = s->level;
    a.bodies = s->bodies_d;
    a.req = nullptr;
    a.req_cnt = nullptr;
    a.req_cap = 0;
    const bool host = s->host_geom && s->level >= 4 && s->nbodies > 0;
    if (host) {
        // every FLUID cell of a quantity may need a back-projection
        if ((long)s->hreq_cap < maxn) {
            if (s->hreq) IFL_HIP(hipFree(s->hreq));
            s->hreq = nullptr;
            s->hreq_cap = 0;
            IFL_HIP(hipMalloc((void**)&s->hreq, 4 * maxn * sizeof(HostReq)));
            s->hreq_cap = maxn;
        }
        a.req = s->hreq;
        a.req_cnt = s->hreq_cnt;
        a.req_cap = (long)s->hreq_cap;
        IFL_HIP(hipMemsetAsync(s->hreq_cnt, 0, 4 * sizeof(int), s->stream));
    }
    dim3 grid(grid_blocks(maxn, 256), a.nq);
    IFL_TRY(prof_begin(s, IFL_KERNEL_ADVECT, -1));
    if (s->cfg.precision != IFL_PREC_FP64 && !host)
        advect32_kernel<<<grid, 256, 0, s->stream>>>(a);
    else
        advect_kernel<<<grid, 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    if (host) IFL_TRY(advect_host_finish(s, a, maxn));
    return prof_end(s);
}

// ------------------------------------------------------------ addInflow 7:456-473
struct InflowArgs {
    QtyView q;
    int xa, xb, ya, yb;  // the cell range
    double x0, y0, x1, y1, v;
    int sequential;
};

IFL_DEV void inflow_cell(const InflowArgs& a, int x, int y) {
    const QtyView& q = a.q;
    const double x0 = a.x0, y0 = a.y0, x1 = a.x1, y1 = a.y1;
    long n = (long)q.w * q.h;
    double l = length2((2.0 * (x + 0.5) * q.hx - (x0 + x1)) / (x1 - x0), (2.0 * (y + 0.5) * q.hx - (y0 + y1)) / (y1 - y0));
    double vi = cubic_pulse(l) * a.v;
    long idx = x + (long)y * q.w;
    if (idx < 0 || idx >= n) return;  // the reference writes past its array here
    if (fabs(q.src[idx]) < fabs(vi)) q.src[idx] = vi;
}

IFL_DEV void inflow_body(const InflowArgs& a, long gi) {
    if (a.sequential) {  // x range wider than the row (Q3): rows alias, keep loop order
        if (gi == 0)
            for (int y = a.ya; y < a.yb; y++)
                for (int x = a.xa; x < a.xb; x++) inflow_cell(a, x, y);
        return;
    }
    const int nx = a.xb - a.xa, ny = a.yb - a.ya;
    if (gi >= (long)nx * ny) return;
    const int i = (int)gi;
    inflow_cell(a, a.xa + i % nx, a.ya + i / nx);
}

__global__ void inflow_kernel(InflowArgs a) { inflow_body(a, blockIdx.x * (long)blockDim.x + threadIdx.x); }

// the launch of quantity qi's inflow (7:456-473); false when its cell range
// is empty
static bool inflow_args(const Solver* s, int qi, double x0, double y0, double x1, double y1, double v,
                        InflowArgs& a) {
    const QtyView& q = s->q[qi];
    int ix0 = (int)(x0 / q.hx - q.ox);
    int iy0 = (int)(y0 / q.hx - q.oy);
    int ix1 = (int)(x1 / q.hx - q.ox);
    int iy1 = (int)(y1 / q.hx - q.oy);
    int ya = std::max(iy0, 0), yb = std::min(iy1, q.h);
    int xa = std::max(ix0, 0), xb = std::min(ix1, q.h);  // x bounded by _h, Q3 (7:463)
    if (yb <= ya || xb <= xa) return false;
    a.q = q;
    a.xa = xa;
    a.xb = xb;
    a.ya = ya;
    a.yb = yb;
    a.x0 = x0;
    a.y0 = y0;
    a.x1 = x1;
    a.y1 = y1;
    a.v = v;
    a.sequential = xb > q.w ? 1 : 0;
    return true;
}

static long inflow_threads(const InflowArgs& a) {
    return a.sequential ? 1 : (long)(a.xb - a.xa) * (a.yb - a.ya);
}

int k_add_inflow(Solver* s, double x0, double y0, double x1, double y1, const double vals[4]) {
    for (int qi = 0; qi < 4; qi++) {
        if (qi == IFL_Q_T && s->level < 6) continue;
        InflowArgs a;
        if (!inflow_args(s, qi, x0, y0, x1, y1, vals[qi], a)) continue;
        inflow_kernel<<<grid_blocks(inflow_threads(a), 256), 256, 0, s->stream>>>(a);
        IFL_LAUNCH_CHECK();
    }
    return IFL_OK;
}

// ------------------------------------------------------------ buildRhs
// 3:208-217 (level 3), 4:614-627 (level 4), 7:624-653 (levels 5-8)
struct RhsArgs {
    const double *u, *v, *uvol, *vvol, *dvol;
    const uint8_t *cell, *body;
    double* r;
    Skew sk;
    int level, nbodies;
    double hx;
    const Body* bodies;
};

IFL_DEV void build_rhs_body(const RhsArgs& a, long gi) {
    int W = a.sk.W, H = a.sk.H;
    if (gi >= (long)W * H) return;
    int i = (int)gi;
    int x = i % W, y = i / W;
    int uw = W + 1;
    double scale = 1.0 / a.hx;
    double res;
    if (a.level == 3 || (a.level == 4 && a.cell[i] == IFL_CELL_FLUID)) {
        res = -scale * (a.u[x + 1 + y * uw] - a.u[x + y * uw] + a.v[x + (y + 1) * W] - a.v[x + y * W]);
    } else if (a.cell[i] != IFL_CELL_FLUID) {
        res = 0.0;
    } else {
        double uv0 = a.uvol[x + y * uw], uv1 = a.uvol[x + 1 + y * uw];
        double vv0 = a.vvol[x + y * W], vv1 = a.vvol[x + (y + 1) * W];
        res = -scale * (uv1 * a.u[x + 1 + y * uw] - uv0 * a.u[x + y * uw] + vv1 * a.v[x + (y + 1) * W] -
                        vv0 * a.v[x + y * W]);
        if (a.nbodies > 0) {
            double vol = a.dvol[i];
            double hx = a.hx;
            if (x > 0) res -= (uv0 - vol) * body_velx(a.bodies[a.body[i - 1]], (y + 0.5) * hx);
            if (y > 0) res -= (vv0 - vol) * body_vely(a.bodies[a.body[i - W]], (x + 0.5) * hx);
            if (x < W - 1) res += (uv1 - vol) * body_velx(a.bodies[a.body[i + 1]], (y + 0.5) * hx);
            if (y < H - 1) res += (vv1 - vol) * body_vely(a.bodies[a.body[i + W]], (x + 0.5) * hx);
        }
    }
    a.r[a.sk.pos(x, y)] = res;
}
__global__ void build_rhs_kernel(RhsArgs a) { build_rhs_body(a, blockIdx.x * (long)blockDim.x + threadIdx.x); }

int k_build_rhs(Solver* s) {
    RhsArgs a;
    a.u = s->q[IFL_Q_U].src;
    a.v = s->q[IFL_Q_V].src;
    a.uvol = s->q[IFL_Q_U].volume;
    a.vvol = s->q[IFL_Q_V].volume;
    a.dvol = s->q[IFL_Q_D].volume;
    a.cell = s->q[IFL_Q_D].cell;
    a.body = s->q[IFL_Q_D].body;
    a.r = s->r;
    a.sk = s->sk;
    a.level = s->level;
    a.nbodies = s->nbodies;
    a.hx = s->hx;
    a.bodies = s->bodies_d;
    build_rhs_kernel<<<grid_blocks((long)s->W * s->H, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ computeDensities 7:658-675
// Face (x, y) of the u grid receives 0.5*rho from cell (x-1, y) and then from
// cell (x, y), in the reference's loop order; v faces from (x, y-1) then (x, y).
IFL_DEV double cell_density(const double* t, const double* d, int i, double rho_air, double t_amb, double alpha) {
    double density = rho_air * t_amb / t[i] * (1.0 + alpha * d[i]);
    return smax(density, 0.05 * rho_air);
}

IFL_DEV void densities_body(const double* __restrict__ t, const double* __restrict__ d, double* __restrict__ ud,
                           double* __restrict__ vd, int W, int H, double rho_air, double t_amb, double alpha, long i) {
    long nu = (long)(W + 1) * H;
    long nv = (long)W * (H + 1);
    if (i < nu) {
        int x = (int)(i % (W + 1)), y = (int)(i / (W + 1));
        double acc = 0.0;
        if (x >= 1) acc += 0.5 * cell_density(t, d, x - 1 + y * W, rho_air, t_amb, alpha);
        if (x < W) acc += 0.5 * cell_density(t, d, x + y * W, rho_air, t_amb, alpha);
        ud[i] = acc;
    } else if (i < nu + nv) {
        long k = i - nu;
        int x = (int)(k % W), y = (int)(k / W);
        double acc = 0.0;
        if (y >= 1) acc += 0.5 * cell_density(t, d, x + (y - 1) * W, rho_air, t_amb, alpha);
        if (y < H) acc += 0.5 * cell_density(t, d, x + y * W, rho_air, t_amb, alpha);
        vd[k] = acc;
    }
}
__global__ void densities_kernel(const double* __restrict__ t, const double* __restrict__ d, double* __restrict__ ud,
                                 double* __restrict__ vd, int W, int H, double rho_air, double t_amb, double alpha) {
    densities_body(t, d, ud, vd, W, H, rho_air, t_amb, alpha, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

int k_compute_densities(Solver* s) {
    double rho = s->cfg.density;
    double alpha = (s->cfg.density_soot - rho) / rho;
    long n = (long)(s->W + 1) * s->H + (long)s->W * (s->H + 1);
    densities_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->q[IFL_Q_T].src, s->q[IFL_Q_D].src, s->udens,
                                                                  s->vdens, s->W, s->H, rho, s->cfg.t_amb, alpha);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ matrix assembly
// buildPressureMatrix 3:222-244 / 4:629-654 / 5:686-713 / 7:680-707 and
// buildHeatDiffusionMatrix 7:709-737. The diagonal of cell c receives, from
// its initial value, the face terms in the scatter loop's order:
//   ((f_y(c-W) + f_x(c-1)) + f_x(c)) + f_y(c)
// each only where the face joins two FLUID cells (every face at level 3).
struct MatArgs {
    const uint8_t* cell;
    const double *uvol, *vvol, *ud, *vd;
    double *adiag, *apx, *apy;
    Skew sk;
    int level;
    int heat;
    double scale;
};

IFL_DEV bool face_x(const MatArgs& a, int x, int y) {  // face between (x, y) and (x+1, y)
    int W = a.sk.W;
    if (x < 0 || x >= W - 1) return false;
    if (a.level == 3 && !a.heat) return true;
    int i = x + y * W;
    return a.cell[i] == IFL_CELL_FLUID && a.cell[i + 1] == IFL_CELL_FLUID;
}
IFL_DEV bool face_y(const MatArgs& a, int x, int y) {  // face between (x, y) and (x, y+1)
    int W = a.sk.W, H = a.sk.H;
    if (y < 0 || y >= H - 1) return false;
    if (a.level == 3 && !a.heat) return true;
    int i = x + y * W;
    return a.cell[i] == IFL_CELL_FLUID && a.cell[i + W] == IFL_CELL_FLUID;
}
IFL_DEV double factor_x(const MatArgs& a, int x, int y) {
    if (a.heat || a.level <= 4) return a.scale;
    int uw = a.sk.W + 1;
    if (a.level <= 6) return a.scale * a.uvol[x + 1 + y * uw];
    return a.scale * a.uvol[x + 1 + y * uw] / a.ud[x + 1 + y * uw];
}
IFL_DEV double factor_y(const MatArgs& a, int x, int y) {
    if (a.heat || a.level <= 4) return a.scale;
    int W = a.sk.W;
    if (a.level <= 6) return a.scale * a.vvol[x + (y + 1) * W];
    return a.scale * a.vvol[x + (y + 1) * W] / a.vd[x + (y + 1) * (W + 1)];  // u-grid stride, Q2 (7:700)
}

IFL_DEV void matrix_body(const MatArgs& a, long gi) {
    int W = a.sk.W, H = a.sk.H;
    if (gi >= (long)W * H) return;
    int i = (int)gi;
    int x = i % W, y = i / W;
    bool fluid = a.level == 3 || a.cell[i] == IFL_CELL_FLUID;
    double diag = a.heat ? 1.0 : 0.0;
    double px = 0.0, py = 0.0;
    if (fluid || a.heat) {
        if (face_y(a, x, y - 1)) diag += factor_y(a, x, y - 1);
        if (face_x(a, x - 1, y)) diag += factor_x(a, x - 1, y);
        if (face_x(a, x, y)) {
            double f = factor_x(a, x, y);
            diag += f;
            px = -f;
        }
        if (face_y(a, x, y)) {
            double f = factor_y(a, x, y);
            diag += f;
            py = -f;
        }
    }
    long p = a.sk.pos(x, y);
    a.adiag[p] = diag;
    a.apx[p] = px;
    a.apy[p] = py;
}
__global__ void matrix_kernel(MatArgs a) { matrix_body(a, blockIdx.x * (long)blockDim.x + threadIdx.x); }

static int launch_matrix(Solver* s, int heat, double scale) {
    MatArgs a;
    a.cell = s->q[IFL_Q_D].cell;
    a.uvol = s->q[IFL_Q_U].volume;
    a.vvol = s->q[IFL_Q_V].volume;
    a.ud = s->udens;
    a.vd = s->vdens;
    a.adiag = s->adiag;
    a.apx = s->aplusx;
    a.apy = s->aplusy;
    a.sk = s->sk;
    a.level = s->level;
    a.heat = heat;
    a.scale = scale;
    matrix_kernel<<<grid_blocks((long)s->W * s->H, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

int k_build_pressure(Solver* s, double dt) {
    double hx = s->hx;
    double scale = s->level >= 7 ? dt / (hx * hx) : dt / (s->cfg.density * hx * hx);
    return launch_matrix(s, 0, scale);
}

int k_build_heat(Solver* s, double dt) {
    double hx = s->hx;
    double scale = s->cfg.diffusion * dt * 1.0 / (hx * hx);
    return launch_matrix(s, 1, scale);
}

// ------------------------------------------------------------ applyPressure
// 3:382-398 / 4:794-809 / 6:863-879 / 7:891-906. Face u(x, y) receives
// "+= term" from cell (x-1, y), then "-= term" from cell (x, y); v faces from
// (x, y-1) then (x, y). Level 3 then zeroes the walls.
struct PressArgs {
    double *u, *v;
    const double* p;
    const uint8_t* cell;
    const double *ud, *vd;
    Skew sk;
    int level;
    double scale;
};

IFL_DEV bool pcell_on(const PressArgs& a, int x, int y) {
    return a.level == 3 || a.cell[x + y * a.sk.W] == IFL_CELL_FLUID;
}

IFL_DEV void apply_pressure_body(const PressArgs& a, long i) {
    int W = a.sk.W, H = a.sk.H;
    long nu = (long)(W + 1) * H;
    long nv = (long)W * (H + 1);
    const bool dens = a.level >= 7;
    if (i < nu) {
        int x = (int)(i % (W + 1)), y = (int)(i / (W + 1));
        double u = a.u[i];
        if (x >= 1 && pcell_on(a, x - 1, y)) {
            double t = a.scale * a.p[a.sk.pos(x - 1, y)];
            u += dens ? t / a.ud[i] : t;
        }
        if (x < W && pcell_on(a, x, y)) {
            double t = a.scale * a.p[a.sk.pos(x, y)];
            u -= dens ? t / a.ud[i] : t;
        }
        if (a.level == 3 && (x == 0 || x == W)) u = 0.0;
        a.u[i] = u;
    } else if (i < nu + nv) {
        long k = i - nu;
        int x = (int)(k % W), y = (int)(k / W);
        double v = a.v[k];
        if (y >= 1 && pcell_on(a, x, y - 1)) {
            double t = a.scale * a.p[a.sk.pos(x, y - 1)];
            v += dens ? t / a.vd[k] : t;
        }
        if (y < H && pcell_on(a, x, y)) {
            double t = a.scale * a.p[a.sk.pos(x, y)];
            v -= dens ? t / a.vd[k] : t;
        }
        if (a.level == 3 && (y == 0 || y == H)) v = 0.0;
        a.v[k] = v;
    }
}
__global__ void apply_pressure_kernel(PressArgs a) { apply_pressure_body(a, blockIdx.x * (long)blockDim.x + threadIdx.x); }

int k_apply_pressure(Solver* s, double dt) {
    PressArgs a;
    a.u = s->q[IFL_Q_U].src;
    a.v = s->q[IFL_Q_V].src;
    a.p = s->p;
    a.cell = s->q[IFL_Q_D].cell;
    a.ud = s->udens;
    a.vd = s->vdens;
    a.sk = s->sk;
    a.level = s->level;
    a.scale = s->level >= 7 ? dt / s->hx : dt / (s->cfg.density * s->hx);
    long n = (long)(s->W + 1) * s->H + (long)s->W * (s->H + 1);
    apply_pressure_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(a);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ addBuoyancy 7:908-919
// v(x, y) += b(x, y-1)*0.5 and then b(x, y)*0.5, loop order.
IFL_DEV void buoyancy_body(double* __restrict__ v, const double* __restrict__ d, const double* __restrict__ t, int W,
                           int H, double dtg, double alpha, double t_amb, long k) {
    if (k >= (long)W * (H + 1)) return;
    int x = (int)(k % W), y = (int)(k / W);
    double val = v[k];
    if (y >= 1) {
        int c = x + (y - 1) * W;
        double b = dtg * (alpha * d[c] - (t[c] - t_amb) / t_amb);
        val += b * 0.5;
    }
    if (y < H) {
        int c = x + y * W;
        double b = dtg * (alpha * d[c] - (t[c] - t_amb) / t_amb);
        val += b * 0.5;
    }
    v[k] = val;
}
__global__ void buoyancy_kernel(double* __restrict__ v, const double* __restrict__ d, const double* __restrict__ t, int W,
                                int H, double dtg, double alpha, double t_amb) {
    buoyancy_body(v, d, t, W, H, dtg, alpha, t_amb, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

int k_add_buoyancy(Solver* s, double dt) {
    double rho = s->cfg.density;
    double alpha = (s->cfg.density_soot - rho) / rho;
    double dtg = dt * s->cfg.gravity;
    buoyancy_kernel<<<grid_blocks((long)s->W * (s->H + 1), 256), 256, 0, s->stream>>>(
        s->q[IFL_Q_V].src, s->q[IFL_Q_D].src, s->q[IFL_Q_T].src, s->W, s->H, dtg, alpha, s->cfg.t_amb);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ setBoundaryCondition 7:921-942
// A face shared by two SOLID cells takes the value written last in loop
// order (cell (x, y) after (x-1, y) / (x, y-1)); walls are zeroed last.
IFL_DEV void boundary_body(double* __restrict__ u, double* __restrict__ v, const uint8_t* __restrict__ cell,
                           const uint8_t* __restrict__ body, const Body* __restrict__ bodies, int W, int H, double hx,
                           long i) {
    long nu = (long)(W + 1) * H;
    long nv = (long)W * (H + 1);
    if (i < nu) {
        int x = (int)(i % (W + 1)), y = (int)(i / (W + 1));
        if (x == 0 || x == W) {
            u[i] = 0.0;
            return;
        }
        int c = -1;
        if (cell[x + y * W] == IFL_CELL_SOLID) c = x + y * W;
        else if (cell[x - 1 + y * W] == IFL_CELL_SOLID) c = x - 1 + y * W;
        if (c >= 0) u[i] = body_velx(bodies[body[c]], (y + 0.5) * hx);
    } else if (i < nu + nv) {
        long k = i - nu;
        int x = (int)(k % W), y = (int)(k / W);
        if (y == 0 || y == H) {
            v[k] = 0.0;
            return;
        }
        int c = -1;
        if (cell[x + y * W] == IFL_CELL_SOLID) c = x + y * W;
        else if (cell[x + (y - 1) * W] == IFL_CELL_SOLID) c = x + (y - 1) * W;
        if (c >= 0) v[k] = body_vely(bodies[body[c]], (x + 0.5) * hx);
    }
}
__global__ void boundary_kernel(double* __restrict__ u, double* __restrict__ v, const uint8_t* __restrict__ cell,
                                const uint8_t* __restrict__ body, const Body* __restrict__ bodies, int W, int H, double hx) {
    boundary_body(u, v, cell, body, bodies, W, H, hx, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

int k_set_boundary(Solver* s) {
    long n = (long)(s->W + 1) * s->H + (long)s->W * (s->H + 1);
    const Body* b = s->bodies_d;
    boundary_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->q[IFL_Q_U].src, s->q[IFL_Q_V].src, s->q[IFL_Q_D].cell,
                                                                s->q[IFL_Q_D].body, b, s->W, s->H, s->hx);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ layout conversion
IFL_DEV void skew_to_rows_body(const double* __restrict__ sk_arr, double* __restrict__ rows, const Skew& sk, long gi) {
    if (gi >= (long)sk.W * sk.H) return;
    int i = (int)gi;
    rows[i] = sk_arr[sk.pos(i % sk.W, i / sk.W)];
}
IFL_DEV void rows_to_skew_body(const double* __restrict__ rows, double* __restrict__ sk_arr, const Skew& sk, long gi) {
    if (gi >= (long)sk.W * sk.H) return;
    int i = (int)gi;
    sk_arr[sk.pos(i % sk.W, i / sk.W)] = rows[i];
}
__global__ void skew_to_rows_kernel(const double* __restrict__ sk_arr, double* __restrict__ rows, Skew sk) {
    skew_to_rows_body(sk_arr, rows, sk, blockIdx.x * (long)blockDim.x + threadIdx.x);
}
__global__ void rows_to_skew_kernel(const double* __restrict__ rows, double* __restrict__ sk_arr, Skew sk) {
    rows_to_skew_body(rows, sk_arr, sk, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

int k_skew_to_rows(Solver* s, const double* skew, double* rows) {
    skew_to_rows_kernel<<<grid_blocks((long)s->W * s->H, 256), 256, 0, s->stream>>>(skew, rows, s->sk);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}
int k_rows_to_skew(Solver* s, const double* rows, double* skew) {
    rows_to_skew_kernel<<<grid_blocks((long)s->W * s->H, 256), 256, 0, s->stream>>>(rows, skew, s->sk);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// memcpy(_r, _t->src()) 7:1001 and memcpy(_t->src(), _p) 7:1005
int k_copy_t_to_r(Solver* s) { return k_rows_to_skew(s, s->q[IFL_Q_T].src, s->r); }
int k_copy_p_to_t(Solver* s) { return k_skew_to_rows(s, s->p, s->q[IFL_Q_T].src); }

IFL_DEV void dflag_body(const uint8_t* __restrict__ cell, uint8_t* __restrict__ dflag, const Skew& sk, int all_fluid,
                        long gi) {
    if (gi >= (long)sk.W * sk.H) return;
    int i = (int)gi;
    dflag[sk.pos(i % sk.W, i / sk.W)] = (all_fluid || cell[i] == IFL_CELL_FLUID) ? 1 : 0;
}
__global__ void dflag_kernel(const uint8_t* __restrict__ cell, uint8_t* __restrict__ dflag, Skew sk, int all_fluid) {
    dflag_body(cell, dflag, sk, all_fluid, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

// One wave per strip column: the column's 64 flags as a lane mask, the
// form in which the sweeps select per lane without vector work.
IFL_DEV void fmask_body(const uint8_t* __restrict__ dflag, uint64_t* __restrict__ fmask, long n, long i) {
    bool f = i < n && dflag[i] != 0;
    uint64_t m = __ballot(f);
    if ((threadIdx.x & 63) == 0 && i < n) fmask[i >> 6] = m;
}
__global__ __launch_bounds__(256) void fmask_kernel(const uint8_t* __restrict__ dflag, uint64_t* __restrict__ fmask, long n) {
    fmask_body(dflag, fmask, n, blockIdx.x * (long)blockDim.x + threadIdx.x);
}

int k_update_dflag(Solver* s) {
    dflag_kernel<<<grid_blocks((long)s->W * s->H, 256), 256, 0, s->stream>>>(s->q[IFL_Q_D].cell, s->dflag, s->sk,
                                                                             s->level == 3 ? 1 : 0);
    IFL_LAUNCH_CHECK();
    long n = (long)s->sk.nstrips * s->sk.strip_sz;
    fmask_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->dflag, s->fmask, n);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}


// ------------------------------------------------------------ batched stages
// The batched form's grid stages: each stage of update() is one launch over
// a table of every solver's arguments, each workgroup mapped to (solver,
// block of that solver's grid; a second index for the quantity where the
// single-solver kernel has one) by a table built once per grid shape. The
// bodies are the single-solver kernels', on the same arrays, so every value
// is the one the solver would compute stepping alone.
struct FillArgs {
    QtyView q;
    const Body* bodies;
    int nb, level;
};
struct DflagArgs {
    const uint8_t* cell;
    uint8_t* dflag;
    uint64_t* fmask;
    Skew sk;
    long n;
    int all_fluid;
};
struct BoundArgs {
    double *u, *v;
    const uint8_t *cell, *body;
    const Body* bodies;
    int W, H;
    double hx;
};
struct DensArgs {
    const double *t, *d;
    double *ud, *vd;
    int W, H;
    double rho_air, t_amb, alpha;
};
struct BuoyArgs {
    double* v;
    const double *d, *t;
    int W, H;
    double dtg, alpha, t_amb;
};
struct CopyArgs {
    const double* src;
    double* dst;
    Skew sk;
};
struct FinishArgs {
    const double* src[4];  // the quantities checked (by the second index)
    long n[4];
    int* flags;            // this solver's two flags: non-finite, border
    const int* border;     // the border flag of its last extrapolation, or null
};

#define IFL_BATCH_STAGE(name, A, call)                                                                   \
    __global__ __launch_bounds__(256) void name(const A* __restrict__ t, const int2* __restrict__ map) { \
        const int2 m = map[blockIdx.x];                                                                   \
        const A& a = t[m.x];                                                                              \
        const long gi = (long)(m.y & 0xffffff) * 256 + threadIdx.x;                                       \
        const int by = m.y >> 24;                                                                         \
        (void)by;                                                                                         \
        call;                                                                                             \
    }

IFL_DEV void nonfinite_body(const FinishArgs& a, int by, long gi) {
    const bool bad = gi < a.n[by] && !isfinite(a.src[by][gi]);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(a.flags, 1);
}

IFL_BATCH_STAGE(bs_fill_phi_kernel, FillArgs, fill_phi_body(a.q, a.bodies, a.nb, gi))
IFL_BATCH_STAGE(bs_fill_cells_kernel, FillArgs, fill_cells_body(a.q, a.bodies, a.nb, a.level, gi))
IFL_BATCH_STAGE(bs_dflag_kernel, DflagArgs, dflag_body(a.cell, a.dflag, a.sk, a.all_fluid, gi))
IFL_BATCH_STAGE(bs_fmask_kernel, DflagArgs, fmask_body(a.dflag, a.fmask, a.n, gi))
IFL_BATCH_STAGE(bs_boundary_kernel, BoundArgs, boundary_body(a.u, a.v, a.cell, a.body, a.bodies, a.W, a.H, a.hx, gi))
IFL_BATCH_STAGE(bs_rhs_kernel, RhsArgs, build_rhs_body(a, gi))
IFL_BATCH_STAGE(bs_densities_kernel, DensArgs,
                densities_body(a.t, a.d, a.ud, a.vd, a.W, a.H, a.rho_air, a.t_amb, a.alpha, gi))
IFL_BATCH_STAGE(bs_matrix_kernel, MatArgs, matrix_body(a, gi))
IFL_BATCH_STAGE(bs_pressure_kernel, PressArgs, apply_pressure_body(a, gi))
IFL_BATCH_STAGE(bs_buoyancy_kernel, BuoyArgs, buoyancy_body(a.v, a.d, a.t, a.W, a.H, a.dtg, a.alpha, a.t_amb, gi))
IFL_BATCH_STAGE(bs_rows_to_skew_kernel, CopyArgs, rows_to_skew_body(a.src, a.dst, a.sk, gi))
IFL_BATCH_STAGE(bs_skew_to_rows_kernel, CopyArgs, skew_to_rows_body(a.src, a.dst, a.sk, gi))
IFL_BATCH_STAGE(bs_advect_kernel, AdvectArgs, advect_body(a, by, gi))
IFL_BATCH_STAGE(bs_nonfinite_kernel, FinishArgs, nonfinite_body(a, by, gi))

__global__ void bs_flags_kernel(const FinishArgs* __restrict__ t, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i].flags[0] = 0;
    t[i].flags[1] = t[i].border ? *t[i].border : 0;
}

// A block table: for each second index y, each solver's blocks of 256
// elements, counts[y * n + i] elements of solver i. Kept while the counts
// stay the same.
struct StageMap {
    std::vector<long> counts;
    int2* map = nullptr;
    int nblk = 0;
};

struct StageDev {
    std::vector<Solver*> solvers;
    hipStream_t stream = nullptr;
    char* arena_h = nullptr;  // pinned: the step's argument tables
    char* arena_d = nullptr;
    size_t cap = 0, off = 0;
    std::vector<StageMap> maps;
    int* flags_d = nullptr;   // 2 per solver
    int* flags_h = nullptr;
    std::vector<hipEvent_t> ready;
};

void stage_dev_free(StageDev* d) {
    if (!d) return;
    if (d->arena_h) (void)hipHostFree(d->arena_h);
    if (d->arena_d) (void)hipFree(d->arena_d);
    for (StageMap& m : d->maps)
        if (m.map) (void)hipFree(m.map);
    if (d->flags_d) (void)hipFree(d->flags_d);
    if (d->flags_h) (void)hipHostFree(d->flags_h);
    for (hipEvent_t e : d->ready) (void)hipEventDestroy(e);
    delete d;
}

namespace {

// the block table of map slot `slot` for these counts (ny second indices)
int stage_map(StageDev* d, int slot, const std::vector<long>& counts, const StageMap** out) {
    Solver* s = d->solvers[0];  // error reporting
    if ((int)d->maps.size() <= slot) d->maps.resize(slot + 1);
    StageMap& m = d->maps[slot];
    if (m.map && m.counts == counts) {
        *out = &m;
        return IFL_OK;
    }
    const int n = (int)d->solvers.size(), ny = (int)(counts.size() / n);
    std::vector<int2> h;
    for (int y = 0; y < ny; y++)
        for (int i = 0; i < n; i++)
            for (long b = 0; b < (counts[(size_t)y * n + i] + 255) / 256; b++) h.push_back(make_int2(i, (int)b | (y << 24)));
    if (m.map) IFL_HIP(hipFree(m.map));
    m.map = nullptr;
    m.nblk = 0;
    m.counts.clear();
    if (!h.empty()) {
        IFL_HIP(hipMalloc((void**)&m.map, h.size() * sizeof(int2)));
        IFL_HIP(hipMemcpy(m.map, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    m.nblk = (int)h.size();
    m.counts = counts;
    *out = &m;
    return IFL_OK;
}

// the step's table of A for every solver, written on the host and copied to
// the device on the stream (the arena is reused from the next step on, after
// the step has waited for the stream)
template <typename A, typename F>
int stage_table(StageDev* d, F fill, const A** out) {
    Solver* s = d->solvers[0];
    const int n = (int)d->solvers.size();
    const size_t at = (d->off + 255) & ~(size_t)255, bytes = n * sizeof(A);
    if (at + bytes > d->cap) {
        s->err = "batched stages: argument tables exceed their arena";
        return IFL_ERR_ALLOC;
    }
    A* h = (A*)(d->arena_h + at);
    for (int i = 0; i < n; i++) fill(d->solvers[i], h[i]);
    IFL_HIP(hipMemcpyAsync(d->arena_d + at, h, bytes, hipMemcpyHostToDevice, d->stream));
    d->off = at + bytes;
    *out = (const A*)(d->arena_d + at);
    return IFL_OK;
}

// map slots
enum {
    MS_CELLS = 0,   // W*H
    MS_FACES,       // (W+1)*H + W*(H+1)
    MS_VFACES,      // W*(H+1)
    MS_SKEW,        // the skewed arrays' length
    MS_FILL,        // + qi: q.w*q.h of solvers with bodies
    MS_PHI = MS_FILL + 4,  // + qi: (q.w+1)*(q.h+1) of solvers with bodies (levels 5+)
    MS_ADVECT = MS_PHI + 4,  // + qmask
    MS_NONFINITE = MS_ADVECT + 16,
    MS_COUNT
};

template <typename G>
std::vector<long> per_solver(StageDev* d, G g) {
    std::vector<long> c;
    for (Solver* s : d->solvers) c.push_back(g(s));
    return c;
}

#define IFL_BS_LAUNCH(kernel, tab, m)                                           \
    do {                                                                         \
        if ((m)->nblk > 0) {                                                     \
            kernel<<<(m)->nblk, 256, 0, d->stream>>>(tab, (m)->map);             \
            IFL_LAUNCH_CHECK();                                                  \
        }                                                                        \
    } while (0)

}  // namespace

int kb_begin(Solver* const* ss, int n, hipStream_t stream, StageDev** dev) {
    Solver* s = ss[0];
    StageDev* d = *dev;
    if (!d || d->solvers.size() != (size_t)n || !std::equal(ss, ss + n, d->solvers.begin())) {
        stage_dev_free(d);
        d = *dev = new StageDev();
        d->solvers.assign(ss, ss + n);
        // the largest table is advect's; a step writes fewer than 24 tables
        d->cap = (size_t)n * 24 * std::max<size_t>(sizeof(AdvectArgs), 256) + 24 * 256;
        IFL_HIP(hipHostMalloc((void**)&d->arena_h, d->cap, hipHostMallocDefault));
        IFL_HIP(hipMalloc((void**)&d->arena_d, d->cap));
        IFL_HIP(hipMalloc((void**)&d->flags_d, 2 * n * sizeof(int)));
        IFL_HIP(hipHostMalloc((void**)&d->flags_h, 2 * n * sizeof(int), hipHostMallocDefault));
        d->ready.assign(n, nullptr);
        for (int i = 0; i < n; i++) IFL_HIP(hipEventCreateWithFlags(&d->ready[i], hipEventDisableTiming));
        d->maps.resize(MS_COUNT);
    }
    d->stream = stream;
    // the previous step's table copies are complete before the arena is reused
    IFL_HIP(hipStreamSynchronize(stream));
    d->off = 0;
    // work the caller queued on a solver's own stream comes first
    for (int i = 0; i < n; i++) {
        IFL_HIP(hipEventRecord(d->ready[i], ss[i]->stream));
        IFL_HIP(hipStreamWaitEvent(stream, d->ready[i], 0));
    }
    return IFL_OK;
}

// k_fill_solid and (for d) k_update_dflag
int kb_fill_solid(StageDev* d, int qi) {
    Solver* s = d->solvers[0];
    const int lv = s->level;
    const FillArgs* tab;
    IFL_TRY(stage_table<FillArgs>(d, [&](Solver* si, FillArgs& a) {
        a.q = si->q[qi];
        a.bodies = si->bodies_d;
        a.nb = si->nbodies;
        a.level = si->level;
    }, &tab));
    const StageMap* m;
    if (lv >= 5) {
        IFL_TRY(stage_map(d, MS_PHI + qi, per_solver(d, [&](Solver* si) {
            return si->nbodies ? (long)(si->q[qi].w + 1) * (si->q[qi].h + 1) : 0L; }), &m));
        IFL_BS_LAUNCH(bs_fill_phi_kernel, tab, m);
    }
    IFL_TRY(stage_map(d, MS_FILL + qi, per_solver(d, [&](Solver* si) {
        return si->nbodies ? (long)si->q[qi].w * si->q[qi].h : 0L; }), &m));
    IFL_BS_LAUNCH(bs_fill_cells_kernel, tab, m);
    if (qi != IFL_Q_D) return IFL_OK;
    const DflagArgs* ft;
    IFL_TRY(stage_table<DflagArgs>(d, [&](Solver* si, DflagArgs& a) {
        a.cell = si->q[IFL_Q_D].cell;
        a.dflag = si->dflag;
        a.fmask = si->fmask;
        a.sk = si->sk;
        a.n = (long)si->sk.nstrips * si->sk.strip_sz;
        a.all_fluid = si->level == 3 ? 1 : 0;
    }, &ft));
    IFL_TRY(stage_map(d, MS_CELLS, per_solver(d, [](Solver* si) { return (long)si->W * si->H; }), &m));
    IFL_BS_LAUNCH(bs_dflag_kernel, ft, m);
    IFL_TRY(stage_map(d, MS_SKEW, per_solver(d, [](Solver* si) { return (long)si->sk.nstrips * si->sk.strip_sz; }),
                      &m));
    IFL_BS_LAUNCH(bs_fmask_kernel, ft, m);
    return IFL_OK;
}

static long faces(const Solver* si) { return (long)(si->W + 1) * si->H + (long)si->W * (si->H + 1); }

int kb_set_boundary(StageDev* d) {
    Solver* s = d->solvers[0];
    const BoundArgs* tab;
    IFL_TRY(stage_table<BoundArgs>(d, [&](Solver* si, BoundArgs& a) {
        a.u = si->q[IFL_Q_U].src;
        a.v = si->q[IFL_Q_V].src;
        a.cell = si->q[IFL_Q_D].cell;
        a.body = si->q[IFL_Q_D].body;
        a.bodies = si->bodies_d;
        a.W = si->W;
        a.H = si->H;
        a.hx = si->hx;
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_FACES, per_solver(d, faces), &m));
    IFL_BS_LAUNCH(bs_boundary_kernel, tab, m);
    return IFL_OK;
}

int kb_build_rhs(StageDev* d) {
    Solver* s = d->solvers[0];
    const RhsArgs* tab;
    IFL_TRY(stage_table<RhsArgs>(d, [&](Solver* si, RhsArgs& a) {
        a.u = si->q[IFL_Q_U].src;
        a.v = si->q[IFL_Q_V].src;
        a.uvol = si->q[IFL_Q_U].volume;
        a.vvol = si->q[IFL_Q_V].volume;
        a.dvol = si->q[IFL_Q_D].volume;
        a.cell = si->q[IFL_Q_D].cell;
        a.body = si->q[IFL_Q_D].body;
        a.r = si->r;
        a.sk = si->sk;
        a.level = si->level;
        a.nbodies = si->nbodies;
        a.hx = si->hx;
        a.bodies = si->bodies_d;
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_CELLS, per_solver(d, [](Solver* si) { return (long)si->W * si->H; }), &m));
    IFL_BS_LAUNCH(bs_rhs_kernel, tab, m);
    return IFL_OK;
}

int kb_compute_densities(StageDev* d) {
    Solver* s = d->solvers[0];
    const DensArgs* tab;
    IFL_TRY(stage_table<DensArgs>(d, [&](Solver* si, DensArgs& a) {
        const double rho = si->cfg.density;
        a.t = si->q[IFL_Q_T].src;
        a.d = si->q[IFL_Q_D].src;
        a.ud = si->udens;
        a.vd = si->vdens;
        a.W = si->W;
        a.H = si->H;
        a.rho_air = rho;
        a.t_amb = si->cfg.t_amb;
        a.alpha = (si->cfg.density_soot - rho) / rho;
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_FACES, per_solver(d, faces), &m));
    IFL_BS_LAUNCH(bs_densities_kernel, tab, m);
    return IFL_OK;
}

int kb_build_matrix(StageDev* d, int heat, double dt) {
    Solver* s = d->solvers[0];
    const MatArgs* tab;
    IFL_TRY(stage_table<MatArgs>(d, [&](Solver* si, MatArgs& a) {
        const double hx = si->hx;
        a.cell = si->q[IFL_Q_D].cell;
        a.uvol = si->q[IFL_Q_U].volume;
        a.vvol = si->q[IFL_Q_V].volume;
        a.ud = si->udens;
        a.vd = si->vdens;
        a.adiag = si->adiag;
        a.apx = si->aplusx;
        a.apy = si->aplusy;
        a.sk = si->sk;
        a.level = si->level;
        a.heat = heat;
        // k_build_heat / k_build_pressure
        a.scale = heat ? si->cfg.diffusion * dt * 1.0 / (hx * hx)
                       : si->level >= 7 ? dt / (hx * hx) : dt / (si->cfg.density * hx * hx);
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_CELLS, per_solver(d, [](Solver* si) { return (long)si->W * si->H; }), &m));
    IFL_BS_LAUNCH(bs_matrix_kernel, tab, m);
    return IFL_OK;
}

int kb_apply_pressure(StageDev* d, double dt) {
    Solver* s = d->solvers[0];
    const PressArgs* tab;
    IFL_TRY(stage_table<PressArgs>(d, [&](Solver* si, PressArgs& a) {
        a.u = si->q[IFL_Q_U].src;
        a.v = si->q[IFL_Q_V].src;
        a.p = si->p;
        a.cell = si->q[IFL_Q_D].cell;
        a.ud = si->udens;
        a.vd = si->vdens;
        a.sk = si->sk;
        a.level = si->level;
        a.scale = si->level >= 7 ? dt / si->hx : dt / (si->cfg.density * si->hx);
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_FACES, per_solver(d, faces), &m));
    IFL_BS_LAUNCH(bs_pressure_kernel, tab, m);
    return IFL_OK;
}

int kb_add_buoyancy(StageDev* d, double dt) {
    Solver* s = d->solvers[0];
    const BuoyArgs* tab;
    IFL_TRY(stage_table<BuoyArgs>(d, [&](Solver* si, BuoyArgs& a) {
        const double rho = si->cfg.density;
        a.v = si->q[IFL_Q_V].src;
        a.d = si->q[IFL_Q_D].src;
        a.t = si->q[IFL_Q_T].src;
        a.W = si->W;
        a.H = si->H;
        a.dtg = dt * si->cfg.gravity;
        a.alpha = (si->cfg.density_soot - rho) / rho;
        a.t_amb = si->cfg.t_amb;
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_VFACES, per_solver(d, [](Solver* si) { return (long)si->W * (si->H + 1); }), &m));
    IFL_BS_LAUNCH(bs_buoyancy_kernel, tab, m);
    return IFL_OK;
}

int kb_copy_t_to_r(StageDev* d) {
    Solver* s = d->solvers[0];
    const CopyArgs* tab;
    IFL_TRY(stage_table<CopyArgs>(d, [&](Solver* si, CopyArgs& a) {
        a.src = si->q[IFL_Q_T].src;
        a.dst = si->r;
        a.sk = si->sk;
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_CELLS, per_solver(d, [](Solver* si) { return (long)si->W * si->H; }), &m));
    IFL_BS_LAUNCH(bs_rows_to_skew_kernel, tab, m);
    return IFL_OK;
}

int kb_copy_p_to_t(StageDev* d) {
    Solver* s = d->solvers[0];
    const CopyArgs* tab;
    IFL_TRY(stage_table<CopyArgs>(d, [&](Solver* si, CopyArgs& a) {
        a.src = si->p;
        a.dst = si->q[IFL_Q_T].src;
        a.sk = si->sk;
    }, &tab));
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_CELLS, per_solver(d, [](Solver* si) { return (long)si->W * si->H; }), &m));
    IFL_BS_LAUNCH(bs_skew_to_rows_kernel, tab, m);
    return IFL_OK;
}

// k_advect with device-evaluated bodies in the FP64 form
int kb_advect(StageDev* d, int qmask, double dt) {
    Solver* s = d->solvers[0];
    int qs[4], nq = 0;
    for (int qi = 0; qi < 4; qi++)
        if (qmask & (1 << qi)) qs[nq++] = qi;
    if (nq == 0) return IFL_OK;
    const AdvectArgs* tab;
    IFL_TRY(stage_table<AdvectArgs>(d, [&](Solver* si, AdvectArgs& a) {
        a.nq = nq;
        for (int k = 0; k < nq; k++) a.q[k] = si->q[qs[k]];
        const QtyView& u = si->q[IFL_Q_U];
        const QtyView& v = si->q[IFL_Q_V];
        a.u = u.src;
        a.v = v.src;
        a.uw = u.w; a.uh = u.h; a.vw = v.w; a.vh = v.h;
        a.hx = si->hx;
        a.dt = dt;
        a.level = si->level;
        a.bodies = si->bodies_d;
        a.req = nullptr;
        a.req_cnt = nullptr;
        a.req_cap = 0;
    }, &tab));
    const int n = (int)d->solvers.size();
    std::vector<long> c((size_t)nq * n);
    for (int k = 0; k < nq; k++)
        for (int i = 0; i < n; i++) c[(size_t)k * n + i] = (long)d->solvers[i]->q[qs[k]].w * d->solvers[i]->q[qs[k]].h;
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_ADVECT + qmask, c, &m));
    IFL_BS_LAUNCH(bs_advect_kernel, tab, m);
    return IFL_OK;
}

int kb_finish(StageDev* d, bool border, int* nonfinite, int* border_solid) {
    Solver* s = d->solvers[0];
    const int n = (int)d->solvers.size();
    int qs[4], nq = 0;
    for (int qi = 0; qi < 4; qi++)
        if (qi != IFL_Q_T || s->level >= 6) qs[nq++] = qi;
    const FinishArgs* tab;
    int i0 = 0;
    IFL_TRY(stage_table<FinishArgs>(d, [&](Solver* si, FinishArgs& a) {
        for (int k = 0; k < 4; k++) {
            a.src[k] = k < nq ? si->q[qs[k]].src : nullptr;
            a.n[k] = k < nq ? (long)si->q[qs[k]].w * si->q[qs[k]].h : 0;
        }
        a.flags = d->flags_d + 2 * i0++;
        a.border = border ? si->front_cnt + 1 : nullptr;
    }, &tab));
    bs_flags_kernel<<<grid_blocks(n, 256), 256, 0, d->stream>>>(tab, n);
    IFL_LAUNCH_CHECK();
    std::vector<long> c((size_t)nq * n);
    for (int k = 0; k < nq; k++)
        for (int i = 0; i < n; i++) c[(size_t)k * n + i] = (long)d->solvers[i]->q[qs[k]].w * d->solvers[i]->q[qs[k]].h;
    const StageMap* m;
    IFL_TRY(stage_map(d, MS_NONFINITE, c, &m));
    IFL_BS_LAUNCH(bs_nonfinite_kernel, tab, m);
    IFL_HIP(hipMemcpyAsync(d->flags_h, d->flags_d, 2 * n * sizeof(int), hipMemcpyDeviceToHost, d->stream));
    IFL_HIP(hipStreamSynchronize(d->stream));
    for (int i = 0; i < n; i++) {
        nonfinite[i] = d->flags_h[2 * i];
        border_solid[i] = d->flags_h[2 * i + 1];
    }
    return IFL_OK;
}

// ------------------------------------------------------------ batched addInflow
// addInflow (7:456-473) of many solvers at once: one launch per rank (every
// solver's first inflow, then every second one ...), a row per (inflow,
// quantity) in the table, so a solver's inflows apply in their order.
IFL_BATCH_STAGE(bs_inflow_kernel, InflowArgs, inflow_body(a, gi))

struct InflowDev {
    InflowArgs* rows_h = nullptr;  // pinned
    InflowArgs* rows_d = nullptr;
    int2* map_h = nullptr;         // pinned
    int2* map_d = nullptr;
    size_t rows_cap = 0, map_cap = 0;
};

void inflow_dev_free(InflowDev* d) {
    if (!d) return;
    if (d->rows_h) (void)hipHostFree(d->rows_h);
    if (d->rows_d) (void)hipFree(d->rows_d);
    if (d->map_h) (void)hipHostFree(d->map_h);
    if (d->map_d) (void)hipFree(d->map_d);
    delete d;
}

int k_add_inflow_batch(Solver* const* ss, int n, hipStream_t stream, InflowDev** dev, int count, const int* inst,
                       const double* rect, const double* vals) {
    Solver* s = ss[0];  // error reporting
    if (!*dev) *dev = new InflowDev();
    InflowDev* d = *dev;
    std::vector<int> rank(count), seen(n, 0);
    int ranks = 0;
    for (int k = 0; k < count; k++) {
        rank[k] = seen[inst[k]]++;
        ranks = std::max(ranks, rank[k] + 1);
    }
    // rows and blocks, rank by rank
    std::vector<InflowArgs> rows;
    std::vector<int2> map;
    std::vector<size_t> row0(ranks + 1), blk0(ranks + 1);
    for (int r = 0; r < ranks; r++) {
        row0[r] = rows.size();
        blk0[r] = map.size();
        for (int k = 0; k < count; k++) {
            if (rank[k] != r) continue;
            const Solver* si = ss[inst[k]];
            const double* rc = rect + 4L * k;
            const double x0 = rc[0], y0 = rc[1], x1 = rc[0] + rc[2], y1 = rc[1] + rc[3];  // ifl_add_inflow
            for (int qi = 0; qi < 4; qi++) {
                if (qi == IFL_Q_T && si->level < 6) continue;
                InflowArgs a;
                if (!inflow_args(si, qi, x0, y0, x1, y1, vals[4L * k + qi], a)) continue;
                const int row = (int)(rows.size() - row0[r]);
                for (long b = 0; b < grid_blocks(inflow_threads(a), 256); b++) map.push_back(make_int2(row, (int)b));
                rows.push_back(a);
            }
        }
    }
    row0[ranks] = rows.size();
    blk0[ranks] = map.size();
    if (rows.size() > d->rows_cap) {
        if (d->rows_h) IFL_HIP(hipHostFree(d->rows_h));
        if (d->rows_d) IFL_HIP(hipFree(d->rows_d));
        d->rows_h = nullptr;
        d->rows_d = nullptr;
        d->rows_cap = 0;
        IFL_HIP(hipHostMalloc((void**)&d->rows_h, rows.size() * sizeof(InflowArgs), hipHostMallocDefault));
        IFL_HIP(hipMalloc((void**)&d->rows_d, rows.size() * sizeof(InflowArgs)));
        d->rows_cap = rows.size();
    }
    if (map.size() > d->map_cap) {
        if (d->map_h) IFL_HIP(hipHostFree(d->map_h));
        if (d->map_d) IFL_HIP(hipFree(d->map_d));
        d->map_h = nullptr;
        d->map_d = nullptr;
        d->map_cap = 0;
        IFL_HIP(hipHostMalloc((void**)&d->map_h, map.size() * sizeof(int2), hipHostMallocDefault));
        IFL_HIP(hipMalloc((void**)&d->map_d, map.size() * sizeof(int2)));
        d->map_cap = map.size();
    }
    if (rows.empty() || map.empty()) return IFL_OK;
    std::copy(rows.begin(), rows.end(), d->rows_h);
    std::copy(map.begin(), map.end(), d->map_h);
    IFL_HIP(hipMemcpyAsync(d->rows_d, d->rows_h, rows.size() * sizeof(InflowArgs), hipMemcpyHostToDevice, stream));
    IFL_HIP(hipMemcpyAsync(d->map_d, d->map_h, map.size() * sizeof(int2), hipMemcpyHostToDevice, stream));
    for (int r = 0; r < ranks; r++) {
        const int nblk = (int)(blk0[r + 1] - blk0[r]);
        if (nblk == 0) continue;
        bs_inflow_kernel<<<nblk, 256, 0, stream>>>(d->rows_d + row0[r], d->map_d + blk0[r]);
        IFL_LAUNCH_CHECK();
    }
    // the tables are rewritten by the next call
    IFL_HIP(hipStreamSynchronize(stream));
    return IFL_OK;
}

}  // namespace ifl
