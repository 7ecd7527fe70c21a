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
__global__ void fill_phi_kernel(QtyView q, const Body* __restrict__ bodies, int nb) {
    int w1 = q.w + 1, h1 = q.h + 1;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w1 * h1) return;
    int ix = i % w1, iy = i / w1;
    double x = (ix + q.ox - 0.5) * q.hx;
    double y = (iy + q.oy - 0.5) * q.hx;
    double phi = body_distance(bodies[0], x, y);
    for (int k = 1; k < nb; k++) phi = smin(phi, body_distance(bodies[k], x, y));
    q.phi[i] = phi;
}

// body, volume, normal, cell per cell: 4:453-479 (level 4), 7:490-521 (5+)
__global__ void fill_cells_kernel(QtyView q, const Body* __restrict__ bodies, int nb, int level) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.w * q.h) return;
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
    ExtrapInst* inst_h[3] = {nullptr, nullptr, nullptr};  // pinned, one table per quantity
    ExtrapInst* inst_d[3] = {nullptr, nullptr, nullptr};
    int2* mask_map[3] = {nullptr, nullptr, nullptr};
    int2* border_map[3] = {nullptr, nullptr, nullptr};
    int nmask[3] = {0, 0, 0}, nborder[3] = {0, 0, 0};
    std::vector<hipEvent_t> ready;
    hipEvent_t joined = nullptr;
};

void extrap_dev_free(ExtrapDev* d) {
    if (!d) return;
    for (int k = 0; k < 3; k++) {
        if (d->inst_h[k]) (void)hipHostFree(d->inst_h[k]);
        if (d->inst_d[k]) (void)hipFree(d->inst_d[k]);
        if (d->mask_map[k]) (void)hipFree(d->mask_map[k]);
        if (d->border_map[k]) (void)hipFree(d->border_map[k]);
    }
    for (hipEvent_t e : d->ready) (void)hipEventDestroy(e);
    if (d->joined) (void)hipEventDestroy(d->joined);
    delete d;
}

int k_extrapolate_batch(Solver* const* ss, int n, hipStream_t stream, ExtrapDev** dev) {
    static const int kQ[3] = {IFL_Q_D, IFL_Q_U, IFL_Q_V};
    Solver* s = ss[0];  // error reporting
    ExtrapDev* d = *dev;
    if (!d || d->solvers.size() != (size_t)n || !std::equal(ss, ss + n, d->solvers.begin())) {
        extrap_dev_free(d);
        d = *dev = new ExtrapDev();
        d->solvers.assign(ss, ss + n);
        for (int k = 0; k < 3; k++) {
            std::vector<int2> mm, bm;
            for (int i = 0; i < n; i++) {
                const QtyView& q = ss[i]->q[kQ[k]];
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
        d->ready.assign(n, nullptr);
        for (int i = 0; i < n; i++) IFL_HIP(hipEventCreateWithFlags(&d->ready[i], hipEventDisableTiming));
        IFL_HIP(hipEventCreateWithFlags(&d->joined, hipEventDisableTiming));
    }
    // the solvers' src arrays swap every step, so the tables are written
    // afresh (the previous step's copies completed before it returned)
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < n; i++) {
            ExtrapInst& e = d->inst_h[k][i];
            e.q = ss[i]->q[kQ[k]];
            e.front = ss[i]->front;
            e.front_cnt = ss[i]->front_cnt;
            e.cap = (long)(ss[i]->W + 1) * (ss[i]->H + 1);
        }
    for (int i = 0; i < n; i++) {
        IFL_HIP(hipEventRecord(d->ready[i], ss[i]->stream));
        IFL_HIP(hipStreamWaitEvent(stream, d->ready[i], 0));
    }
    for (int k = 0; k < 3; k++) {
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
    IFL_HIP(hipEventRecord(d->joined, stream));
    for (int i = 0; i < n; i++) IFL_HIP(hipStreamWaitEvent(ss[i]->stream, d->joined, 0));
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
__global__ void advect_kernel(AdvectArgs a) {
    const QtyView& q = a.q[blockIdx.y];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.w * q.h) return;
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
                const int k = atomicAdd(&a.req_cnt[blockIdx.y], 1);
                HostReq* r = a.req + blockIdx.y * a.req_cap + k;
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
__global__ void inflow_kernel(QtyView q, int xa, int xb, int ya, int yb, double x0, double y0, double x1, double y1,
                              double v, int sequential) {
    long n = (long)q.w * q.h;
    int nx = xb - xa, ny = yb - ya;
    auto body = [&](int x, int y) {
        double l = length2((2.0 * (x + 0.5) * q.hx - (x0 + x1)) / (x1 - x0), (2.0 * (y + 0.5) * q.hx - (y0 + y1)) / (y1 - y0));
        double vi = cubic_pulse(l) * v;
        long idx = x + (long)y * q.w;
        if (idx < 0 || idx >= n) return;  // the reference writes past its array here
        if (fabs(q.src[idx]) < fabs(vi)) q.src[idx] = vi;
    };
    if (sequential) {  // x range wider than the row (Q3): rows alias, keep loop order
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int y = ya; y < yb; y++)
                for (int x = xa; x < xb; x++) body(x, y);
        return;
    }
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny) return;
    body(xa + i % nx, ya + i / nx);
}

int k_add_inflow(Solver* s, double x0, double y0, double x1, double y1, const double vals[4]) {
    for (int qi = 0; qi < 4; qi++) {
        if (qi == IFL_Q_T && s->level < 6) continue;
        QtyView q = s->q[qi];
        int ix0 = (int)(x0 / q.hx - q.ox);
        int iy0 = (int)(y0 / q.hx - q.oy);
        int ix1 = (int)(x1 / q.hx - q.ox);
        int iy1 = (int)(y1 / q.hx - q.oy);
        int ya = std::max(iy0, 0), yb = std::min(iy1, q.h);
        int xa = std::max(ix0, 0), xb = std::min(ix1, q.h);  // x bounded by _h, Q3 (7:463)
        if (yb <= ya || xb <= xa) continue;
        int seq = xb > q.w ? 1 : 0;
        long cnt = (long)(xb - xa) * (yb - ya);
        inflow_kernel<<<seq ? 1 : grid_blocks(cnt, 256), 256, 0, s->stream>>>(q, xa, xb, ya, yb, x0, y0, x1, y1, vals[qi], seq);
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

__global__ void build_rhs_kernel(RhsArgs a) {
    int W = a.sk.W, H = a.sk.H;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
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

__global__ void densities_kernel(const double* __restrict__ t, const double* __restrict__ d, double* __restrict__ ud,
                                 double* __restrict__ vd, int W, int H, double rho_air, double t_amb, double alpha) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
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

__global__ void matrix_kernel(MatArgs a) {
    int W = a.sk.W, H = a.sk.H;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
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

__global__ void apply_pressure_kernel(PressArgs a) {
    int W = a.sk.W, H = a.sk.H;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
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
__global__ void buoyancy_kernel(double* __restrict__ v, const double* __restrict__ d, const double* __restrict__ t, int W,
                                int H, double dtg, double alpha, double t_amb) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
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
__global__ void boundary_kernel(double* __restrict__ u, double* __restrict__ v, const uint8_t* __restrict__ cell,
                                const uint8_t* __restrict__ body, const Body* __restrict__ bodies, int W, int H, double hx) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
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

int k_set_boundary(Solver* s) {
    long n = (long)(s->W + 1) * s->H + (long)s->W * (s->H + 1);
    const Body* b = s->bodies_d;
    boundary_kernel<<<grid_blocks(n, 256), 256, 0, s->stream>>>(s->q[IFL_Q_U].src, s->q[IFL_Q_V].src, s->q[IFL_Q_D].cell,
                                                                s->q[IFL_Q_D].body, b, s->W, s->H, s->hx);
    IFL_LAUNCH_CHECK();
    return IFL_OK;
}

// ------------------------------------------------------------ layout conversion
__global__ void skew_to_rows_kernel(const double* __restrict__ sk_arr, double* __restrict__ rows, Skew sk) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sk.W * sk.H) return;
    rows[i] = sk_arr[sk.pos(i % sk.W, i / sk.W)];
}
__global__ void rows_to_skew_kernel(const double* __restrict__ rows, double* __restrict__ sk_arr, Skew sk) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sk.W * sk.H) return;
    sk_arr[sk.pos(i % sk.W, i / sk.W)] = rows[i];
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

__global__ void dflag_kernel(const uint8_t* __restrict__ cell, uint8_t* __restrict__ dflag, Skew sk, int all_fluid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sk.W * sk.H) return;
    dflag[sk.pos(i % sk.W, i / sk.W)] = (all_fluid || cell[i] == IFL_CELL_FLUID) ? 1 : 0;
}

// One wave per strip column: the column's 64 flags as a lane mask, the
// form in which the sweeps select per lane without vector work.
__global__ __launch_bounds__(256) void fmask_kernel(const uint8_t* __restrict__ dflag, uint64_t* __restrict__ fmask, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    bool f = i < n && dflag[i] != 0;
    uint64_t m = __ballot(f);
    if ((threadIdx.x & 63) == 0 && i < n) fmask[i >> 6] = m;
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

}  // namespace ifl
