// Shared device helpers for the gfx950 solver kernels.
//
// Arithmetic policy: every translation unit is compiled with
// -ffp-contract=off and without fast-math, so a*b+c is never fused and
// '/' and sqrt are the correctly rounded IEEE binary64 operations, matching
// the reference built with g++ -O2 on x86-64 (SSE2, no FMA). Expressions keep
// the reference's operand order; the citations (N:line = N-<solver>/Fluid.cpp)
// name the reference lines each helper restates.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ifl.h"

#define IFL_DEV __device__ __forceinline__
// also evaluated on the host (the host-evaluated bodies' solid fields)
#define IFL_HD __host__ __device__ __forceinline__

namespace ifl {

// std::min / std::max semantics (b < a ? b : a), including NaN behaviour.
IFL_HD double smin(double a, double b) { return (b < a) ? b : a; }
IFL_HD double smax(double a, double b) { return (a < b) ? b : a; }
IFL_DEV int imin(int a, int b) { return (b < a) ? b : a; }
IFL_DEV int imax(int a, int b) { return (a < b) ? b : a; }
// sgn / nsgn 7:37-43
IFL_DEV int sgn(double v) { return (0.0 < v) - (v < 0.0); }
IFL_DEV int nsgn(double v) { return v < 0.0 ? -1 : 1; }
// length 7:45-47
IFL_DEV double length2(double x, double y) { return sqrt(x * x + y * y); }
// cubicPulse 7:49-52
IFL_DEV double cubic_pulse(double x) {
    x = smin(fabs(x), 1.0);
    return 1.0 - x * x * (3.0 - 2.0 * x);
}
// (int) conversion of a coordinate that the caller already clamped; NaN maps
// to 0 so a corrupted input can never index outside an array.
IFL_DEV int trunc_idx(double x) { return (x == x) ? (int)x : 0; }

// frand's generator 8:38-46: seed = (seed*1103515245 + 12345) & 0x7FFFFFFF.
// Masking each step is arithmetic modulo 2^31, so n steps are one affine map
// s -> A*s + C (mod 2^31), built by repeated squaring (jump-ahead).
__host__ __device__ inline uint32_t lcg_next(uint32_t s) { return (s * 1103515245u + 12345u) & 0x7FFFFFFFu; }
__host__ __device__ inline uint32_t lcg_jump(uint32_t s, unsigned long long n) {
    if (n == 0) return s;
    uint32_t A = 1u, C = 0u, a = 1103515245u, c = 12345u;
    while (n) {
        if (n & 1) {
            C = a * C + c;
            A = a * A;
        }
        c = a * c + c;
        a = a * a;
        n >>= 1;
    }
    return (A * s + C) & 0x7FFFFFFFu;
}

// ---------------------------------------------------------------- bodies
// Device copy of a SolidBody (7:109-257) with host-evaluated cos/sin.
struct Body {
    int type;
    double px, py, sx, sy, theta, vx, vy, vt;
    double cn, sn;  // cos(-theta), sin(-theta)
    double cp, sp;  // cos(theta), sin(theta)
};

// rotate(x, y, phi) 7:54-58 with c = cos(phi), s = sin(phi) from the host
IFL_DEV void rotate_cs(double& x, double& y, double c, double s) {
    double tx = x, ty = y;
    x = c * tx + s * ty;
    y = -s * tx + c * ty;
}

// SolidBox::distance 7:175-186, SolidSphere::distance 7:226-228
IFL_DEV double body_distance(const Body& b, double x, double y) {
    if (b.type == IFL_BODY_SPHERE) return length2(x - b.px, y - b.py) - b.sx * 0.5;
    x -= b.px;
    y -= b.py;
    rotate_cs(x, y, b.cn, b.sn);
    double dx = fabs(x) - b.sx * 0.5;
    double dy = fabs(y) - b.sy * 0.5;
    if (dx >= 0.0 || dy >= 0.0) return length2(smax(dx, 0.0), smax(dy, 0.0));
    return smax(dx, dy);
}

// closestSurfacePoint: SolidBox 7:188-203, SolidSphere 7:230-243
IFL_DEV void body_closest(const Body& b, double& x, double& y) {
    if (b.type == IFL_BODY_SPHERE) {
        x -= b.px;
        y -= b.py;
        rotate_cs(x, y, b.cn, b.sn);
        x /= b.sx;
        y /= b.sy;
        double r = length2(x, y);
        if (r < 1e-4) {
            x = 0.5;
            y = 0.0;
        } else {
            x /= 2.0 * r;
            y /= 2.0 * r;
        }
        x *= b.sx;
        y *= b.sy;
        rotate_cs(x, y, b.cp, b.sp);
        x += b.px;
        y += b.py;
        return;
    }
    x -= b.px;
    y -= b.py;
    rotate_cs(x, y, b.cn, b.sn);
    double dx = fabs(x) - b.sx * 0.5;
    double dy = fabs(y) - b.sy * 0.5;
    if (dx > dy)
        x = nsgn(x) * 0.5 * b.sx;
    else
        y = nsgn(y) * 0.5 * b.sy;
    rotate_cs(x, y, b.cp, b.sp);
    x += b.px;
    y += b.py;
}

// distanceNormal: SolidBox 7:205-217, SolidSphere 7:245-256 (float r, Q4)
IFL_DEV void body_normal(const Body& b, double& nx, double& ny, double x, double y) {
    if (b.type == IFL_BODY_SPHERE) {
        x -= b.px;
        y -= b.py;
        float r = (float)length2(x, y);
        if (r < 1e-4) {
            nx = 1.0;
            ny = 0.0;
        } else {
            nx = x / (double)r;
            ny = y / (double)r;
        }
        return;
    }
    x -= b.px;
    y -= b.py;
    rotate_cs(x, y, b.cn, b.sn);
    if (fabs(x) - b.sx * 0.5 > fabs(y) - b.sy * 0.5) {
        nx = nsgn(x);
        ny = 0.0;
    } else {
        nx = 0.0;
        ny = nsgn(y);
    }
    rotate_cs(nx, ny, b.cp, b.sp);
}

// velocityX / velocityY 7:149-155
IFL_DEV double body_velx(const Body& b, double y) { return (b.py - y) * b.vt + b.vx; }
IFL_DEV double body_vely(const Body& b, double x) { return (x - b.px) * b.vt + b.vy; }

// occupancy 7:60-102
IFL_HD double tri_occ(double out1, double in, double out2) {
    return 0.5 * in * in / ((out1 - in) * (out2 - in));
}
IFL_HD double trap_occ(double out1, double out2, double in1, double in2) {
    return 0.5 * (-in1 / (out1 - in1) - in2 / (out2 - in2));
}
IFL_HD double occupancy(double d11, double d12, double d21, double d22) {
    unsigned b = (d11 < 0.0 ? 1u : 0u) | (d12 < 0.0 ? 2u : 0u) | (d22 < 0.0 ? 4u : 0u) | (d21 < 0.0 ? 8u : 0u);
    switch (b) {
    case 0x0: return 0.0;
    case 0x1: return tri_occ(d21, d11, d12);
    case 0x2: return tri_occ(d11, d12, d22);
    case 0x4: return tri_occ(d12, d22, d21);
    case 0x8: return tri_occ(d22, d21, d11);
    case 0xE: return 1.0 - tri_occ(-d21, -d11, -d12);
    case 0xD: return 1.0 - tri_occ(-d11, -d12, -d22);
    case 0xB: return 1.0 - tri_occ(-d12, -d22, -d21);
    case 0x7: return 1.0 - tri_occ(-d22, -d21, -d11);
    case 0x3: return trap_occ(d21, d22, d11, d12);
    case 0x6: return trap_occ(d11, d21, d12, d22);
    case 0x9: return trap_occ(d12, d22, d11, d21);
    case 0xC: return trap_occ(d11, d12, d21, d22);
    case 0x5: return tri_occ(d11, d12, d22) + tri_occ(d22, d21, d11);
    case 0xA: return tri_occ(d21, d11, d12) + tri_occ(d12, d22, d21);
    default: return 1.0;
    }
}

// ---------------------------------------------------------------- grids
// One staggered quantity (FluidQuantity 7:259-589): row-major x + y*w.
struct QtyView {
    int w, h;
    double ox, oy, hx;
    double* src;
    double* dst;
    double* phi;
    double* volume;
    double* nx;
    double* ny;
    uint8_t* cell;
    uint8_t* body;
    uint8_t* mask;
};

IFL_DEV double qat(const QtyView& q, int x, int y) { return q.src[x + y * q.w]; }

// lerp(a, b, x) 7:277-279
IFL_DEV double lerp1(double a, double b, double x) { return a * (1.0 - x) + b * x; }

// cerp(a, b, c, d, x) 7:281-295
IFL_DEV double cerp1(double a, double b, double c, double d, double x) {
    double xsq = x * x;
    double xcu = xsq * x;
    double minv = smin(a, smin(b, smin(c, d)));
    double maxv = smax(a, smax(b, smax(c, d)));
    double t = a * (0.0 - 0.5 * x + 1.0 * xsq - 0.5 * xcu) + b * (1.0 + 0.0 * x - 2.5 * xsq + 1.5 * xcu) +
               c * (0.0 + 0.5 * x + 2.0 * xsq - 1.5 * xcu) + d * (0.0 + 0.0 * x - 0.5 * xsq + 0.5 * xcu);
    return smin(smax(t, minv), maxv);
}

// lerp(x, y) 7:390-402
IFL_DEV double qlerp(const double* __restrict__ src, int w, int h, double ox, double oy, double x, double y) {
    x = smin(smax(x - ox, 0.0), w - 1.001);
    y = smin(smax(y - oy, 0.0), h - 1.001);
    int ix = trunc_idx(x);
    int iy = trunc_idx(y);
    x -= ix;
    y -= iy;
    const double* r0 = src + (size_t)iy * w + ix;
    double x00 = r0[0], x10 = r0[1];
    double x01 = r0[w], x11 = r0[w + 1];
    return lerp1(lerp1(x00, x10, x), lerp1(x01, x11, x), y);
}

// cerp(x, y) 7:404-422
IFL_DEV double qcerp(const double* __restrict__ src, int w, int h, double ox, double oy, double x, double y) {
    x = smin(smax(x - ox, 0.0), w - 1.001);
    y = smin(smax(y - oy, 0.0), h - 1.001);
    int ix = trunc_idx(x);
    int iy = trunc_idx(y);
    x -= ix;
    y -= iy;
    int x0 = imax(ix - 1, 0), x1 = ix, x2 = ix + 1, x3 = imin(ix + 2, w - 1);
    int ys[4] = {imax(iy - 1, 0), iy, iy + 1, imin(iy + 2, h - 1)};
    double qv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double* row = src + (size_t)ys[k] * w;
        qv[k] = cerp1(row[x0], row[x1], row[x2], row[x3], x);
    }
    return cerp1(qv[0], qv[1], qv[2], qv[3], y);
}

// The same samplers with the interpolation in single precision (the mixed
// form's advection): the sample point is located in double, the double
// fields are read and rounded to float, and the weights and sums are float.
IFL_DEV float lerp1f(float a, float b, float x) { return a * (1.0f - x) + b * x; }
IFL_DEV float cerp1f(float a, float b, float c, float d, float x) {
    float xsq = x * x;
    float xcu = xsq * x;
    float minv = fminf(a, fminf(b, fminf(c, d)));
    float maxv = fmaxf(a, fmaxf(b, fmaxf(c, d)));
    float t = a * (0.0f - 0.5f * x + 1.0f * xsq - 0.5f * xcu) + b * (1.0f + 0.0f * x - 2.5f * xsq + 1.5f * xcu) +
              c * (0.0f + 0.5f * x + 2.0f * xsq - 1.5f * xcu) + d * (0.0f + 0.0f * x - 0.5f * xsq + 0.5f * xcu);
    return fminf(fmaxf(t, minv), maxv);
}
IFL_DEV float qlerpf(const double* __restrict__ src, int w, int h, double ox, double oy, double x, double y) {
    x = smin(smax(x - ox, 0.0), w - 1.001);
    y = smin(smax(y - oy, 0.0), h - 1.001);
    int ix = trunc_idx(x);
    int iy = trunc_idx(y);
    const float fx = (float)(x - ix), fy = (float)(y - iy);
    const double* r0 = src + (size_t)iy * w + ix;
    float x00 = (float)r0[0], x10 = (float)r0[1];
    float x01 = (float)r0[w], x11 = (float)r0[w + 1];
    return lerp1f(lerp1f(x00, x10, fx), lerp1f(x01, x11, fx), fy);
}
IFL_DEV float qcerpf(const double* __restrict__ src, int w, int h, double ox, double oy, double x, double y) {
    x = smin(smax(x - ox, 0.0), w - 1.001);
    y = smin(smax(y - oy, 0.0), h - 1.001);
    int ix = trunc_idx(x);
    int iy = trunc_idx(y);
    const float fx = (float)(x - ix), fy = (float)(y - iy);
    int x0 = imax(ix - 1, 0), x1 = ix, x2 = ix + 1, x3 = imin(ix + 2, w - 1);
    int ys[4] = {imax(iy - 1, 0), iy, iy + 1, imin(iy + 2, h - 1)};
    float qv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double* row = src + (size_t)ys[k] * w;
        qv[k] = cerp1f((float)row[x0], (float)row[x1], (float)row[x2], (float)row[x3], fx);
    }
    return cerp1f(qv[0], qv[1], qv[2], qv[3], fy);
}

// ---------------------------------------------------------------- PCG layout
// The PCG vectors and matrix live in a strip-skewed layout: rows are grouped
// in strips of 64 (one wave, one lane per row); inside a strip, cell
// (x, y0 + l) is stored at strip_base + (x + l) * 64 + l. A wavefront step of
// the MIC(0) sweeps, which advances every lane by one column, then reads one
// contiguous 512-byte run per array. Unused slots (x < 0, x >= W, rows >= H)
// are kept at +0.0.
constexpr int kStrip = 64;

struct Skew {
    int W, H, nstrips;
    int cols;       // W + 63 skewed columns per strip
    long strip_sz;  // cols * 64 elements
    IFL_DEV long pos(int x, int y) const {
        int s = y >> 6, l = y & 63;
        return (long)s * strip_sz + (long)(x + l) * kStrip + l;
    }
};

// ---------------------------------------------------------------- lane moves
// Wave-wide shifts by one lane through DPP (GFX9 wave_shr:1 / wave_shl:1).
// Lanes without a source (lane 0 for shr, lane 63 for shl) receive `edge`.
IFL_DEV double wave_shr1(double v, double edge) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(__double2loint(edge), lo, 0x138, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(__double2hiint(edge), hi, 0x138, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
IFL_DEV double wave_shl1(double v, double edge) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(__double2loint(edge), lo, 0x130, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(__double2hiint(edge), hi, 0x130, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
IFL_DEV float wave_shr1f(float v, float edge) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
IFL_DEV float wave_shl1f(float v, float edge) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// Wave reductions in a fixed lane order (deterministic).
IFL_DEV double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
IFL_DEV double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace ifl
