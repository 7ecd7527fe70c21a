/*
 * ifl_fluid.hpp -- source-level drop-in for the reference's solver classes.
 *
 * A program written against one of the reference solvers
 * (N-<solver>/Fluid.cpp) keeps its main() and replaces the class
 * definitions with
 *
 *     #define IFL_SOLVER_LEVEL 7     // which reference solver: 3..8
 *     #include "ifl_fluid.hpp"
 *
 * The classes keep the reference's names, constructors and public methods
 * (SolidBox / SolidSphere 7:169-257, FluidSolver 7:591-1082); update() runs
 * on the GPU through the C ABI in ifl.h and prints the reference's stdout
 * lines. Solid bodies stay host objects that the caller owns and moves, as
 * in the reference (7:1124-1125); their state is sent with every update().
 *
 * User-defined SolidBody subclasses derive from SolidBody through the
 * reference's protected constructor and implement its three virtuals; while
 * the body list holds one, the solver evaluates every body's virtuals on the
 * host (ifl_set_host_body_ops) and the device the rest. Levels 4-7; the
 * device evaluates SolidBox and SolidSphere itself.
 *
 * Differences: errors throw std::runtime_error (the reference has no error
 * path); user-defined bodies are rejected at level 8; ~FluidSolver is safe at
 * every level (the reference's level-8 destructor double-deletes,
 * 8:887-888).
 * The environment variable IFL_DOT_ORDER=sequential selects dot products in
 * the reference's summation order.
 */
#ifndef IFL_FLUID_HPP
#define IFL_FLUID_HPP

#ifndef IFL_SOLVER_LEVEL
#define IFL_SOLVER_LEVEL 7
#endif

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "ifl.h"

namespace ifl_detail {
inline void check(int rc, const ifl_solver* s) {
    if (rc != IFL_OK) throw std::runtime_error(std::string("ifl: ") + ifl_last_error(s));
}
inline double length(double x, double y) { return sqrt(x * x + y * y); }
inline int nsgn(double a) { return a < 0.0 ? -1 : 1; }
inline void rotate(double& x, double& y, double phi) {
    double tx = x, ty = y;
    x = cos(phi) * tx + sin(phi) * ty;
    y = -sin(phi) * tx + cos(phi) * ty;
}
}  // namespace ifl_detail

enum CellType { CELL_FLUID = IFL_CELL_FLUID, CELL_SOLID = IFL_CELL_SOLID, CELL_EMPTY = IFL_CELL_EMPTY };

/* SolidBody 7:109-167 */
class SolidBody {
protected:
    double _posX, _posY, _scaleX, _scaleY, _theta;
    double _velX, _velY, _velTheta;
    int _type;

    /* globalToLocal / localToGlobal 7:119-133 */
    void globalToLocal(double& x, double& y) const {
        x -= _posX;
        y -= _posY;
        ifl_detail::rotate(x, y, -_theta);
        x /= _scaleX;
        y /= _scaleY;
    }
    void localToGlobal(double& x, double& y) const {
        x *= _scaleX;
        y *= _scaleY;
        ifl_detail::rotate(x, y, _theta);
        x += _posX;
        y += _posY;
    }

    /* the reference's constructor (7:135-138): a user-defined body, whose
     * virtuals the solver evaluates on the host */
    SolidBody(double posX, double posY, double scaleX, double scaleY, double theta, double velX, double velY,
              double velTheta)
        : _posX(posX), _posY(posY), _scaleX(scaleX), _scaleY(scaleY), _theta(theta), _velX(velX), _velY(velY),
          _velTheta(velTheta), _type(IFL_BODY_HOST) {}
    SolidBody(double posX, double posY, double scaleX, double scaleY, double theta, double velX, double velY,
              double velTheta, int type)
        : _posX(posX), _posY(posY), _scaleX(scaleX), _scaleY(scaleY), _theta(theta), _velX(velX), _velY(velY),
          _velTheta(velTheta), _type(type) {}

public:
    virtual ~SolidBody() {}
    /* 7:143-146 */
    virtual double distance(double x, double y) const = 0;
    virtual void closestSurfacePoint(double& x, double& y) const = 0;
    virtual void distanceNormal(double& nx, double& ny, double x, double y) const = 0;
    int deviceType() const { return _type; }

    double velocityX(double x, double y) const { (void)x; return (_posY - y) * _velTheta + _velX; }
    double velocityY(double x, double y) const { (void)y; return (x - _posX) * _velTheta + _velY; }
    void velocity(double& vx, double& vy, double x, double y) const {
        vx = velocityX(x, y);
        vy = velocityY(x, y);
    }
    void update(double timestep) {
        _posX += _velX * timestep;
        _posY += _velY * timestep;
        _theta += _velTheta * timestep;
    }
    /* the state the device needs, with cos/sin from this host's libm */
    ifl_body pod() const {
        ifl_body b;
        memset(&b, 0, sizeof b);
        b.type = _type;
        b.pos_x = _posX;
        b.pos_y = _posY;
        b.scale_x = _scaleX;
        b.scale_y = _scaleY;
        b.theta = _theta;
        b.vel_x = _velX;
        b.vel_y = _velY;
        b.vel_theta = _velTheta;
        b.cos_neg = cos(-_theta);
        b.sin_neg = sin(-_theta);
        b.cos_pos = cos(_theta);
        b.sin_pos = sin(_theta);
        return b;
    }
};

/* SolidBox 7:169-218 */
class SolidBox : public SolidBody {
public:
    SolidBox(double x, double y, double sx, double sy, double t, double vx, double vy, double vt)
        : SolidBody(x, y, sx, sy, t, vx, vy, vt, IFL_BODY_BOX) {}
    double distance(double x, double y) const {
        x -= _posX;
        y -= _posY;
        ifl_detail::rotate(x, y, -_theta);
        double dx = fabs(x) - _scaleX * 0.5;
        double dy = fabs(y) - _scaleY * 0.5;
        if (dx >= 0.0 || dy >= 0.0) return ifl_detail::length(std::max(dx, 0.0), std::max(dy, 0.0));
        return std::max(dx, dy);
    }
    /* 7:188-203 */
    void closestSurfacePoint(double& x, double& y) const {
        x -= _posX;
        y -= _posY;
        ifl_detail::rotate(x, y, -_theta);
        double dx = fabs(x) - _scaleX * 0.5;
        double dy = fabs(y) - _scaleY * 0.5;
        if (dx > dy)
            x = ifl_detail::nsgn(x) * 0.5 * _scaleX;
        else
            y = ifl_detail::nsgn(y) * 0.5 * _scaleY;
        ifl_detail::rotate(x, y, _theta);
        x += _posX;
        y += _posY;
    }
    /* 7:205-217 */
    void distanceNormal(double& nx, double& ny, double x, double y) const {
        x -= _posX;
        y -= _posY;
        ifl_detail::rotate(x, y, -_theta);
        if (fabs(x) - _scaleX * 0.5 > fabs(y) - _scaleY * 0.5) {
            nx = ifl_detail::nsgn(x);
            ny = 0.0;
        } else {
            nx = 0.0;
            ny = ifl_detail::nsgn(y);
        }
        ifl_detail::rotate(nx, ny, _theta);
    }
};

/* SolidSphere 7:220-257 */
class SolidSphere : public SolidBody {
public:
    SolidSphere(double x, double y, double s, double t, double vx, double vy, double vt)
        : SolidBody(x, y, s, s, t, vx, vy, vt, IFL_BODY_SPHERE) {}
    double distance(double x, double y) const { return ifl_detail::length(x - _posX, y - _posY) - _scaleX * 0.5; }
    /* 7:230-243 */
    void closestSurfacePoint(double& x, double& y) const {
        globalToLocal(x, y);
        double r = ifl_detail::length(x, y);
        if (r < 1e-4) {
            x = 0.5;
            y = 0.0;
        } else {
            x /= 2.0 * r;
            y /= 2.0 * r;
        }
        localToGlobal(x, y);
    }
    /* 7:245-256 (the radius is rounded to float, as there) */
    void distanceNormal(double& nx, double& ny, double x, double y) const {
        x -= _posX;
        y -= _posY;
        float r = ifl_detail::length(x, y);
        if (r < 1e-4) {
            nx = 1.0;
            ny = 0.0;
        } else {
            nx = x / r;
            ny = y / r;
        }
    }
};

/* FluidSolver: 3:187-466, 4:590-917, 5:628-970, 6:587-1058, 7:591-1082, 8:942-1467 */
class FluidSolver {
    ifl_solver* _s;
    const std::vector<const SolidBody*>* _bodies;
    int _w, _h;
    double _tAmb;

    void create(double density, double soot, double diffusion) {
        ifl_config cfg;
        ifl_detail::check(ifl_config_default(&cfg, IFL_SOLVER_LEVEL, _w, _h), nullptr);
        cfg.density = density;
        if (IFL_SOLVER_LEVEL >= 6) {
            cfg.density_soot = soot;
            cfg.diffusion = diffusion;
        }
        const char* order = getenv("IFL_DOT_ORDER");
        if (order && !strcmp(order, "sequential")) cfg.dot_order = IFL_DOT_SEQUENTIAL;
        _tAmb = cfg.t_amb;
        ifl_detail::check(ifl_create(&cfg, &_s), nullptr);
        push_bodies();
        /* the FLIP constructor seeds its particles from the bodies (8:1308-1314) */
        if (IFL_SOLVER_LEVEL == 8) ifl_detail::check(ifl_init_particles(_s), _s);
    }
    /* the virtuals of the caller's bodies, for lists that hold a user-defined body */
    static const SolidBody* body_at(void* user, int i) {
        return (*(const std::vector<const SolidBody*>*)user)[i];
    }
    static void cb_distance(void* user, int b, const double* x, const double* y, double* out, size_t n) {
        const SolidBody* body = body_at(user, b);
        for (size_t i = 0; i < n; i++) out[i] = body->distance(x[i], y[i]);
    }
    static void cb_closest(void* user, int b, double* x, double* y, size_t n) {
        const SolidBody* body = body_at(user, b);
        for (size_t i = 0; i < n; i++) body->closestSurfacePoint(x[i], y[i]);
    }
    static void cb_normal(void* user, int b, const double* x, const double* y, double* nx, double* ny, size_t n) {
        const SolidBody* body = body_at(user, b);
        for (size_t i = 0; i < n; i++) body->distanceNormal(nx[i], ny[i], x[i], y[i]);
    }
    void push_bodies() {
        if (!_bodies) return;
        std::vector<ifl_body> pods;
        bool host = false;
        for (size_t i = 0; i < _bodies->size(); i++) {
            pods.push_back((*_bodies)[i]->pod());
            host = host || pods.back().type == IFL_BODY_HOST;
        }
        if (host) {
            ifl_host_body_ops ops;
            ops.user = (void*)_bodies;
            ops.distance = cb_distance;
            ops.closest = cb_closest;
            ops.normal = cb_normal;
            ifl_detail::check(ifl_set_host_body_ops(_s, &ops), _s);
        }
        ifl_detail::check(ifl_set_bodies(_s, pods.empty() ? nullptr : &pods[0], (int)pods.size()), _s);
    }
    static void print_solve(const ifl_solve_stats& st) {
        if (st.status == IFL_SOLVE_CONVERGED)
            printf("Exiting solver after %d iterations, maximum error is %f\n", st.iterations, st.max_error);
        else if (st.status == IFL_SOLVE_BUDGET)
            printf("Exceeded budget of %d iterations, maximum error was %f\n", st.iterations, st.max_error);
        else if (st.status == IFL_SOLVE_INITIAL_SMALL && IFL_SOLVER_LEVEL >= 6)
            printf("Initial guess sufficiently small\n");
    }
    std::vector<double> field(int f) const {
        size_t n = 0;
        int e = 0;
        ifl_detail::check(ifl_field_info(_s, f, &n, &e), _s);
        std::vector<double> v(n);
        ifl_detail::check(ifl_download(_s, f, &v[0], n * sizeof(double)), _s);
        return v;
    }

public:
#if IFL_SOLVER_LEVEL == 3
    FluidSolver(int w, int h, double density) : _s(nullptr), _bodies(nullptr), _w(w), _h(h), _tAmb(294.0) {
        create(density, 0.0, 0.0);
    }
#elif IFL_SOLVER_LEVEL <= 5
    FluidSolver(int w, int h, double density, const std::vector<const SolidBody*>& bodies)
        : _s(nullptr), _bodies(&bodies), _w(w), _h(h), _tAmb(294.0) {
        create(density, 0.0, 0.0);
    }
#else
    FluidSolver(int w, int h, double rhoAir, double rhoSoot, double diffusion,
                const std::vector<const SolidBody*>& bodies)
        : _s(nullptr), _bodies(&bodies), _w(w), _h(h), _tAmb(294.0) {
        create(rhoAir, rhoSoot, diffusion);
    }
#endif
    ~FluidSolver() { ifl_destroy(_s); }

    void update(double timestep) {
        push_bodies();
        ifl_step_stats st;
        ifl_detail::check(ifl_update(_s, timestep, &st), _s);
        if (IFL_SOLVER_LEVEL == 8) printf("Particle count: %d\n", st.particle_count);
        if (IFL_SOLVER_LEVEL >= 6) print_solve(st.heat);
        print_solve(st.pressure);
    }

#if IFL_SOLVER_LEVEL <= 5
    void addInflow(double x, double y, double w, double h, double d, double u, double v) {
        double vals[4] = {d, 0.0, u, v};
        ifl_detail::check(ifl_add_inflow(_s, x, y, w, h, vals), _s);
    }
#else
    void addInflow(double x, double y, double w, double h, double d, double t, double u, double v) {
        double vals[4] = {d, t, u, v};
        ifl_detail::check(ifl_add_inflow(_s, x, y, w, h, vals), _s);
    }
    double ambientT() { return _tAmb; }
#endif

    /* toImage 3:455 / 4:903 / 5:961 / 6:1021 / 7:1047 / 8:1426, rendered on
     * the device (ifl_to_image) into the caller's w*h*4 (w*h*8 with the heat
     * panel) bytes */
#if IFL_SOLVER_LEVEL <= 5
    void toImage(unsigned char* rgba) {
        ifl_detail::check(ifl_to_image(_s, rgba, (size_t)_w * _h * 4, 0), _s);
    }
#else
    void toImage(unsigned char* rgba, bool renderHeat) {
        ifl_detail::check(ifl_to_image(_s, rgba, (size_t)_w * _h * 4 * (renderHeat ? 2 : 1), renderHeat ? 1 : 0), _s);
    }
#endif

    /* access for callers that read the state (the reference's fields are private) */
    ifl_solver* handle() { return _s; }
};

#endif /* IFL_FLUID_HPP */
