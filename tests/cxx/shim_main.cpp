// A reference-style main() (after 4:919-965, 7:1084-1129, 8:1469-1513) built against
// the source-level drop-in header instead of Fluid.cpp. Usage:
//   shim_main <w> <h> <steps> <density-out.bin>
//   shim_main <w> <h> <steps> <density-out.bin> user
// adds a user-defined SolidBody subclass (a sphere written out as a
// caller would write it), which the solver evaluates on the host.
// Prints what the reference prints per update() and writes the final density
// field (row-major doubles) for comparison by the tests.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "ifl_fluid.hpp"

#if IFL_SOLVER_LEVEL >= 4
// A caller's own body type, derived through the reference's protected
// constructor; its geometry is that of a sphere of diameter s.
class UserSphere : public SolidBody {
public:
    UserSphere(double x, double y, double s, double t, double vx, double vy, double vt)
        : SolidBody(x, y, s, s, t, vx, vy, vt) {}
    double distance(double x, double y) const {
        return sqrt((x - _posX) * (x - _posX) + (y - _posY) * (y - _posY)) - _scaleX * 0.5;
    }
    void closestSurfacePoint(double& x, double& y) const {
        globalToLocal(x, y);
        double r = sqrt(x * x + y * y);
        if (r < 1e-4) {
            x = 0.5;
            y = 0.0;
        } else {
            x /= 2.0 * r;
            y /= 2.0 * r;
        }
        localToGlobal(x, y);
    }
    void distanceNormal(double& nx, double& ny, double x, double y) const {
        x -= _posX;
        y -= _posY;
        float r = sqrt(x * x + y * y);
        if (r < 1e-4) {
            nx = 1.0;
            ny = 0.0;
        } else {
            nx = x / r;
            ny = y / r;
        }
    }
};
#endif

int main(int argc, char** argv) {
    if (argc != 5 && argc != 6) {
        fprintf(stderr, "usage: %s w h steps out.bin [user]\n", argv[0]);
        return 2;
    }
    const int w = atoi(argv[1]), h = atoi(argv[2]), steps = atoi(argv[3]);
    const bool user = argc == 6 && !strcmp(argv[5], "user");
    const double dt = IFL_SOLVER_LEVEL == 8 ? 0.0025 : 0.005;  // 8:1477
    std::vector<const SolidBody*> bodies;
#if IFL_SOLVER_LEVEL >= 4
    SolidBox box(0.5, 0.6, 0.7, 0.1, M_PI * 0.25, 0.0, 0.0, 0.0);
    bodies.push_back(&box);
    UserSphere ball(0.3, 0.35, 0.15, 0.3, 0.0, 0.0, 0.0);
    if (user) bodies.push_back(&ball);
#else
    (void)user;
#endif
#if IFL_SOLVER_LEVEL == 3
    FluidSolver solver(w, h, 0.1);
#elif IFL_SOLVER_LEVEL <= 5
    FluidSolver solver(w, h, 0.1, bodies);
#elif IFL_SOLVER_LEVEL <= 7
    FluidSolver solver(w, h, 0.1, 1.0, 0.01, bodies);
#else
    FluidSolver solver(w, h, 0.1, 0.25, 0.01, bodies);  // 8:1490; the inflow is inside update() (8:1368)
#endif
    for (int k = 0; k < steps; k++) {
#if IFL_SOLVER_LEVEL <= 5
        solver.addInflow(0.45, 0.2, 0.15, 0.03, 1.0, 0.0, 3.0);
#elif IFL_SOLVER_LEVEL <= 7
        solver.addInflow(0.45, 0.2, 0.1, 0.05, 1.0, solver.ambientT(), 0.0, 0.0);
#endif
        solver.update(dt);
#if IFL_SOLVER_LEVEL >= 4
        if ((k + 1) % 4 == 0) {
            box.update(dt);
            ball.update(dt);
        }
#endif
    }
    std::vector<unsigned char> image((size_t)w * h * 4 * 2);
#if IFL_SOLVER_LEVEL <= 5
    solver.toImage(&image[0]);
#else
    solver.toImage(&image[0], true);
#endif
    size_t n = (size_t)w * h;
    std::vector<double> d(n);
    if (ifl_download(solver.handle(), IFL_FIELD_Q(IFL_Q_D, IFL_QF_SRC), &d[0], n * sizeof(double)) != IFL_OK) {
        fprintf(stderr, "download failed: %s\n", ifl_last_error(solver.handle()));
        return 1;
    }
    FILE* f = fopen(argv[4], "wb");
    if (!f || fwrite(&d[0], sizeof(double), n, f) != n) return 1;
    fclose(f);
    return 0;
}
