// A reference-style main() (after 4:919-965, 7:1084-1129, 8:1469-1513) built against
// the source-level drop-in header instead of Fluid.cpp. Usage:
//   shim_main <w> <h> <steps> <density-out.bin>
// Prints what the reference prints per update() and writes the final density
// field (row-major doubles) for comparison by the tests.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "ifl_fluid.hpp"

int main(int argc, char** argv) {
    if (argc != 5) {
        fprintf(stderr, "usage: %s w h steps out.bin\n", argv[0]);
        return 2;
    }
    const int w = atoi(argv[1]), h = atoi(argv[2]), steps = atoi(argv[3]);
    const double dt = IFL_SOLVER_LEVEL == 8 ? 0.0025 : 0.005;  // 8:1477
    std::vector<const SolidBody*> bodies;
#if IFL_SOLVER_LEVEL >= 4
    SolidBox box(0.5, 0.6, 0.7, 0.1, M_PI * 0.25, 0.0, 0.0, 0.0);
    bodies.push_back(&box);
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
        if ((k + 1) % 4 == 0) box.update(dt);
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
