// Test infrastructure only: the reference's solver 1 (1-matrixless/Fluid.cpp)
// compiled from its own source tree, exposing maxTimestep (1:310-328) and its
// velocity grids so that the oracle's restatement of the CFL bound can be
// checked against it. Never linked into the product.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "lodepng/lodepng.h"

void* operator new[](std::size_t n) {
    void* p = std::calloc(n ? n : 1, 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

extern "C" int ifl_ref1_printf(const char*, ...) { return 0; }

#define main ifl_reference_main_unused
#define printf ifl_ref1_printf
#define class struct
#define private public
#define protected public
#include REF_SRC
#undef class
#undef private
#undef protected
#undef printf
#undef main

extern "C" {

void* ref1_create(int w, int h, double density) { return new FluidSolver(w, h, density); }

void ref1_add_inflow(void* s, double x, double y, double w, double h, double d, double u, double v) {
    static_cast<FluidSolver*>(s)->addInflow(x, y, w, h, d, u, v);
}

void ref1_update(void* s, double dt) { static_cast<FluidSolver*>(s)->update(dt); }

double ref1_max_timestep(void* s) { return static_cast<FluidSolver*>(s)->maxTimestep(); }

/* q: 2 = u ((w+1) x h), 3 = v (w x (h+1)); returns _src and its length */
double* ref1_velocity(void* s, int q, long* n) {
    FluidSolver* f = static_cast<FluidSolver*>(s);
    FluidQuantity* g = q == 2 ? f->_u : f->_v;
    *n = (long)g->_w * g->_h;
    return g->_src;
}

}  // extern "C"
