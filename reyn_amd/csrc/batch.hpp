// Internal interface between the scene batch and the solver steps.
#pragma once

#include "sph.hpp"

namespace reyn {

// The device state of a batch: concatenated fluid arrays, the scene table and
// the solver work arrays (p, a_ii, s_i only for IISPH).
struct BatchTarget {
    FluidDev f;
    SceneArg a;
    uint32_t nblocks, nscenes;
    float *rho, *p, *a_ii, *s_i, *ax, *ay, *az;
};

struct BatchSolver {
    int solver, kernel;
    float omega, eta;
    uint32_t min_iter, max_iter, check_every;
    int stop_rule, warm;
};

// One pressure-solver step of every scene (densities through integration),
// each scene with its own time step and loop control; iterations and the
// final loop statistic (PCISPH: mean density, IISPH: mean error) per scene.
int batch_solve(reyn_ctx* ctx, const BatchTarget& bt, const BatchSolver& bs, uint32_t* iters, double* errs);

} // namespace reyn
