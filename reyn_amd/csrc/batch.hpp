// Internal interface between the scene batch and the solver steps.
#pragma once

#include "sph.hpp"

namespace reyn {

// The device state of a batch: concatenated fluid arrays, the scene table, the
// solver work arrays (p, a_ii, s_i only for IISPH) and the batch's own scratch
// (its pointers must not move while a captured step refers to them).
struct BatchTarget {
    FluidDev f;
    SceneArg a; // a.blk_list: launch only the listed workgroups
    uint32_t nblocks, nscenes; // all workgroups, scenes
    uint32_t nlaunch; // workgroups launched with a.blk_list
    float *rho, *p, *a_ii, *s_i, *ax, *ay, *az;
    void* scratch;
};

struct BatchSolver {
    int solver, kernel;
    float omega, eta;
    uint32_t min_iter, max_iter, check_every;
    int stop_rule, warm;
};

// Scratch a batch of n particles in nblocks workgroups and nscenes scenes needs.
size_t batch_scratch_bytes(uint32_t n, uint32_t nblocks, uint32_t nscenes);

enum { BATCH_BEGIN = 0, BATCH_ITERATE = 1, BATCH_END = 2 };

// The solver step in phases, for every scene of the batch:
//   BATCH_BEGIN    densities and everything before the pressure loop, the
//                  loop-control reset, then `count` loop iterations
//   BATCH_ITERATE  `count` loop iterations (a scene whose loop is done skips)
//   BATCH_END      the final pressure accelerations and integration
int batch_phase(reyn_ctx* ctx, const BatchTarget& bt, const BatchSolver& bs, int phase, uint32_t count);

// The per-scene loop-control blocks in the batch's scratch, their size, and
// their host copy read out: iterations, the loop statistic (PCISPH: mean
// predicted density, IISPH: mean error) and whether every scene is done.
const void* batch_loop_control(const BatchTarget& bt);
size_t batch_loop_control_bytes(uint32_t nscenes);
void batch_read_control(const void* host, uint32_t nscenes, int solver, uint32_t* iters, double* stats,
    uint8_t* scene_done, int* done);

// Copies every scene's iteration count of the finished loop to `iters_dev`.
int batch_save_iterations(reyn_ctx* ctx, const BatchTarget& bt, uint32_t* iters_dev);

// Device-side dam-break scenes (scene.hip): sizes and generation into device
// arrays.
int dam_break_sizes(reyn_ctx* ctx, const reyn_dam_break* d, uint32_t* n_lattice, uint32_t* n_bdy);
int dam_break_generate(reyn_ctx* ctx, const reyn_dam_break* d, float* x, float* y, float* z, uint32_t* n_fluid,
    float* bx, float* by, float* bz);

} // namespace reyn
