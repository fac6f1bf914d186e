// Internal interface of the uniform-grid builder, shared by the single-scene
// entry points and the scene batch.
#pragma once

#include "sph.hpp"

struct reyn_grid_builder {
    reyn_ctx* ctx = nullptr;
    reyn_grid geom {}; // prefix/sorted/n filled by construct
    float search_radius = 0.f;
    size_t ncells = 0;
    size_t nblocks_scan = 0;
    uint32_t* counts = nullptr; // ncells
    uint32_t* prefix = nullptr; // ncells + 1
    uint32_t* block_sums = nullptr; // nblocks_scan
    uint32_t* status = nullptr; // [0] out-of-bounds points, [1] max cell count, [2] oversized cells
    // deferred checks: construct calls copy the status here and return before
    // it arrives; reyn_grid_builder_check (or the next construct) reads it
    uint32_t* status_pinned = nullptr;
    bool defer = false;
    bool pending = false; // a deferred build whose status is still unread
    uint32_t cell_cap = 0; // most points per cell a build accepts (walk_cell_cap of the range); 0: no limit
    uint32_t cap = 0; // per-point capacity of the arrays below
    uint32_t* keys = nullptr;
    uint32_t* slots = nullptr;
    uint32_t* tmp = nullptr;
    uint32_t* sorted = nullptr;
};

namespace reyn {

// The reference's grid geometry (uniformgrid.cu:65-96) with the cell size
// scaled by `cell_ratio`; no memory.
int grid_geometry(reyn_ctx* ctx, reyn_float3 bmin, reyn_float3 bmax, float search_radius, float cell_ratio,
    reyn_grid* out, size_t* ncells, int order = REYN_CELL_ORDER_XYZ);

// A builder whose cell tables cover `ncells` cells (a batch concatenates the
// cells of all its scenes); its geometry fields are unused.
int builder_alloc(reyn_ctx* ctx, size_t ncells, float search_radius, reyn_grid_builder** out);

// Counting sort of the n points of a scene batch over the builder's cells:
// point t of scene s gets the key cell_off[s] + its cell in scene s's grid.
// The build status is copied to `status_host` unless it is NULL (then it
// stays in b->status).
int sort_batch(reyn_grid_builder* b, const SceneArg& sa, uint32_t nblocks, const uint32_t* cell_off,
    const float* xx, const float* xy, const float* xz, uint32_t n, uint32_t* status_host);

// Waits for a build and turns its status into an error code.
int finish_build(reyn_grid_builder* b, const uint32_t* status_host);

int launch_gather(reyn_ctx* ctx, const uint32_t* sorted, uint32_t n, const float* const* src,
    float* const* dst, uint32_t nfields);

} // namespace reyn
