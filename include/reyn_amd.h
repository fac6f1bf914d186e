/*
 * reyn_amd.h — C-ABI of the MI355X (gfx950) SPH per-timestep hot path.
 *
 * Every entry point takes plain device pointers and sizes, enqueues its work
 * on the stream owned by a `reyn_ctx`, returns an `int` status (0 = ok,
 * negative = error) and never throws or exits. `reyn_last_error(ctx)` holds a
 * message for the last failing call on that context; after a successful
 * solver step it may hold a note that the step walked the grid in every pass
 * because the input exceeded the 32-bit offsets of the neighbour list or the
 * packed records. Caller memory is never
 * freed by the library; buffers the library allocates itself (grid-builder
 * tables, solver scratch) belong to the handle that allocated them.
 *
 * The reference (JulianKarrer/reyn) exposes this path only as C++ templates
 * with device lambdas; each declaration below cites the reference interface
 * it replaces as `path:line` in that repository. The C++ host API in
 * include/reyn/ restores the reference's class shapes on top of these calls.
 */
#ifndef REYN_AMD_H_
#define REYN_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the library is built with hidden visibility; everything declared here is
 * exported */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#define REYN_ABI_VERSION 3

/* ---- status codes ------------------------------------------------------ */
enum {
    REYN_OK = 0,
    REYN_ERR_INVALID_ARGUMENT = -1,
    REYN_ERR_HIP = -2,
    REYN_ERR_OUT_OF_MEMORY = -3,
    /* a particle lay outside the cells a query may touch; it was clamped to
     * the grid so no memory was accessed out of bounds */
    REYN_ERR_OUT_OF_BOUNDS = -4,
    REYN_ERR_NOT_FINITE = -5,
    REYN_ERR_NOT_CONVERGED = -6,
    REYN_ERR_CAPACITY = -7
};

/* ---- plain value types ------------------------------------------------- */

/* Smoothing kernel families; replaces the CRTP `Kernel<D>` classes and the
 * `FOREACH_KERNEL` list (src/kernels.cuh:19-174). */
typedef enum reyn_kernel {
    REYN_KERNEL_C2 = 0, /* Wendland C2, src/kernels.cuh:78-95 */
    REYN_KERNEL_B3 = 1, /* cubic B-spline, src/kernels.cuh:101-118 */
    REYN_KERNEL_W6 = 2, /* Wendland C6, src/kernels.cuh:121-145 */
    REYN_KERNEL_COS = 3 /* double cosine, src/kernels.cuh:148-165 */
} reyn_kernel;

typedef struct reyn_float3 {
    float x, y, z;
} reyn_float3;

/* The order of the cell keys, and so of a reordered state.
 * REYN_CELL_ORDER_XYZ is the reference's key z nx ny + y nx + x
 * (src/datastructure/uniformgrid.cuh:33-40). REYN_CELL_ORDER_BRICK groups the
 * cells in 2 x 2 x 2 bricks: key = 8 brick + 4 (z & 1) + 2 (y & 1) + (x & 1)
 * with brick = (z/2) (nx/2) (ny/2) + (y/2) (nx/2) + x/2, the counts nx, ny, nz
 * rounded up to even (a one-level Z-order: the 64 points of a wavefront sit
 * in about one brick instead of one row of cells). No counterpart in the
 * reference, whose only Morton codes order mesh faces
 * (src/datastructure/lbvh.cuh:17-45). */
enum { REYN_CELL_ORDER_XYZ = 0, REYN_CELL_ORDER_BRICK = 1 };

/* Query-side view of a uniform grid; replaces the POD `UniformGrid<R>` and
 * `MaybeSorted<R>` (src/datastructure/uniformgrid.cuh:70-106).
 * `sorted == NULL` is `Resort::yes` (the indexed arrays are stored in cell
 * order); otherwise it is `Resort::no` and `sorted[k]` is the index of the
 * k-th point in cell order. `prefix` has ncells + 1 entries; the last one is
 * the number of points. */
typedef struct reyn_grid {
    reyn_float3 bound_min; /* padded lower corner of cell (0,0,0) */
    float cell_size;
    float r_c_2; /* squared search radius, formed in float */
    int32_t nx, ny, nz;
    int32_t range; /* cells searched on each side of a query's cell */
    const uint32_t* prefix;
    const uint32_t* sorted;
    uint32_t n; /* number of indexed points */
    int32_t order; /* REYN_CELL_ORDER_* */
} reyn_grid;

/* Fluid particle state in SoA layout; replaces `Particles`
 * (src/particles.cuh:25-75). All pointers are device pointers of n floats. */
typedef struct reyn_particles {
    float *xx, *xy, *xz;
    float *vx, *vy, *vz;
    float* m;
    uint32_t n;
} reyn_particles;

/* Static boundary samples with their own cell-ordered grid; replaces the POD
 * `Boundary` (src/scene/scene.cuh:14-20). `grid.sorted` must be NULL. */
typedef struct reyn_boundary {
    const float *xx, *xy, *xz, *m;
    uint32_t n;
    reyn_grid grid;
    /* optional: per cell of `grid`, nonzero when a sample lies in the
     * (2 range + 1)^3 cells around it (`reyn_grid_occupancy`); a fluid
     * particle in a cell marked zero has no boundary neighbour and skips the
     * boundary walk. NULL: every particle walks the boundary grid. */
    const uint8_t* near;
} reyn_boundary;

/* Constants shared by every solver (the constructor arguments of SESPH,
 * PCISPH and IISPH: src/solvers/SESPH.cuh:37-40, PCISPH.cuh:80-82,
 * IISPH.cuh:51-55). `h` is the particle spacing, `h_bar` the kernel support. */
typedef struct reyn_sph_params {
    int32_t kernel; /* reyn_kernel */
    float h_bar;
    float h;
    float nu;
    float rho_0;
    reyn_float3 g;
} reyn_sph_params;

/* ---- context, memory and timing ---------------------------------------- */

typedef struct reyn_ctx reyn_ctx;

/* Creates a context bound to HIP device `device` with its own stream of
 * default priority. No reference counterpart (the reference uses the default
 * stream and global state, src/common.h:11-36). */
int reyn_ctx_create(int device, reyn_ctx** out);
int reyn_ctx_destroy(reyn_ctx* ctx);
const char* reyn_last_error(const reyn_ctx* ctx);
int reyn_ctx_synchronize(reyn_ctx* ctx);
/* The context's hipStream_t, as an opaque pointer. */
void* reyn_ctx_stream(reyn_ctx* ctx);
int reyn_device_count(int* count);
/* Free and total device memory of the context's GPU, in bytes. */
int reyn_device_memory(reyn_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
int reyn_abi_version(void);

/* Device memory helpers (replace `DeviceBuffer<T>` storage,
 * src/buffer.cuh:15-269, for hosts that have no HIP binding of their own). */
int reyn_malloc(reyn_ctx* ctx, void** ptr, size_t bytes);
int reyn_free(reyn_ctx* ctx, void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device; synchronous with
 * respect to the context's stream. */
int reyn_memcpy(reyn_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
int reyn_fill_u32(reyn_ctx* ctx, uint32_t* ptr, uint32_t value, size_t count);
int reyn_fill_f32(reyn_ctx* ctx, float* ptr, float value, size_t count);

/* Event timing on the context's stream: `reyn_timer_start` records a start
 * event, `reyn_timer_stop` records a stop event, waits for it and returns the
 * elapsed milliseconds. */
int reyn_timer_start(reyn_ctx* ctx);
int reyn_timer_stop(reyn_ctx* ctx, float* ms);

/* Per-launch profiling with HIP events on the context's stream (no
 * reference counterpart; the reference has no profiler hooks, see its
 * nanobench micro-benchmarks, src/datastructure/uniformgrid.cu:385-423).
 * `reyn_profile_enable` clears the records and switches recording on or off;
 * `reyn_profile_select` limits recording to the launches of one kernel
 * (`name`, e.g. "k_iisph_jacobi"; NULL or "" records every kernel), so the
 * events of the other launches do not weigh on a timed run;
 * `reyn_profile_summary` waits for the stream and writes one line per kernel,
 * "name launches total_ms working_launches working_total_ms", into buf (at
 * most len bytes, NUL-terminated); *needed receives the full size. Working
 * launches are those that took at least a fifth of the kernel's longest
 * launch: a solver-loop launch that finds its loop already converged returns
 * at once and is not one of them (on scenes large enough that a working
 * launch takes more than five times as long; on small scenes some are). */
int reyn_profile_enable(reyn_ctx* ctx, int on);
int reyn_profile_select(reyn_ctx* ctx, const char* name);
int reyn_profile_summary(reyn_ctx* ctx, char* buf, size_t len, size_t* needed);

/* ---- uniform grid (src/datastructure/uniformgrid.cu:65-181) ------------- */

typedef struct reyn_grid_builder reyn_grid_builder;

/* Replaces `UniformGridBuilder::UniformGridBuilder(bound_min, bound_max,
 * search_radius)` (src/datastructure/uniformgrid.cu:65-96). The reference
 * fixes cell_ratio = 0.5 (cell = search_radius / 2, :69); this entry point
 * takes the ratio of cell size to search radius (0.5, 1.0 and 2.0 are the
 * supported sweep values; any ratio in [0.25, 4] works). The lower bound is
 * padded by 2 * search_radius and the cell counts are ceil-divided as in
 * :73-80. Fails with REYN_ERR_INVALID_ARGUMENT where the reference throws
 * (:93-95). */
int reyn_grid_builder_create(reyn_ctx* ctx, reyn_float3 bound_min,
    reyn_float3 bound_max, float search_radius, float cell_ratio,
    reyn_grid_builder** out);
/* The same with the cell order `order` (REYN_CELL_ORDER_*); the builder above
 * is REYN_CELL_ORDER_XYZ. */
int reyn_grid_builder_create_ordered(reyn_ctx* ctx, reyn_float3 bound_min,
    reyn_float3 bound_max, float search_radius, float cell_ratio, int32_t order,
    reyn_grid_builder** out);
int reyn_grid_builder_destroy(reyn_grid_builder* b);
/* Geometry of the builder (prefix/sorted are NULL until a construct call). */
int reyn_grid_builder_geometry(const reyn_grid_builder* b, reyn_grid* out);

/* Replaces `UniformGridBuilder::construct` (uniformgrid.cu:145-156) and its
 * `_construct` (uniformgrid.cu:98-143): counts points per cell, scans, and
 * writes `sorted` — a stable counting sort (ascending point index inside a
 * cell; the reference's in-cell order is racy, uniformgrid.cu:59). The
 * returned grid is `Resort::no` and points into builder-owned tables that
 * stay valid until the next construct call on the same builder. Returns
 * REYN_ERR_OUT_OF_BOUNDS (after finishing the build with clamped cells) if a
 * point lies outside the cells that queries may touch. */
int reyn_grid_construct(reyn_ctx* ctx, reyn_grid_builder* b,
    float search_radius, const float* xx, const float* xy, const float* xz,
    uint32_t n, reyn_grid* out);

/* Replaces `construct_and_reorder(search_radius, tmp, xx, xy, xz, bufs...)`
 * and the `Particles&` overloads (uniformgrid.cuh:243-483,
 * uniformgrid.cu:158-181; `DeviceBuffer::reorder`, buffer.cuh:184-209).
 * fields[0..2] must be x, y, z; fields[3..nfields) are further per-point
 * float arrays. Every field is gathered into tmp[k] in cell order and then
 * the caller's pointers fields[k] and tmp[k] are swapped, as the reference
 * swaps storage with its tmp buffer. The returned grid is `Resort::yes`. */
int reyn_grid_construct_and_reorder(reyn_ctx* ctx, reyn_grid_builder* b,
    float search_radius, float** fields, float** tmp, uint32_t nfields,
    uint32_t n, reyn_grid* out);

/* The most points one cell of a grid whose queries search `range` cells on
 * each side may hold: a build in which a cell holds more fails with
 * REYN_ERR_CAPACITY (after finishing its tables), so that no neighbour walk
 * over a built grid, (2 range + 1)^3 cells, exceeds its budget of 2^15
 * candidates. Host only. No reference counterpart (the reference's walk has
 * no bound, uniformgrid.cuh:113-178). */
int reyn_grid_cell_capacity(int32_t range, uint32_t* cap);

/* Deferred build checks (on != 0): the construct calls above enqueue the
 * build and return REYN_OK at once, without waiting for its status; the grid
 * they return may be used by work enqueued after them. The errors a build
 * would have returned (REYN_ERR_OUT_OF_BOUNDS, REYN_ERR_CAPACITY) come from
 * reyn_grid_builder_check, which waits for the stream; a construct call
 * whose predecessor was not checked checks it first and, if that build
 * failed, returns its error without building (nor reordering).
 * A caller that reads the device anyway (a time-step read-back) checks the
 * build after that read and saves one wait per step. Switching deferral off
 * checks a pending build. No reference counterpart (the reference's build
 * reports nothing, uniformgrid.cu:98-143). */
int reyn_grid_builder_defer_check(reyn_grid_builder* b, int32_t on);
/* The status of the last deferred build: REYN_OK, or the error its construct
 * call would have returned; REYN_OK when no build is pending. */
int reyn_grid_builder_check(reyn_ctx* ctx, reyn_grid_builder* b);

/* Device copies of the builder's last per-cell counts (ncells) and per-point
 * cell keys (n), for tests. */
int reyn_grid_debug_tables(reyn_grid_builder* b, const uint32_t** counts,
    const uint32_t** keys);

/* Gather: dst[k][i] = src[k][sorted[i]] for every field k
 * (buffer.cuh:184-209, thrust::gather). */
int reyn_gather_f32(reyn_ctx* ctx, const uint32_t* sorted, uint32_t n,
    const float* const* src, float* const* dst, uint32_t nfields);

/* The reference's grid test and micro-benchmark kernels
 * (uniformgrid.cu:185-245): for every point i of (xx,xy,xz) the number of
 * neighbours j != i with |x_ij|^2 <= r_c_2, the sum of |x_ij|^2 and of x_ij,
 * found through `grid`. Any output may be NULL. */
int reyn_grid_query_stats(reyn_ctx* ctx, const reyn_grid* grid,
    const float* xx, const float* xy, const float* xz, uint32_t n,
    uint32_t* count, float* len2, float* vx, float* vy, float* vz);
/* near[c] = 1 when a point of `grid` lies in the (2 range + 1)^3 cells
 * around cell c, else 0, for all nx ny nz cells (the boundary view's
 * `near`). No reference counterpart: the reference walks the boundary grid
 * for every fluid particle (solver.cuh:18-38). */
int reyn_grid_occupancy(reyn_ctx* ctx, const reyn_grid* grid, uint8_t* near);
/* res[i] = sum_j x_i.x * x_j.x * W(x_ij) with the given kernel
 * (`_benchmark_kernel`, uniformgrid.cu:231-245). */
int reyn_grid_query_bench(reyn_ctx* ctx, const reyn_grid* grid,
    int32_t kernel, float h_bar, const float* xx, const float* xy,
    const float* xz, uint32_t n, float* res);

/* ---- smoothing kernels (src/kernels.cuh:19-174) ------------------------- */

/* Evaluates W and grad W at n offsets (the reference's `_test_kernels`,
 * src/kernels.cu:6-15). */
int reyn_kernel_eval(reyn_ctx* ctx, int32_t kernel, float h_bar,
    const float* dx, const float* dy, const float* dz, uint32_t n, float* w,
    float* gx, float* gy, float* gz);

/* ---- shared solver phases (src/solvers/solver.cuh, solver.cu) ---------- */

/* rho_i = sum_j m_j W_ij + sum_k m_k W_ik, compensated sum
 * (`compute_densities`, src/solvers/solver.cuh:18-38). bdy may be NULL. */
int reyn_density(reyn_ctx* ctx, const reyn_sph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid,
    const reyn_boundary* bdy, float* rho);

/* Gravity plus fluid-fluid viscosity (`write_non_prs_acc_ff`,
 * src/solvers/solver.cuh:40-69). */
int reyn_nonpressure_acc(reyn_ctx* ctx, const reyn_sph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid, const float* rho,
    float* ax, float* ay, float* az);

/* PCISPH predicted density (`predict_rho`, src/solvers/solver.cuh:71-105). */
int reyn_predict_rho(reyn_ctx* ctx, const reyn_sph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid,
    const reyn_boundary* bdy, float dt, float* rho);

/* Semi-implicit Euler (`integrate`, `integrate_a_v_only`,
 * src/solvers/solver.cu:3-32). */
int reyn_integrate(reyn_ctx* ctx, const reyn_particles* fluid,
    const float* ax, const float* ay, const float* az, float dt);
int reyn_integrate_a_v_only(reyn_ctx* ctx, const reyn_particles* fluid,
    const float* ax, const float* ay, const float* az, float dt);

/* ---- SESPH (src/solvers/SESPH.cuh:11-70, SESPH.cu:6-104) ---------------- */

typedef struct reyn_sesph_params {
    reyn_sph_params sph;
    float k; /* stiffness; the reference hard-codes 500 (SESPH.cuh:50) */
} reyn_sesph_params;

typedef struct reyn_sesph_work {
    float* rho;
    float *ax, *ay, *az;
} reyn_sesph_work;

/* One `SESPH<K,R>::step` (SESPH.cu:67-95): densities, accelerations with
 * the normalised EOS max(0, k(rho/rho_0 - 1)), integration. */
int reyn_sesph_step(reyn_ctx* ctx, const reyn_sesph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid,
    const reyn_boundary* bdy, const reyn_sesph_work* work, float dt);

/* reyn_sesph_step that allocates nothing. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The same
 * arguments and step as reyn_sesph_step, with results equal to FP32 rounding
 * (every pass walks the grid instead of reading a neighbour list). Its scratch
 * is a static device array of 17 MiB that every GPU the library runs on holds
 * until the process ends, for at most 2^20 fluid particles (more:
 * REYN_ERR_CAPACITY). Calls that overlap on one GPU share that array and race;
 * nothing in it needs to start at zero. Per-launch profiling
 * (reyn_profile_enable) creates events, so it must be off. */
int reyn_sesph_step_noalloc(reyn_ctx* ctx, const reyn_sesph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid,
    const reyn_boundary* bdy, const reyn_sesph_work* work, float dt);

/* ---- PCISPH (src/solvers/PCISPH.cuh:8-111, PCISPH.cu:15-118) ------------ */

/* `k_c_pcisph` (PCISPH.cuh:49-77): delta = 1 / (2 V0^2 (|sum grad W|^2 +
 * sum |grad W|^2)) over a 5^3 template lattice of spacing h, rounded to
 * float. Host-only, no context needed. */
int reyn_pcisph_delta(int32_t kernel, float h_bar, float h, double* delta);

typedef struct reyn_pcisph_params {
    reyn_sph_params sph;
    double delta;      /* from reyn_pcisph_delta */
    float eta_rho_max; /* 0.001 (PCISPH.cuh:36) */
    uint32_t min_iter; /* 3 (PCISPH.cuh:38) */
    uint32_t max_iter; /* 0 = no cap, as in the reference */
    uint32_t check_every; /* iterations launched between host checks; 0 = 4 */
} reyn_pcisph_params;

typedef struct reyn_pcisph_work {
    float* rho;
    float *ax, *ay, *az;
} reyn_pcisph_work;

typedef struct reyn_solve_info {
    uint32_t iterations;
    uint32_t n_violating;   /* particles above the threshold, last iterate */
    double mean_rel_error;  /* FP64 mean error / rho_0, last iterate */
    double avg_density;     /* PCISPH: FP64 mean predicted density */
} reyn_solve_info;

/* One `PCISPH<K,R>::step` (PCISPH.cu:58-111). The loop
 * `while (l < min_iter || avg(rho) > (1 + eta) rho_0)` runs on the device,
 * with the average formed in FP64. */
int reyn_pcisph_step(reyn_ctx* ctx, const reyn_pcisph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid,
    const reyn_boundary* bdy, const reyn_pcisph_work* work, float dt,
    reyn_solve_info* info);

/* ---- IISPH (src/solvers/IISPH.cuh:8-93, IISPH.cu:10-269) ---------------- */

enum {
    /* the reference rule: stop once no particle has a predicted density error
     * above eta * rho_0 (IISPH.cu:153-161, :246) */
    REYN_IISPH_STOP_ANY = 0,
    /* stop once mean_i max(0, err_i) / rho_0 <= eta, with the mean formed in
     * FP64 on the device */
    REYN_IISPH_STOP_MEAN = 1
};

typedef struct reyn_iisph_params {
    reyn_sph_params sph;
    float omega;       /* 0.5 (IISPH.cuh:45) */
    float eta_rho_max; /* 0.005 (IISPH.cuh:41) */
    uint32_t min_iter; /* 3 (IISPH.cuh:43) */
    uint32_t max_iter; /* 0 = no cap, as in the reference */
    int32_t stop_rule; /* REYN_IISPH_STOP_* */
    int32_t warm_start; /* 1: p *= 0.5 (IISPH.cuh:48, IISPH.cu:199-215) */
    uint32_t check_every; /* iterations launched between host checks; 0 = 4 */
} reyn_iisph_params;

typedef struct reyn_iisph_work {
    float* rho;
    float* p; /* persists across steps (warm start) */
    float* a_ii;
    float* s_i;
    float *ax, *ay, *az;
} reyn_iisph_work;

/* One `IISPH<K,R>::step` (IISPH.cu:164-262): densities, non-pressure
 * accelerations, a_ii and s_i, the relaxed-Jacobi pressure loop with its
 * convergence test on the device, final pressure accelerations and
 * integration. */
int reyn_iisph_step(reyn_ctx* ctx, const reyn_iisph_params* prm,
    const reyn_particles* fluid, const reyn_grid* grid,
    const reyn_boundary* bdy, const reyn_iisph_work* work, float dt,
    reyn_solve_info* info);

/* ---- boundary, time step, bounds, reductions --------------------------- */

/* Akinci boundary masses (`calculate_boundary_masses`,
 * src/scene/sample_boundary.cu:251-297): m starts at 1 and is refined
 * `iterations` times with m_b <- rho_0 m_b / sum_j m_j W_bj using B3 with
 * support 2h. Each refinement reads the previous masses only (Jacobi), so
 * the result is deterministic. */
int reyn_boundary_masses(reyn_ctx* ctx, const reyn_boundary* bdy, float* m,
    float h, float rho_0, uint32_t iterations);

/* `cfl_time_step` (src/timestep/cfl.cu:5-37): dt = min(lambda h / max|v|,
 * sqrt(2 lambda h / |g|)); max|v| is reduced in FP64 on the device. */
int reyn_cfl_time_step(reyn_ctx* ctx, float lambda, float h,
    const reyn_particles* fluid, reyn_float3 g, float* dt, double* v_max);

/* `simple_dt_controller` (src/timestep/cfl.cu:39-52). */
int reyn_simple_dt_controller(reyn_ctx* ctx, uint32_t iter_last,
    float dt_last, float lambda, float h, const reyn_particles* fluid,
    reyn_float3 g, uint32_t iter_target, float change_factor, float* dt);

/* `Scene::hard_enforce_bounds` (src/scene/scene.cu:466-491). */
int reyn_hard_enforce_bounds(reyn_ctx* ctx, reyn_float3 bound_min,
    reyn_float3 bound_max, const reyn_particles* fluid);

/* FP64 reductions of a float array (`DeviceBuffer::min/max/sum/avg`,
 * src/buffer.cuh:213-258). op: 0 sum, 1 min, 2 max, 3 mean. */
int reyn_reduce_f32(reyn_ctx* ctx, const float* x, uint32_t n, int32_t op,
    double* out);

/* ---- lattice scenes on the device ------------------------------------------
 * A dam break: a fluid lattice of counts[0] x counts[1] x counts[2] points at
 * (i + 1/2) s from the origin, each coordinate moved by U(-jitter s, jitter s)
 * from a counter hash of (seed, point, axis); the lattice shell of `layers`
 * layers of the box [0, box] at the same spacing (the points of the
 * box lattice outside its inner index box, x fastest); optionally `layers`
 * Fibonacci-lattice spheres of radii radius - (k + 1/2) s around `center`,
 * with the fluid within radius + s/2 of the centre removed. All coordinates
 * are formed in double and rounded to float. The reference initialises fluid
 * boxes on the device with curand jitter (src/scene/scene.cu:14-54, 239-253)
 * and samples boundaries from meshes; no counterpart for the shell. */
typedef struct reyn_dam_break {
    uint32_t counts[3];
    double spacing;
    double box[3];
    double jitter;           /* fraction of the spacing */
    uint32_t seed;
    uint32_t layers;         /* wall (and sphere) layers, at most 8 */
    int32_t sphere;          /* 1: sphere obstacle */
    double center[3];
    double radius;
} reyn_dam_break;

/* The lattice size (the fluid count before the sphere is cut out) and the
 * number of boundary samples. Host only. */
int reyn_dam_break_sizes(const reyn_dam_break* d, uint32_t* n_lattice, uint32_t* n_bdy);
/* Generates into device arrays: fluid x, y, z of at least n_lattice entries
 * (n_fluid receives the count kept), boundary bx, by, bz of n_bdy. */
int reyn_dam_break_generate(reyn_ctx* ctx, const reyn_dam_break* d, float* x, float* y, float* z,
    uint32_t* n_fluid, float* bx, float* by, float* bz);

/* ---- scene set-up --------------------------------------------------------------
 * Device helpers from which the C++ `reyn::Scene` builds the reference's scenes
 * (src/scene/scene.cu:239-464). All work on device arrays. */
/* Point i of an nx x ny x nz lattice at origin + (ix h, iy h, iz h), x fastest
 * (`_init_box_kernel`, scene.cu:14-37). */
int reyn_lattice_box(reyn_ctx* ctx, reyn_float3 origin, uint32_t nx, uint32_t ny, uint32_t nz,
    float h, float* xs, float* ys, float* zs);
/* x += sigma N(0, 1) per axis (`jitter_positions`, scene.cu:225-237): normal
 * deviates from a counter hash of (seed, point, draw) by the Box-Muller
 * transform, in place of thrust's random engine. */
int reyn_jitter_normal(reyn_ctx* ctx, float* xs, float* ys, float* zs, uint32_t n, float sigma,
    uint32_t seed);
/* keep[k] = 0 when a boundary sample lies within `radius` of point k
 * (`CullPredicate`, scene.cu:55-83); radius at most the boundary grid's search
 * radius. */
int reyn_boundary_cull(reyn_ctx* ctx, const reyn_boundary* b, float radius, const float* xs,
    const float* ys, const float* zs, uint32_t n, uint8_t* keep);
/* Stable in-place compaction of nfields >= 1 float arrays of n entries to the
 * entries with keep[k] != 0 (`cull_particles_by_predicate`, scene.cu:183-206,
 * thrust::remove_if); `count` receives the number kept. Synchronises. */
int reyn_compact_f32(reyn_ctx* ctx, const uint8_t* keep, uint32_t n, float* const* fields,
    uint32_t nfields, uint32_t* count);

/* ---- triangle meshes ---------------------------------------------------------
 * `Mesh`, `load_mesh_from_obj` (src/scene/loader.h, loader.cpp:13-107): vertices
 * in double, triangles as vertex indices. OBJ: `v` and `f` lines (corners "v",
 * "v/t", "v//n", "v/t/n"; negative indices count back), polygons triangulated
 * as fans from their first corner (the default of tinyobjloader v1.0.6, which
 * the reference uses), shapes (`o`/`g`) skipped when their lower-case name
 * contains an `ignore` entry or lacks `only`. On failure `err` (may be NULL)
 * receives the reason. */
typedef struct reyn_mesh reyn_mesh;
int reyn_mesh_load_obj(const char* path, const char* const* ignore, uint32_t n_ignore,
    const char* only, reyn_mesh** out, char* err, size_t err_len);
int reyn_mesh_create(const double* x, const double* y, const double* z, uint32_t n_vertices,
    const uint32_t* faces, uint32_t n_faces, reyn_mesh** out);
int reyn_mesh_destroy(reyn_mesh* m);
/* Sizes, total area and signed volume (sum of a . (b x c) / 6,
 * `DeviceMesh::get_volume`, sample_boundary.cuh:33-60); any output may be NULL. */
int reyn_mesh_info(const reyn_mesh* m, uint32_t* n_vertices, uint32_t* n_faces, double* area,
    double* volume);
int reyn_mesh_data(const reyn_mesh* m, double* x, double* y, double* z, uint32_t* faces);
/* Boundary samples of the surface (`uniform_sample_mesh`,
 * sample_boundary.cu:179-249, 299-389): n = ceil(area / (1.5 sqrt(3) h_bdy^2))
 * with h_bdy = h / oversampling; sample i takes the area coordinate
 * (i / n + u / n) * area, the face whose inclusive area CDF is first >= it, and
 * barycentric (s, t) folded into the triangle, u, s, t from a counter hash of
 * (seed, i); in double, rounded to float. Device arrays of `capacity`. */
int reyn_mesh_sample_count(const reyn_mesh* m, float h, float oversampling, uint32_t* count);
int reyn_mesh_sample(reyn_ctx* ctx, const reyn_mesh* m, float h, float oversampling,
    uint32_t seed, float* xs, float* ys, float* zs, uint32_t capacity, uint32_t* count);
/* Relaxes samples of the mesh surface (device arrays, any order; left in the
 * cell order of their grid) `relax_sampling` (sample_boundary.cu:100-166):
 * per iteration a grid over the samples, their B3 number densities (support
 * 2 h_bdy), a shift -factor h_bdy^2 sum_j (n_i + n_j) / n_0^2 grad W_ij with n_0
 * the hexagonal-plane rest density, and the projection of every shifted sample
 * to the closest point of the mesh (through an LBVH built for the call).
 * `ratio` (may be NULL) receives max/min number density of the last
 * iteration. */
int reyn_mesh_relax(reyn_ctx* ctx, const reyn_mesh* m, float h, float oversampling,
    float* xs, float* ys, float* zs, uint32_t n, uint32_t iterations, float factor, float* ratio);
/* Moves every point (device arrays, in place) to the closest point of the
 * mesh by testing every face in double; among faces at the same distance the
 * lowest index. The brute-force counterpart of `reyn_lbvh_project` (the
 * reference tests its LBVH against such a kernel, lbvh.cu:278-311). */
int reyn_mesh_project(reyn_ctx* ctx, const reyn_mesh* m, float* xs, float* ys, float* zs,
    uint32_t n);
/* inside[k] = 1 when the generalized winding number of the mesh at point k
 * (device arrays) exceeds 1/2, summed over every face: exact for any closed
 * mesh, O(points x faces). `reyn_lbvh_inside` is the fast form. */
int reyn_mesh_inside(reyn_ctx* ctx, const reyn_mesh* m, const float* xs, const float* ys,
    const float* zs, uint32_t n, uint8_t* inside);

/* Linear BVH over the faces of a mesh, built on the GPU of `ctx` (`LBVH`,
 * src/datastructure/lbvh.cuh:614-758): face centroids' 30-bit Morton codes
 * in the mesh's box, made unique by the face index (code << 32 | face), a
 * radix sort, the binary radix tree of the sorted keys and the boxes
 * bottom-up (double, exact). It keeps its own copy of the mesh. */
typedef struct reyn_lbvh reyn_lbvh;
int reyn_lbvh_create(reyn_ctx* ctx, const reyn_mesh* m, reyn_lbvh** out);
int reyn_lbvh_destroy(reyn_lbvh* b);
/* Number of faces (leaves) and the tree height (levels from the root to the
 * deepest leaf, the leaf included). */
int reyn_lbvh_info(const reyn_lbvh* b, uint32_t* n_faces, uint32_t* height);
/* Host copies of the tree (any may be NULL): for internal node i (n_faces - 1
 * of them, the root is 0) its children left[i], right[i] as index << 1 | 1 for
 * an internal node and index << 1 for a leaf (the reference's `ChildNode`,
 * lbvh.cuh:162-193); leaf k's face; boxes as lo x, y, z, hi x, y, z, 6 doubles
 * per internal node and per leaf. */
int reyn_lbvh_nodes(reyn_ctx* ctx, const reyn_lbvh* b, uint32_t* left, uint32_t* right,
    uint32_t* leaf_face, double* internal_box, double* leaf_box);
/* `LBVH::project` (lbvh.cuh:728-742): every point (device arrays, in place)
 * to the closest point of the mesh; the same result as `reyn_mesh_project`.
 * Asynchronous on the context's stream. */
int reyn_lbvh_project(reyn_ctx* ctx, const reyn_lbvh* b, float* xs, float* ys, float* zs,
    uint32_t n);
/* The reference's `InsidePredicate` (scene.cu:88-140) through the LBVH:
 * inside[k] = 1 when point k lies behind the closest feature of the surface,
 * decided by the angle-weighted pseudo-normal of the face, edge or vertex that
 * holds the closest point (the reference reads the closest face's normal);
 * points on the surface count as outside. For closed, consistently oriented
 * meshes with shared vertices it agrees with `reyn_mesh_inside`.
 * Asynchronous on the context's stream. */
int reyn_lbvh_inside(reyn_ctx* ctx, const reyn_lbvh* b, const float* xs, const float* ys,
    const float* zs, uint32_t n, uint8_t* inside);

/* ---- scene batches ---------------------------------------------------------
 * Many independent scenes advanced together: every phase of a step is one
 * launch over the particles of all scenes (workgroups aligned to scene
 * starts), and each scene keeps its own grid, boundary, spacing, time step and
 * pressure-loop control. A scene's result is the same alone (a batch of one)
 * and in any batch. No counterpart in the reference, which runs one scene per
 * process (src/main.cu:46-78); the step it batches is the reference's
 * PCISPH or IISPH step with the main loop's grid, time-step controller and
 * bounds (main.cu:59-78, cfl.cu:39-52, scene.cu:466-491). */

enum { REYN_SOLVER_PCISPH = 1, REYN_SOLVER_IISPH = 2 };

typedef struct reyn_batch reyn_batch;

typedef struct reyn_batch_scene_info {
    uint32_t steps;          /* steps taken */
    uint32_t iterations;     /* pressure iterations of the last step */
    float dt;                /* time step of the last step */
    double time;             /* simulated time */
    double loop_stat;        /* PCISPH: mean predicted density; IISPH: mean error */
} reyn_batch_scene_info;

typedef struct reyn_batch_scene {
    uint32_t n;                  /* fluid particles */
    const float *x, *y, *z;      /* host arrays of n: initial positions */
    const float *vx, *vy, *vz;   /* host arrays of n, or NULL for rest */
    float mass;                  /* per fluid particle */
    uint32_t n_bdy;              /* boundary samples */
    const float *bx, *by, *bz;   /* host arrays of n_bdy */
    reyn_float3 bound_min;       /* the fluid domain: grid extent and the */
    reyn_float3 bound_max;       /* hard_enforce_bounds box */
    float h;                     /* particle spacing; kernel support 2h */
    const float* p;              /* host array of n: initial pressure (IISPH warm
                                    start), or NULL for zero */
    /* NULL for a scene that starts now; or the record of a scene downloaded
     * from a batch (with its state in the order downloaded and its pressure):
     * the scene continues from there exactly as if it had stayed (the time
     * step controller resumes from `dt` and `iterations`) */
    const reyn_batch_scene_info* resume;
    int32_t device;              /* 1: every array above is a device pointer */
} reyn_batch_scene;

typedef struct reyn_batch_params {
    int32_t solver;          /* REYN_SOLVER_* */
    int32_t kernel;          /* REYN_KERNEL_*, shared by all scenes */
    float nu, rho_0;
    reyn_float3 g;
    float cell_ratio;        /* fluid and boundary grids */
    /* time step: the reference main loop's controller (main.cu:59-65) */
    float dt0, cfl_lambda, change_factor;
    uint32_t iter_target;
    /* pressure loop: PCISPH uses eta, min_iter, max_iter; IISPH all */
    float eta_rho_max, omega;
    uint32_t min_iter, max_iter;
    int32_t stop_rule, warm_start;
    /* 1: after the first step, every step runs as captured HIP graphs (time
     * step, grid, reorder and loop start; chunks of loop iterations; the end),
     * with one host check of the loop control per launch */
    int32_t graph;
    /* loop iterations launched between host checks after the first min_iter
     * (0 = 4); later chunks launch only the scenes whose loop is not done */
    uint32_t check_every;
    /* REYN_CELL_ORDER_* of the fluid grids and so of each scene's state (the
     * boundary grids are XYZ) */
    int32_t cell_order;
} reyn_batch_params;

/* Uploads the scenes, builds their boundary grids and Akinci masses. The
 * scenes get the ids 0 .. count-1; `count` may be 0 (an empty batch that
 * scenes join later). */
int reyn_batch_create(reyn_ctx* ctx, const reyn_batch_params* prm,
    const reyn_batch_scene* scenes, uint32_t count, reyn_batch** out);
int reyn_batch_destroy(reyn_batch* b);
/* Scenes join and leave between steps. A joining scene starts from its given
 * state with the batch's dt0, or continues a downloaded scene (`resume`);
 * resident scenes keep their state, time step and loop history, so a scene's
 * result does not depend on which scenes shared its steps. `ids` (may be
 * NULL) receives the new scenes' ids. */
int reyn_batch_add(reyn_batch* b, const reyn_batch_scene* scenes, uint32_t count,
    uint32_t* ids);
int reyn_batch_remove(reyn_batch* b, const uint32_t* ids, uint32_t count);
/* Adds dam-break scenes generated on the device (`reyn_dam_break_generate`)
 * with particle mass rho_0 s^3 (float, (s*s)*s) and the boundary samples' box
 * as bounds; no host arrays are involved. */
int reyn_batch_add_dam_breaks(reyn_batch* b, const reyn_dam_break* scenes, uint32_t count,
    uint32_t* ids);
/* The fluid particle count of a resident scene. */
int reyn_batch_scene_size(const reyn_batch* b, uint32_t scene, uint32_t* n);
/* The ids of the resident scenes (up to `cap`) and their number. */
int reyn_batch_scenes(const reyn_batch* b, uint32_t* ids, uint32_t cap, uint32_t* count);
/* Device bytes a scene occupies while resident in a batch with `prm`: state,
 * work arrays, solver scratch, grid tables and boundary. */
int reyn_batch_scene_bytes(const reyn_batch_params* prm, const reyn_batch_scene* scene,
    uint64_t* bytes);
/* Advances every resident scene by `steps` steps (time step, grid + reorder,
 * solver step, bounds). */
int reyn_batch_step(reyn_batch* b, uint32_t steps);
int reyn_batch_info(const reyn_batch* b, uint32_t scene, reyn_batch_scene_info* out);
/* Copies a scene's state (in its current cell order) to host arrays of its n;
 * any pointer may be NULL. `scene` is the scene's id. */
int reyn_batch_download(reyn_batch* b, uint32_t scene, float* x, float* y, float* z,
    float* vx, float* vy, float* vz, float* rho, float* p);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* REYN_AMD_H_ */
