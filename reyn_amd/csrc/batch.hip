// Scene batches: many independent scenes advanced together, one launch per
// phase over the particles of all scenes.
//
// Layout: the fluid arrays of all scenes are concatenated (scene s owns
// particles [first_s, first_s + n_s)); workgroups are aligned to scene starts
// (a table gives each workgroup its scene), so a workgroup sees one scene's
// constants and forms the partial sums that scene would form alone. The fluid
// grids of all scenes are concatenated into one cell space (scene s's cells
// start at cell_off[s]), so one counting sort over all points keeps every
// scene's particles contiguous and in the order a build of that scene alone
// would give. Each scene keeps its own boundary samples and grid, spacing,
// kernel support, time step (the main loop's controller, computed on the
// device from the scene's max |v| and last iteration count) and pressure-loop
// control.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "batch.hpp"
#include "grid.hpp"

namespace reyn {
namespace {

// Per-workgroup max |v| in FP64 (a NaN velocity counts as infinite), as the
// single-scene CFL reduction forms it.
__global__ __launch_bounds__(kBlock) void k_batch_vmax(SceneArg sa, FluidDev f, double* __restrict__ part)
{
    __shared__ double lds[kBlock / kWave];
    REYN_SCENE(true, sa);
    double v = 0.0;
    if (live) {
        const double a = f.vx[t], b = f.vy[t], c = f.vz[t];
        v = sqrt(fma(a, a, fma(b, b, c * c)));
        if (v != v)
            v = INFINITY;
    }
    v = block_reduce<2>(v, lds);
    if (threadIdx.x == 0)
        part[lb_] = v;
}

struct DtParams {
    f3 g;
    float lambda, change;
    uint32_t target;
    int pcisph;
};

// One workgroup per scene: the main loop's time-step controller
// (cfl.cu:5-52, main.cu:59-65) in the same float and double operations as
// the single-scene entry points, then the PCISPH stiffness (PCISPH.cu:65).
// `bad[s]` is set when the scene's velocities are not finite.
__global__ __launch_bounds__(kBlock) void k_batch_dt(SceneDev* __restrict__ table, const double* __restrict__ part,
    const uint32_t* __restrict__ iters_last, const double* __restrict__ delta, DtParams dp,
    float* __restrict__ dt_out, uint32_t* __restrict__ bad)
{
    __shared__ double lds[kBlock / kWave];
    SceneDev& sc = table[blockIdx.x];
    double v = 0.0;
    for (uint32_t k = threadIdx.x; k < sc.nblk; k += blockDim.x)
        v = fmax(v, part[sc.blk_first + k]);
    v = block_reduce<2>(v, lds);
    if (threadIdx.x == 0) {
        const float h = sc.c.h;
        const float dt_vel = static_cast<float>(double(dp.lambda * h) / v);
        const double gn = sqrt(double(fmaf(dp.g.x, dp.g.x, fmaf(dp.g.y, dp.g.y, dp.g.z * dp.g.z))));
        const float dt_g = sqrtf(static_cast<float>(2.0 * double(dp.lambda) * double(h) / gn));
        const float cfl = fminf(dt_vel, dt_g);
        const float factor = iters_last[sc.idx] < dp.target ? 1.f + dp.change : 1.f - dp.change;
        const float dt = fminf(cfl, sc.dt * factor);
        const bool ok = (v <= 1.0e300) && (dt > 0.f);
        bad[sc.idx] = ok ? 0u : 1u;
        sc.dt = dt;
        if (dp.pcisph)
            sc.kk = ok ? static_cast<float>((delta[sc.idx] / dt) / dt) : 0.f;
        dt_out[sc.idx] = dt;
    }
}

// `hard_enforce_bounds` (scene.cu:466-491) with each scene's own box.
__global__ __launch_bounds__(kBlock) void k_batch_bounds(SceneArg sa, FluidDev f)
{
    REYN_SCENE(true, sa);
    if (!live)
        return;
    const f3 lo = sc.box_lo, hi = sc.box_hi;
    const f3 x = ld3(f.x, f.y, f.z, t);
    if (x.x <= lo.x || x.y <= lo.y || x.z <= lo.z || hi.x <= x.x || hi.y <= x.y || hi.z <= x.z) {
        const float c = static_cast<float>(-0.1);
        st3(f.vx, f.vy, f.vz, t, c * ld3(f.vx, f.vy, f.vz, t));
    }
    const f3 cl { fmaxf(fminf(x.x, hi.x), lo.x), fmaxf(fminf(x.y, hi.y), lo.y), fmaxf(fminf(x.z, hi.z), lo.z) };
    st3(f.x, f.y, f.z, t, cl);
}

// Packed boundary records {x, y, z, m} of one scene (the samples are static).
__global__ __launch_bounds__(kBlock) void k_batch_pack_bdy(BdyDev b, float4* __restrict__ B4)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < b.n)
        B4[k] = make_float4(b.x[k], b.y[k], b.z[k], b.m[k]);
}

constexpr int kFields = 8; // x y z vx vy vz m p

} // namespace
} // namespace reyn

using namespace reyn;

namespace {

// A resident scene: its boundary (own cell-ordered grid, Akinci masses and
// packed records, in one allocation), its fluid grid geometry and solver
// constants, and the state that carries over when the batch's layout is
// rebuilt (time step, stiffness, last iteration count, step record).
struct SceneRes {
    uint32_t id = 0;
    uint32_t n = 0;
    float mass = 0.f;
    reyn_grid_builder* bgrid = nullptr;
    void* bmem = nullptr;
    uint8_t* near = nullptr; // boundary cells with samples within reach
    BdyDev bd {};
    GridDev g {}; // prefix set by the layout
    size_t ncells = 0;
    SphK c {};
    float thr = 0.f;
    f3 box_lo {}, box_hi {};
    double delta = 0.0;
    float dt = 0.f, kk = 0.f;
    uint32_t iters_last = 0;
    reyn_batch_scene_info info {};
    uint32_t iters = 0;
    double stat = 0.0;
};

void free_scene(SceneRes* r)
{
    if (!r)
        return;
    reyn_grid_builder_destroy(r->bgrid);
    if (r->bmem)
        (void)hipFree(r->bmem);
    if (r->near)
        (void)hipFree(r->near);
    delete r;
}

// The concatenated device state of the resident scenes: fluid arrays (scene
// slot s owns [first[s], first[s] + n_s)), scene table, per-workgroup and
// per-scene arrays, the solver scratch and the fluid grid over the
// concatenated cells. Rebuilt whenever scenes join or leave.
struct Layout {
    uint32_t S = 0, N = 0, nblocks = 0;
    std::vector<void*> allocs;
    float* st[2][kFields] = {};
    int cur = 0;
    float *rho = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *a_ii = nullptr, *s_i = nullptr;
    SceneDev* table = nullptr;
    uint32_t* blk_scene = nullptr;
    uint32_t* cell_off = nullptr;
    double* delta = nullptr;
    uint32_t* iters_dev = nullptr;
    double* vpart = nullptr;
    float* dt_dev = nullptr;
    uint32_t* bad = nullptr;
    void* scratch = nullptr;
    uint32_t* blk_list = nullptr; // device: the workgroups of a compacted launch
    uint32_t* blk_list_host = nullptr; // pinned staging of that list
    void* pinned = nullptr; // status, flags, time steps, loop control
    reyn_grid_builder* fgrid = nullptr;
    std::vector<uint32_t> first, blk_first, nblk;
    std::vector<uint8_t> scene_done;
};

void free_layout(Layout& L)
{
    reyn_grid_builder_destroy(L.fgrid);
    for (void* p : L.allocs)
        (void)hipFree(p);
    if (L.pinned)
        (void)hipHostFree(L.pinned);
    if (L.blk_list_host)
        (void)hipHostFree(L.blk_list_host);
    L = Layout {};
}

} // namespace

struct reyn_batch {
    reyn_ctx* ctx = nullptr;
    reyn_batch_params prm {};
    std::vector<SceneRes*> scenes; // by slot
    uint32_t next_id = 0;
    Layout L;
    // captured steps, per state buffer the step starts from: the time step,
    // grid, reorder and loop start; one chunk of loop iterations; the end
    bool graphs = false;
    bool direct_next = true; // the first step after a layout change runs directly
    hipGraphExec_t g_begin[2] = {}, g_chunk[2] = {}, g_end[2] = {};
    std::vector<uint32_t> hint; // per scene: the loop iterations of its last step
};

namespace {

int lalloc(reyn_ctx* ctx, Layout& L, void** p, size_t bytes)
{
    REYN_HIP(ctx, hipMalloc(p, bytes ? bytes : 4));
    L.allocs.push_back(*p);
    return REYN_OK;
}
template <class T> int lalloc_n(reyn_ctx* ctx, Layout& L, T** p, size_t n)
{
    return lalloc(ctx, L, reinterpret_cast<void**>(p), n * sizeof(T));
}

// Copies host or device data to the device.
int upload(reyn_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    return REYN_OK;
}

void destroy_graphs(reyn_batch* b)
{
    for (int k = 0; k < 2; ++k)
        for (hipGraphExec_t* g : { &b->g_begin[k], &b->g_chunk[k], &b->g_end[k] })
            if (*g) {
                (void)hipGraphExecDestroy(*g);
                *g = nullptr;
            }
    b->graphs = false;
}

SceneArg batch_arg(const reyn_batch* b)
{
    SceneArg a {};
    a.table = b->L.table;
    a.blk_scene = b->L.blk_scene;
    return a;
}

FluidDev batch_fluid(const reyn_batch* b)
{
    float* const* s = b->L.st[b->L.cur];
    return FluidDev { s[0], s[1], s[2], s[3], s[4], s[5], s[6], b->L.N };
}

int check_scene(reyn_ctx* ctx, const reyn_batch_scene& sc, uint32_t s)
{
    if (sc.n == 0 || !sc.x || !sc.y || !sc.z)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u has no fluid positions", s);
    if (sc.n_bdy && (!sc.bx || !sc.by || !sc.bz))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u has NULL boundary arrays", s);
    if (!(sc.h > 0.f) || !(sc.mass > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u needs a positive spacing and mass", s);
    if (sc.resume && !(sc.resume->dt > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u resumes with a non-positive time step", s);
    return REYN_OK;
}

int check_params(reyn_ctx* ctx, const reyn_batch_params* prm)
{
    if (prm->solver != REYN_SOLVER_PCISPH && prm->solver != REYN_SOLVER_IISPH)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown solver %d", prm->solver);
    if (prm->kernel < 0 || prm->kernel > 3)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown kernel %d", prm->kernel);
    if (!(prm->dt0 > 0.f) || !(prm->cfl_lambda > 0.f) || !(prm->rho_0 > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: dt0, cfl_lambda and rho_0 must be positive");
    if (prm->cell_ratio < 0.f)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: negative cell ratio");
    if (prm->solver == REYN_SOLVER_IISPH && prm->stop_rule != REYN_IISPH_STOP_ANY
        && prm->stop_rule != REYN_IISPH_STOP_MEAN)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown stop rule %d", prm->stop_rule);
    return REYN_OK;
}

// Bounding box of a scene's boundary samples (host arrays).
void bdy_box(const reyn_batch_scene& sc, reyn_float3* lo_out, reyn_float3* hi_out)
{
    reyn_float3 lo { sc.bx[0], sc.by[0], sc.bz[0] }, hi = lo;
    for (size_t k = 0; k < sc.n_bdy; ++k) {
        lo = { std::min(lo.x, sc.bx[k]), std::min(lo.y, sc.by[k]), std::min(lo.z, sc.bz[k]) };
        hi = { std::max(hi.x, sc.bx[k]), std::max(hi.y, sc.by[k]), std::max(hi.z, sc.bz[k]) };
    }
    *lo_out = lo;
    *hi_out = hi;
}

// Boundary bytes per sample: x, y, z, m and the packed record.
constexpr size_t kBdyBytes = 4 * sizeof(float) + sizeof(float4);

// Uploads a scene's boundary, builds its cell-ordered grid and Akinci masses
// (as BoundarySamples), packs its records and fixes its fluid grid geometry.
int make_scene(reyn_batch* b, const reyn_batch_scene& sc, SceneRes** out)
{
    reyn_ctx* ctx = b->ctx;
    const reyn_batch_params& prm = b->prm;
    *out = nullptr;
    SceneRes* r = new (std::nothrow) SceneRes();
    if (!r)
        return REYN_ERR_OUT_OF_MEMORY;
    const float h = sc.h, rad = 2.f * h;
    int rc = REYN_OK;
    void* tmpmem = nullptr;
    do {
        r->n = sc.n;
        r->mass = sc.mass;
        if (sc.n_bdy) {
            const size_t nb = sc.n_bdy;
            rc = check_hip(ctx, hipMalloc(&r->bmem, nb * kBdyBytes), "hipMalloc boundary");
            if (rc == REYN_OK)
                rc = check_hip(ctx, hipMalloc(&tmpmem, nb * 3 * sizeof(float)), "hipMalloc boundary tmp");
            if (rc != REYN_OK)
                break;
            float* base = static_cast<float*>(r->bmem);
            float* fields[3] = { base, base + nb, base + 2 * nb };
            float* bm = base + 3 * nb;
            float4* b4 = reinterpret_cast<float4*>(base + 4 * nb);
            float* tb = static_cast<float*>(tmpmem);
            float* tmp[3] = { tb, tb + nb, tb + 2 * nb };
            const float* src[3] = { sc.bx, sc.by, sc.bz };
            for (int k = 0; k < 3 && rc == REYN_OK; ++k)
                rc = upload(ctx, fields[k], src[k], nb * sizeof(float));
            reyn_float3 lo {}, hi {};
            if (!sc.device) {
                bdy_box(sc, &lo, &hi);
            } else { // device arrays: exact FP64 min/max reductions
                float* lv[3] = { &lo.x, &lo.y, &lo.z };
                float* hv[3] = { &hi.x, &hi.y, &hi.z };
                for (int k = 0; k < 3 && rc == REYN_OK; ++k) {
                    double mn = 0.0, mx = 0.0;
                    rc = reyn_reduce_f32(ctx, fields[k], sc.n_bdy, 1, &mn);
                    if (rc == REYN_OK)
                        rc = reyn_reduce_f32(ctx, fields[k], sc.n_bdy, 2, &mx);
                    *lv[k] = static_cast<float>(mn);
                    *hv[k] = static_cast<float>(mx);
                }
            }
            if (rc == REYN_OK)
                rc = reyn_grid_builder_create(ctx, lo, hi, rad, prm.cell_ratio, &r->bgrid);
            reyn_grid g {};
            if (rc == REYN_OK)
                rc = reyn_grid_construct_and_reorder(ctx, r->bgrid, rad, fields, tmp, 3, sc.n_bdy, &g);
            if (rc != REYN_OK)
                break;
            const size_t bcells = size_t(g.nx) * size_t(g.ny) * size_t(g.nz);
            rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&r->near), bcells ? bcells : 1), "hipMalloc near");
            if (rc == REYN_OK)
                rc = reyn_grid_occupancy(ctx, &g, r->near);
            if (rc != REYN_OK)
                break;
            reyn_boundary view { fields[0], fields[1], fields[2], bm, sc.n_bdy, g, r->near };
            rc = reyn_boundary_masses(ctx, &view, bm, h, prm.rho_0, 5);
            if (rc != REYN_OK)
                break;
            r->bd = to_dev(&view);
            r->bd.p4 = b4;
            {
                ProfScope _prof(ctx, "k_pack_bdy");
                k_batch_pack_bdy<<<blocks_for(sc.n_bdy), kBlock, 0, ctx->stream>>>(r->bd, b4);
            }
            rc = check_hip(ctx, hipGetLastError(), "k_pack_bdy");
            if (rc == REYN_OK)
                rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            if (rc != REYN_OK)
                break;
        }
        reyn_grid geom {};
        rc = grid_geometry(ctx, sc.bound_min, sc.bound_max, rad, prm.cell_ratio, &geom, &r->ncells, prm.cell_order);
        if (rc != REYN_OK)
            break;
        r->g = to_dev(geom);
        const reyn_sph_params sp { prm.kernel, rad, h, prm.nu, prm.rho_0, prm.g };
        r->c = make_sphk(sp);
        r->thr = prm.solver == REYN_SOLVER_PCISPH ? (prm.eta_rho_max + 1.f) * prm.rho_0 : prm.eta_rho_max * prm.rho_0;
        r->box_lo = { sc.bound_min.x, sc.bound_min.y, sc.bound_min.z };
        r->box_hi = { sc.bound_max.x, sc.bound_max.y, sc.bound_max.z };
        r->dt = prm.dt0;
        if (sc.resume) { // the controller continues from the saved step
            r->dt = sc.resume->dt;
            r->iters_last = sc.resume->iterations;
            r->info = *sc.resume;
            r->iters = sc.resume->iterations;
            r->stat = sc.resume->loop_stat;
        }
        if (prm.solver == REYN_SOLVER_PCISPH)
            rc = reyn_pcisph_delta(prm.kernel, rad, h, &r->delta);
    } while (false);
    if (tmpmem)
        (void)hipFree(tmpmem);
    if (rc != REYN_OK) {
        free_scene(r);
        return rc;
    }
    *out = r;
    return REYN_OK;
}

// Rebuilds the layout for the scenes `next` (resident scenes, in their new
// slot order, and new ones). `old_slot[i]` is the slot of next[i] in the
// current layout, or -1 for a new scene whose fluid state comes from
// `src[i]`. The current layout stays intact until the new one is complete.
int relayout(reyn_batch* b, const std::vector<SceneRes*>& next, const std::vector<int>& old_slot,
    const std::vector<const reyn_batch_scene*>& src)
{
    reyn_ctx* ctx = b->ctx;
    Layout& O = b->L;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // carried device state of the resident scenes: time step, stiffness, last
    // iteration count
    if (O.S) {
        std::vector<SceneDev> tab(O.S);
        std::vector<uint32_t> it(O.S);
        REYN_HIP(ctx, hipMemcpy(tab.data(), O.table, sizeof(SceneDev) * O.S, hipMemcpyDeviceToHost));
        REYN_HIP(ctx, hipMemcpy(it.data(), O.iters_dev, sizeof(uint32_t) * O.S, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < next.size(); ++i)
            if (old_slot[i] >= 0) {
                next[i]->dt = tab[old_slot[i]].dt;
                next[i]->kk = tab[old_slot[i]].kk;
                next[i]->iters_last = it[old_slot[i]];
            }
    }
    const uint32_t S = static_cast<uint32_t>(next.size());
    uint64_t N = 0, nblk = 0, cells = 0;
    for (const SceneRes* r : next) {
        N += r->n;
        nblk += blocks_for(r->n);
        cells += r->ncells;
    }
    // the batch's neighbour list addresses 16-bit entries with 32-bit byte offsets
    if (N >= (uint64_t(1) << 25))
        return set_error(ctx, REYN_ERR_CAPACITY, "batch: at most 2^25 - 1 fluid particles in total (got %llu)",
            static_cast<unsigned long long>(N));
    if (cells >= (uint64_t(1) << 32) - 1)
        return set_error(ctx, REYN_ERR_CAPACITY, "batch: the scenes' grids exceed 2^32 - 1 cells");
    Layout L;
    L.S = S;
    L.N = static_cast<uint32_t>(N);
    L.nblocks = static_cast<uint32_t>(nblk);
    auto build = [&]() -> int {
        if (S == 0)
            return REYN_OK;
        for (int k = 0; k < 2; ++k)
            for (int f = 0; f < kFields; ++f)
                REYN_TRY(lalloc_n(ctx, L, &L.st[k][f], N));
        for (float** a : { &L.rho, &L.ax, &L.ay, &L.az, &L.a_ii, &L.s_i })
            REYN_TRY(lalloc_n(ctx, L, a, N));
        REYN_TRY(lalloc_n(ctx, L, &L.table, S));
        REYN_TRY(lalloc_n(ctx, L, &L.blk_scene, L.nblocks));
        REYN_TRY(lalloc_n(ctx, L, &L.cell_off, S));
        REYN_TRY(lalloc_n(ctx, L, &L.delta, S));
        REYN_TRY(lalloc_n(ctx, L, &L.iters_dev, S));
        REYN_TRY(lalloc_n(ctx, L, &L.vpart, L.nblocks));
        REYN_TRY(lalloc_n(ctx, L, &L.dt_dev, S));
        REYN_TRY(lalloc_n(ctx, L, &L.bad, S));
        REYN_TRY(lalloc(ctx, L, &L.scratch, batch_scratch_bytes(L.N, L.nblocks, S)));
        REYN_TRY(lalloc_n(ctx, L, &L.blk_list, L.nblocks));
        REYN_HIP(ctx, hipHostMalloc(&L.pinned, 256 + size_t(S) * 64, hipHostMallocDefault));
        REYN_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&L.blk_list_host), sizeof(uint32_t) * L.nblocks,
            hipHostMallocDefault));
        REYN_TRY(builder_alloc(ctx, static_cast<size_t>(cells), 0.f, &L.fgrid));
        // every scene's grid has the batch's cell ratio, so the same range
        if (S > 0)
            L.fgrid->cell_cap = walk_cell_cap(next[0]->g.range);
        // fluid state: resident scenes copied from the current state buffer,
        // new scenes uploaded
        std::vector<float> zeros;
        std::vector<SceneDev> tab(S);
        std::vector<uint32_t> cell_off(S), blk_scene, iters(S);
        std::vector<double> delta(S);
        uint64_t first = 0, coff = 0;
        for (uint32_t s = 0; s < S; ++s) {
            const SceneRes& r = *next[s];
            const size_t n = r.n, bytes = n * sizeof(float);
            if (old_slot[s] >= 0) {
                const size_t of = O.first[old_slot[s]];
                for (int f = 0; f < kFields; ++f)
                    REYN_HIP(ctx, hipMemcpyAsync(L.st[0][f] + first, O.st[O.cur][f] + of, bytes,
                        hipMemcpyDeviceToDevice, ctx->stream));
                REYN_HIP(ctx, hipMemcpyAsync(L.rho + first, O.rho + of, bytes, hipMemcpyDeviceToDevice,
                    ctx->stream));
            } else {
                const reyn_batch_scene& sc = *src[s];
                const float* in[6] = { sc.x, sc.y, sc.z, sc.vx, sc.vy, sc.vz };
                if (!sc.vx || !sc.vy || !sc.vz)
                    zeros.assign(n, 0.f);
                for (int f = 0; f < 6; ++f)
                    REYN_TRY(upload(ctx, L.st[0][f] + first, in[f] ? in[f] : zeros.data(), bytes));
                REYN_TRY(reyn_fill_f32(ctx, L.st[0][6] + first, r.mass, n));
                if (sc.p)
                    REYN_TRY(upload(ctx, L.st[0][7] + first, sc.p, bytes));
                else
                    REYN_TRY(reyn_fill_f32(ctx, L.st[0][7] + first, 0.f, n));
                REYN_TRY(reyn_fill_f32(ctx, L.rho + first, 0.f, n));
                REYN_HIP(ctx, hipStreamSynchronize(ctx->stream)); // `zeros` is reused
            }
            SceneDev& d = tab[s];
            d = SceneDev {};
            d.g = r.g;
            d.g.prefix = L.fgrid->prefix + coff;
            d.b = r.bd;
            d.c = r.c;
            d.dt = r.dt;
            d.kk = r.kk;
            d.thr = r.thr;
            d.first = static_cast<uint32_t>(first);
            d.n = r.n;
            d.blk_first = static_cast<uint32_t>(blk_scene.size());
            d.nblk = blocks_for(r.n);
            d.idx = s;
            d.box_lo = r.box_lo;
            d.box_hi = r.box_hi;
            blk_scene.insert(blk_scene.end(), d.nblk, s);
            L.first.push_back(d.first);
            L.blk_first.push_back(d.blk_first);
            L.nblk.push_back(d.nblk);
            cell_off[s] = static_cast<uint32_t>(coff);
            delta[s] = r.delta;
            iters[s] = r.iters_last;
            first += n;
            coff += r.ncells;
        }
        REYN_TRY(reyn_fill_f32(ctx, L.st[1][kFields - 1], 0.f, N));
        REYN_TRY(upload(ctx, L.table, tab.data(), sizeof(SceneDev) * S));
        REYN_TRY(upload(ctx, L.blk_scene, blk_scene.data(), sizeof(uint32_t) * blk_scene.size()));
        REYN_TRY(upload(ctx, L.cell_off, cell_off.data(), sizeof(uint32_t) * S));
        REYN_TRY(upload(ctx, L.delta, delta.data(), sizeof(double) * S));
        REYN_TRY(upload(ctx, L.iters_dev, iters.data(), sizeof(uint32_t) * S));
        REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        L.scene_done.assign(S, 0);
        return REYN_OK;
    };
    const int rc = build();
    if (rc != REYN_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        free_layout(L);
        return rc;
    }
    destroy_graphs(b);
    free_layout(O);
    b->L = L;
    b->scenes = next;
    b->direct_next = true;
    return REYN_OK;
}

// Host staging in the pinned buffer.
struct Staging {
    uint32_t* status; // 4: build status
    uint32_t* bad; // S: non-finite velocities per scene
    float* dt; // S: time step per scene
    void* ctl; // S loop-control blocks
};
Staging staging(const reyn_batch* b)
{
    char* p = static_cast<char*>(b->L.pinned);
    const size_t S = b->L.S;
    Staging st;
    st.status = reinterpret_cast<uint32_t*>(p);
    st.bad = reinterpret_cast<uint32_t*>(p + 64);
    st.dt = reinterpret_cast<float*>(p + 64 + 4 * S);
    st.ctl = p + 128 + 8 * S + (64 - (8 * S) % 64) % 64;
    return st;
}

BatchTarget target(const reyn_batch* b)
{
    const Layout& L = b->L;
    BatchTarget bt {};
    bt.f = batch_fluid(b);
    bt.a = batch_arg(b);
    bt.nblocks = L.nblocks;
    bt.nscenes = L.S;
    bt.rho = L.rho;
    bt.p = L.st[L.cur][kFields - 1];
    bt.a_ii = L.a_ii;
    bt.s_i = L.s_i;
    bt.ax = L.ax;
    bt.ay = L.ay;
    bt.az = L.az;
    bt.scratch = L.scratch;
    return bt;
}

BatchSolver solver_of(const reyn_batch* b)
{
    const reyn_batch_params& prm = b->prm;
    return BatchSolver { prm.solver, prm.kernel, prm.omega, prm.eta_rho_max, prm.min_iter, prm.max_iter, 4,
        prm.stop_rule, prm.warm_start };
}

// Time step, grid over all scenes, reorder into the other state buffer and
// the solver's begin phase with the first min_iter loop iterations. Flips
// `cur`. No host synchronisation.
int launch_begin(reyn_batch* b, uint32_t* status_host)
{
    reyn_ctx* ctx = b->ctx;
    Layout& L = b->L;
    const reyn_batch_params& prm = b->prm;
    const SceneArg sa = batch_arg(b);
    {
        ProfScope _prof(ctx, "k_batch_vmax");
        k_batch_vmax<<<L.nblocks, kBlock, 0, ctx->stream>>>(sa, batch_fluid(b), L.vpart);
    }
    REYN_LAUNCHED(ctx, "k_batch_vmax");
    const DtParams dp { f3 { prm.g.x, prm.g.y, prm.g.z }, prm.cfl_lambda, prm.change_factor, prm.iter_target,
        prm.solver == REYN_SOLVER_PCISPH ? 1 : 0 };
    {
        ProfScope _prof(ctx, "k_batch_dt");
        k_batch_dt<<<L.S, kBlock, 0, ctx->stream>>>(L.table, L.vpart, L.iters_dev, L.delta, dp, L.dt_dev, L.bad);
    }
    REYN_LAUNCHED(ctx, "k_batch_dt");
    const FluidDev f = batch_fluid(b);
    REYN_TRY(sort_batch(L.fgrid, sa, L.nblocks, L.cell_off, f.x, f.y, f.z, L.N, status_host));
    const int nf = prm.solver == REYN_SOLVER_IISPH ? kFields : kFields - 1;
    REYN_TRY(launch_gather(ctx, L.fgrid->sorted, L.N, L.st[L.cur], L.st[1 - L.cur], nf));
    L.cur = 1 - L.cur;
    return batch_phase(ctx, target(b), solver_of(b), BATCH_BEGIN, std::max(prm.min_iter, 1u));
}

int launch_chunk(reyn_batch* b, uint32_t count)
{
    return batch_phase(b->ctx, target(b), solver_of(b), BATCH_ITERATE, count);
}

// A chunk of loop iterations over the workgroups of the scenes whose loop is
// not done yet (the done scenes' workgroups would only return at once).
int launch_chunk_active(reyn_batch* b, uint32_t count)
{
    reyn_ctx* ctx = b->ctx;
    Layout& L = b->L;
    uint32_t n = 0;
    for (uint32_t s = 0; s < L.S; ++s)
        if (!L.scene_done[s])
            for (uint32_t k = 0; k < L.nblk[s]; ++k)
                L.blk_list_host[n++] = L.blk_first[s] + k;
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipMemcpyAsync(L.blk_list, L.blk_list_host, sizeof(uint32_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    BatchTarget bt = target(b);
    bt.a.blk_list = L.blk_list;
    bt.nlaunch = n;
    return batch_phase(ctx, bt, solver_of(b), BATCH_ITERATE, count);
}

int launch_end(reyn_batch* b)
{
    reyn_ctx* ctx = b->ctx;
    REYN_TRY(batch_phase(ctx, target(b), solver_of(b), BATCH_END, 0));
    {
        ProfScope _prof(ctx, "k_batch_bounds");
        k_batch_bounds<<<b->L.nblocks, kBlock, 0, ctx->stream>>>(batch_arg(b), batch_fluid(b));
    }
    REYN_LAUNCHED(ctx, "k_batch_bounds");
    return batch_save_iterations(ctx, target(b), b->L.iters_dev);
}

// Copies the loop control (and with `first` the build status, flags and time
// steps) to the host and waits; returns whether every scene's loop is done.
int fetch(reyn_batch* b, bool first, bool* done)
{
    reyn_ctx* ctx = b->ctx;
    Layout& L = b->L;
    const Staging st = staging(b);
    const uint32_t S = L.S;
    if (first) {
        REYN_HIP(ctx, hipMemcpyAsync(st.status, L.fgrid->status, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
            ctx->stream));
        REYN_HIP(ctx, hipMemcpyAsync(st.bad, L.bad, sizeof(uint32_t) * S, hipMemcpyDeviceToHost, ctx->stream));
        REYN_HIP(ctx, hipMemcpyAsync(st.dt, L.dt_dev, sizeof(float) * S, hipMemcpyDeviceToHost, ctx->stream));
    }
    REYN_HIP(ctx, hipMemcpyAsync(st.ctl, batch_loop_control(target(b)), batch_loop_control_bytes(S),
        hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (first) {
        for (uint32_t s = 0; s < S; ++s)
            if (st.bad[s])
                return set_error(ctx, REYN_ERR_NOT_FINITE, "batch: scene %u has non-finite velocities",
                    b->scenes[s]->id);
        REYN_TRY(finish_build(L.fgrid, st.status));
    }
    std::vector<uint32_t> iters(S);
    std::vector<double> stat(S);
    int all = 0;
    batch_read_control(st.ctl, S, b->prm.solver, iters.data(), stat.data(), L.scene_done.data(), &all);
    for (uint32_t s = 0; s < S; ++s) {
        b->scenes[s]->iters = iters[s];
        b->scenes[s]->stat = stat[s];
    }
    *done = all != 0;
    return REYN_OK;
}

template <class F> int capture(reyn_batch* b, hipGraphExec_t* out, F&& launches)
{
    reyn_ctx* ctx = b->ctx;
    const bool prof = ctx->prof;
    ctx->prof = false; // no event records inside a captured graph
    REYN_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    const int rc = launches();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    ctx->prof = prof;
    if (rc != REYN_OK) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    REYN_HIP(ctx, e);
    const hipError_t ei = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    REYN_HIP(ctx, ei);
    return REYN_OK;
}

uint32_t chunk_of(const reyn_batch* b) { return b->prm.check_every ? b->prm.check_every : 4u; }

// Captures the three graphs for both state buffers (after a first step run
// directly, which sized every buffer).
int build_graphs(reyn_batch* b)
{
    Layout& L = b->L;
    const int cur0 = L.cur;
    for (int k = 0; k < 2; ++k) {
        L.cur = k;
        REYN_TRY(capture(b, &b->g_begin[k], [&] { return launch_begin(b, nullptr); }));
        // launch_begin flipped cur: the chunk and end graphs of the buffer it
        // leaves the state in
        REYN_TRY(capture(b, &b->g_chunk[L.cur], [&] { return launch_chunk(b, chunk_of(b)); }));
        REYN_TRY(capture(b, &b->g_end[L.cur], [&] { return launch_end(b); }));
    }
    L.cur = cur0;
    b->graphs = true;
    return REYN_OK;
}

int step(reyn_batch* b, bool use_graph)
{
    reyn_ctx* ctx = b->ctx;
    Layout& L = b->L;
    const uint32_t chunk = chunk_of(b);
    const uint32_t min_it = std::max(b->prm.min_iter, 1u);
    const uint32_t cap = b->prm.max_iter ? b->prm.max_iter : 100000u;
    if (use_graph) {
        REYN_HIP(ctx, hipGraphLaunch(b->g_begin[L.cur], ctx->stream));
        L.cur = 1 - L.cur;
    } else {
        REYN_TRY(launch_begin(b, nullptr));
    }
    // each scene's iteration count of its last step: a scene's loop takes
    // about as many iterations as in the step before
    b->hint.resize(L.S);
    for (uint32_t s = 0; s < L.S; ++s)
        b->hint[s] = b->scenes[s]->iters;
    bool done = false;
    REYN_TRY(fetch(b, true, &done));
    uint32_t launched = min_it;
    while (!done) {
        if (launched >= cap)
            return set_error(ctx, REYN_ERR_NOT_CONVERGED, "batch: a scene did not converge in %u iterations", launched);
        const bool all_active = std::find(L.scene_done.begin(), L.scene_done.end(), 1) == L.scene_done.end();
        // the next chunk runs until the first still-iterating scene is likely
        // one iteration short of its last step's count (at least check_every
        // iterations), over the workgroups of the scenes still iterating, so
        // a step ends after few checks; iterations of a scene whose loop is
        // done return at once, so results do not change
        uint32_t soon = 0xffffffffu;
        for (uint32_t s = 0; s < L.S; ++s)
            if (!L.scene_done[s] && b->hint[s] > launched + 1)
                soon = std::min(soon, b->hint[s] - 1 - launched);
        uint32_t next = (soon == 0xffffffffu) ? chunk : std::max(chunk, soon);
        next = std::min(next, cap - launched);
        if (use_graph && all_active && next == chunk)
            REYN_HIP(ctx, hipGraphLaunch(b->g_chunk[L.cur], ctx->stream));
        else
            REYN_TRY(launch_chunk_active(b, next));
        launched += next;
        REYN_TRY(fetch(b, false, &done));
    }
    if (use_graph)
        REYN_HIP(ctx, hipGraphLaunch(b->g_end[L.cur], ctx->stream));
    else
        REYN_TRY(launch_end(b));
    const Staging st = staging(b);
    for (uint32_t s = 0; s < L.S; ++s) {
        SceneRes& r = *b->scenes[s];
        reyn_batch_scene_info& in = r.info;
        in.steps += 1;
        in.iterations = r.iters;
        in.dt = st.dt[s];
        in.time += double(st.dt[s]);
        in.loop_stat = r.stat;
    }
    return REYN_OK;
}

int slot_of(const reyn_batch* b, uint32_t id)
{
    for (size_t s = 0; s < b->scenes.size(); ++s)
        if (b->scenes[s]->id == id)
            return static_cast<int>(s);
    return -1;
}

// New scenes (their boundaries built) appended after the resident ones.
int add_scenes(reyn_batch* b, const reyn_batch_scene* scenes, uint32_t count, uint32_t* ids)
{
    reyn_ctx* ctx = b->ctx;
    for (uint32_t s = 0; s < count; ++s)
        REYN_TRY(check_scene(ctx, scenes[s], s));
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<SceneRes*> added;
    int rc = REYN_OK;
    for (uint32_t s = 0; s < count && rc == REYN_OK; ++s) {
        SceneRes* r = nullptr;
        rc = make_scene(b, scenes[s], &r);
        if (rc == REYN_OK)
            added.push_back(r);
    }
    if (rc == REYN_OK) {
        std::vector<SceneRes*> next = b->scenes;
        std::vector<int> old_slot(next.size());
        std::vector<const reyn_batch_scene*> src(next.size(), nullptr);
        for (size_t s = 0; s < next.size(); ++s)
            old_slot[s] = static_cast<int>(s);
        for (uint32_t s = 0; s < count; ++s) {
            next.push_back(added[s]);
            old_slot.push_back(-1);
            src.push_back(&scenes[s]);
        }
        rc = relayout(b, next, old_slot, src);
    }
    if (rc != REYN_OK) {
        for (SceneRes* r : added)
            free_scene(r);
        return rc;
    }
    for (uint32_t s = 0; s < count; ++s) {
        added[s]->id = b->next_id++;
        if (ids)
            ids[s] = added[s]->id;
    }
    return REYN_OK;
}

} // namespace

extern "C" {

int reyn_batch_create(reyn_ctx* ctx, const reyn_batch_params* prm, const reyn_batch_scene* scenes,
    uint32_t count, reyn_batch** out)
{
    if (!ctx || !prm || !out || (count && !scenes))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: NULL argument");
    *out = nullptr;
    REYN_TRY(check_params(ctx, prm));
    reyn_batch* b = new (std::nothrow) reyn_batch();
    if (!b)
        return REYN_ERR_OUT_OF_MEMORY;
    b->ctx = ctx;
    b->prm = *prm;
    const int rc = add_scenes(b, scenes, count, nullptr);
    if (rc != REYN_OK) {
        reyn_batch_destroy(b);
        return rc;
    }
    *out = b;
    return REYN_OK;
}

int reyn_batch_destroy(reyn_batch* b)
{
    if (!b)
        return REYN_OK;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    destroy_graphs(b);
    free_layout(b->L);
    for (SceneRes* r : b->scenes)
        free_scene(r);
    delete b;
    return REYN_OK;
}

int reyn_batch_add(reyn_batch* b, const reyn_batch_scene* scenes, uint32_t count, uint32_t* ids)
{
    if (!b || (count && !scenes))
        return set_error(b ? b->ctx : nullptr, REYN_ERR_INVALID_ARGUMENT, "batch add: NULL argument");
    if (count == 0)
        return REYN_OK;
    return add_scenes(b, scenes, count, ids);
}

int reyn_batch_remove(reyn_batch* b, const uint32_t* ids, uint32_t count)
{
    if (!b || (count && !ids))
        return set_error(b ? b->ctx : nullptr, REYN_ERR_INVALID_ARGUMENT, "batch remove: NULL argument");
    reyn_ctx* ctx = b->ctx;
    std::vector<uint8_t> drop(b->scenes.size(), 0);
    for (uint32_t k = 0; k < count; ++k) {
        const int s = slot_of(b, ids[k]);
        if (s < 0 || drop[s])
            return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch remove: scene %u is not in the batch", ids[k]);
        drop[s] = 1;
    }
    if (count == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<SceneRes*> next, gone;
    std::vector<int> old_slot;
    for (size_t s = 0; s < b->scenes.size(); ++s) {
        if (drop[s]) {
            gone.push_back(b->scenes[s]);
        } else {
            next.push_back(b->scenes[s]);
            old_slot.push_back(static_cast<int>(s));
        }
    }
    REYN_TRY(relayout(b, next, old_slot, std::vector<const reyn_batch_scene*>(next.size(), nullptr)));
    for (SceneRes* r : gone)
        free_scene(r);
    return REYN_OK;
}

int reyn_batch_add_dam_breaks(reyn_batch* b, const reyn_dam_break* d, uint32_t count, uint32_t* ids)
{
    if (!b || (count && !d))
        return set_error(b ? b->ctx : nullptr, REYN_ERR_INVALID_ARGUMENT, "batch add: NULL argument");
    reyn_ctx* ctx = b->ctx;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<void*> mem;
    std::vector<reyn_batch_scene> scenes(count);
    int rc = REYN_OK;
    for (uint32_t k = 0; k < count && rc == REYN_OK; ++k) {
        uint32_t nl = 0, nb = 0, nf = 0;
        rc = dam_break_sizes(ctx, &d[k], &nl, &nb);
        void* p = nullptr;
        if (rc == REYN_OK)
            rc = check_hip(ctx, hipMalloc(&p, sizeof(float) * 3 * (size_t(nl) + nb)), "hipMalloc scene");
        if (rc != REYN_OK)
            break;
        mem.push_back(p);
        float* f = static_cast<float*>(p);
        float *x = f, *y = f + nl, *z = f + 2 * size_t(nl), *bx = f + 3 * size_t(nl), *by = bx + nb, *bz = by + nb;
        rc = dam_break_generate(ctx, &d[k], x, y, z, &nf, bx, by, bz);
        const float s = static_cast<float>(d[k].spacing);
        reyn_batch_scene& sc = scenes[k];
        sc = reyn_batch_scene {};
        sc.n = nf;
        sc.x = x;
        sc.y = y;
        sc.z = z;
        sc.mass = b->prm.rho_0 * (s * s * s);
        sc.n_bdy = nb;
        sc.bx = nb ? bx : nullptr;
        sc.by = nb ? by : nullptr;
        sc.bz = nb ? bz : nullptr;
        sc.h = s;
        sc.device = 1;
        if (rc == REYN_OK && nb) { // bounds: the boundary samples' box
            float* lo[3] = { &sc.bound_min.x, &sc.bound_min.y, &sc.bound_min.z };
            float* hi[3] = { &sc.bound_max.x, &sc.bound_max.y, &sc.bound_max.z };
            const float* a[3] = { bx, by, bz };
            for (int c = 0; c < 3 && rc == REYN_OK; ++c) {
                double mn = 0.0, mx = 0.0;
                rc = reyn_reduce_f32(ctx, a[c], nb, 1, &mn);
                if (rc == REYN_OK)
                    rc = reyn_reduce_f32(ctx, a[c], nb, 2, &mx);
                *lo[c] = static_cast<float>(mn);
                *hi[c] = static_cast<float>(mx);
            }
        } else if (rc == REYN_OK) {
            sc.bound_min = { 0.f, 0.f, 0.f };
            sc.bound_max = { static_cast<float>(d[k].box[0]), static_cast<float>(d[k].box[1]),
                static_cast<float>(d[k].box[2]) };
        }
    }
    if (rc == REYN_OK)
        rc = count ? add_scenes(b, scenes.data(), count, ids) : REYN_OK;
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : mem)
        (void)hipFree(p);
    return rc;
}

int reyn_batch_scene_size(const reyn_batch* b, uint32_t scene, uint32_t* n)
{
    const int s = (b && n) ? slot_of(b, scene) : -1;
    if (s < 0)
        return REYN_ERR_INVALID_ARGUMENT;
    *n = b->scenes[s]->n;
    return REYN_OK;
}

int reyn_batch_scenes(const reyn_batch* b, uint32_t* ids, uint32_t cap, uint32_t* count)
{
    if (!b || !count || (cap && !ids))
        return REYN_ERR_INVALID_ARGUMENT;
    *count = static_cast<uint32_t>(b->scenes.size());
    for (uint32_t s = 0; s < cap && s < *count; ++s)
        ids[s] = b->scenes[s]->id;
    return REYN_OK;
}

int reyn_batch_scene_bytes(const reyn_batch_params* prm, const reyn_batch_scene* sc, uint64_t* bytes)
{
    if (!prm || !sc || !bytes || !(sc->h > 0.f))
        return REYN_ERR_INVALID_ARGUMENT;
    *bytes = 0;
    const float rad = 2.f * sc->h;
    reyn_grid geom {};
    size_t cells = 0;
    REYN_TRY(grid_geometry(nullptr, sc->bound_min, sc->bound_max, rad, prm->cell_ratio, &geom, &cells,
        prm->cell_order));
    uint64_t total = 0;
    // fluid: both state buffers, the work arrays, the solver scratch and the
    // grid's per-point arrays; its cells: counts and prefix
    const uint64_t n = sc->n, nb = blocks_for(sc->n);
    total += n * (2 * kFields + 6) * sizeof(float);
    total += batch_scratch_bytes(sc->n, static_cast<uint32_t>(nb), 1);
    total += n * 4 * sizeof(uint32_t) + nb * (sizeof(double) + 3 * sizeof(uint32_t));
    total += uint64_t(cells) * 2 * sizeof(uint32_t);
    if (sc->n_bdy && sc->bx && sc->by && sc->bz) {
        reyn_float3 lo, hi;
        bdy_box(*sc, &lo, &hi);
        size_t bcells = 0;
        REYN_TRY(grid_geometry(nullptr, lo, hi, rad, prm->cell_ratio, &geom, &bcells));
        total += uint64_t(sc->n_bdy) * (kBdyBytes + 4 * sizeof(uint32_t)) + uint64_t(bcells) * 2 * sizeof(uint32_t);
    }
    *bytes = total + 4096;
    return REYN_OK;
}

int reyn_batch_step(reyn_batch* b, uint32_t steps)
{
    if (!b)
        return REYN_ERR_INVALID_ARGUMENT;
    if (b->L.S == 0)
        return REYN_OK;
    REYN_HIP(b->ctx, hipSetDevice(b->ctx->device));
    for (uint32_t k = 0; k < steps; ++k) {
        // the first step after a layout change runs directly (it sizes every
        // buffer), then the captured graphs take over when asked for
        const bool use = b->prm.graph && !b->direct_next;
        if (use && !b->graphs)
            REYN_TRY(build_graphs(b));
        REYN_TRY(step(b, use));
        b->direct_next = false;
    }
    REYN_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return REYN_OK;
}

int reyn_batch_info(const reyn_batch* b, uint32_t scene, reyn_batch_scene_info* out)
{
    if (!b || !out)
        return REYN_ERR_INVALID_ARGUMENT;
    const int s = slot_of(b, scene);
    if (s < 0)
        return REYN_ERR_INVALID_ARGUMENT;
    *out = b->scenes[s]->info;
    return REYN_OK;
}

int reyn_batch_download(reyn_batch* b, uint32_t scene, float* x, float* y, float* z, float* vx, float* vy,
    float* vz, float* rho, float* p)
{
    const int s = b ? slot_of(b, scene) : -1;
    if (s < 0)
        return set_error(b ? b->ctx : nullptr, REYN_ERR_INVALID_ARGUMENT, "batch download: bad scene");
    reyn_ctx* ctx = b->ctx;
    const Layout& L = b->L;
    const size_t off = L.first[s], bytes = size_t(b->scenes[s]->n) * sizeof(float);
    float* const* st = L.st[L.cur];
    float* dst[8] = { x, y, z, vx, vy, vz, rho, p };
    const float* src[8] = { st[0], st[1], st[2], st[3], st[4], st[5], L.rho, st[kFields - 1] };
    for (int f = 0; f < 8; ++f)
        if (dst[f])
            REYN_HIP(ctx, hipMemcpyAsync(dst[f], src[f] + off, bytes, hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return REYN_OK;
}

} // extern "C"
