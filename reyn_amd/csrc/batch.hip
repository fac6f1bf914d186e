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

struct reyn_batch {
    reyn_ctx* ctx = nullptr;
    reyn_batch_params prm {};
    uint32_t S = 0, N = 0, NB = 0, nblocks = 0;
    std::vector<uint32_t> first, count;
    std::vector<reyn_grid_builder*> bgrid;
    reyn_grid_builder* fgrid = nullptr;
    std::vector<SceneDev> table_host;
    std::vector<reyn_batch_scene_info> info;
    std::vector<uint32_t> iters;
    std::vector<double> stat;
    std::vector<uint8_t> scene_done;
    std::vector<uint32_t> blk_first, nblk; // workgroups of every scene
    uint32_t* blk_list = nullptr; // device: the workgroups of a compacted launch
    uint32_t* blk_list_host = nullptr; // pinned staging of that list
    std::vector<void*> allocs;
    float* st[2][kFields] = {};
    int cur = 0;
    float *rho = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr, *a_ii = nullptr, *s_i = nullptr;
    float *bm = nullptr;
    float4* b4 = nullptr; // packed boundary records of all scenes
    SceneDev* table = nullptr;
    uint32_t* blk_scene = nullptr;
    uint32_t* cell_off = nullptr;
    double* delta = nullptr;
    uint32_t* iters_dev = nullptr;
    double* vpart = nullptr;
    float* dt_dev = nullptr;
    uint32_t* bad = nullptr;
    void* scratch = nullptr; // the solver scratch of the batch
    void* pinned = nullptr; // status, flags, time steps, loop control
    // captured steps, per state buffer the step starts from: the time step,
    // grid, reorder and loop start; one chunk of loop iterations; the end
    bool graphs = false;
    hipGraphExec_t g_begin[2] = {}, g_chunk[2] = {}, g_end[2] = {};
};

namespace {

int balloc(reyn_batch* b, void** p, size_t bytes)
{
    REYN_HIP(b->ctx, hipMalloc(p, bytes ? bytes : 4));
    b->allocs.push_back(*p);
    return REYN_OK;
}
template <class T> int balloc_n(reyn_batch* b, T** p, size_t n)
{
    return balloc(b, reinterpret_cast<void**>(p), n * sizeof(T));
}

int upload(reyn_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return REYN_OK;
}

SceneArg batch_arg(const reyn_batch* b)
{
    SceneArg a {};
    a.table = b->table;
    a.blk_scene = b->blk_scene;
    return a;
}

FluidDev batch_fluid(const reyn_batch* b)
{
    float* const* s = b->st[b->cur];
    return FluidDev { s[0], s[1], s[2], s[3], s[4], s[5], s[6], b->N };
}

int create(reyn_batch* b, const reyn_batch_params* prm, const reyn_batch_scene* scenes, uint32_t count)
{
    reyn_ctx* ctx = b->ctx;
    b->prm = *prm;
    b->S = count;
    const float r_ratio = prm->cell_ratio;
    // ranges and workgroups
    uint64_t N = 0, NB = 0, nblk = 0;
    for (uint32_t s = 0; s < count; ++s) {
        const reyn_batch_scene& sc = scenes[s];
        if (sc.n == 0 || !sc.x || !sc.y || !sc.z)
            return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u has no fluid positions", s);
        if (sc.n_bdy && (!sc.bx || !sc.by || !sc.bz))
            return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u has NULL boundary arrays", s);
        if (!(sc.h > 0.f) || !(sc.mass > 0.f))
            return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: scene %u needs a positive spacing and mass", s);
        b->first.push_back(static_cast<uint32_t>(N));
        b->count.push_back(sc.n);
        N += sc.n;
        NB += sc.n_bdy;
        nblk += blocks_for(sc.n);
    }
    if (N >= (uint64_t(1) << 24))
        return set_error(ctx, REYN_ERR_CAPACITY, "batch: at most 2^24 - 1 fluid particles in total (got %llu)",
            static_cast<unsigned long long>(N));
    b->N = static_cast<uint32_t>(N);
    b->NB = static_cast<uint32_t>(NB);
    b->nblocks = static_cast<uint32_t>(nblk);
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    // device arrays
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < kFields; ++f)
            REYN_TRY(balloc_n(b, &b->st[k][f], N));
    for (float** a : { &b->rho, &b->ax, &b->ay, &b->az, &b->a_ii, &b->s_i })
        REYN_TRY(balloc_n(b, a, N));
    float *bx[2], *by[2], *bz[2];
    for (int k = 0; k < 2; ++k) {
        REYN_TRY(balloc_n(b, &bx[k], NB));
        REYN_TRY(balloc_n(b, &by[k], NB));
        REYN_TRY(balloc_n(b, &bz[k], NB));
    }
    REYN_TRY(balloc_n(b, &b->bm, NB));
    REYN_TRY(balloc_n(b, &b->b4, NB));
    REYN_TRY(balloc_n(b, &b->table, count));
    REYN_TRY(balloc_n(b, &b->blk_scene, b->nblocks));
    REYN_TRY(balloc_n(b, &b->cell_off, count));
    REYN_TRY(balloc_n(b, &b->delta, count));
    REYN_TRY(balloc_n(b, &b->iters_dev, count));
    REYN_TRY(balloc_n(b, &b->vpart, b->nblocks));
    REYN_TRY(balloc_n(b, &b->dt_dev, count));
    REYN_TRY(balloc_n(b, &b->bad, count));
    REYN_HIP(ctx, hipHostMalloc(&b->pinned, 256 + size_t(count) * 64, hipHostMallocDefault));
    REYN_TRY(balloc(b, &b->scratch, batch_scratch_bytes(b->N, b->nblocks, count)));
    REYN_TRY(balloc_n(b, &b->blk_list, b->nblocks));
    REYN_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&b->blk_list_host), sizeof(uint32_t) * b->nblocks,
        hipHostMallocDefault));
    // fluid state
    std::vector<float> zeros;
    uint64_t cells = 0, boff = 0;
    std::vector<uint32_t> cell_off(count), blk_scene;
    std::vector<double> delta(count, 0.0);
    b->table_host.resize(count);
    for (uint32_t s = 0; s < count; ++s) {
        const reyn_batch_scene& sc = scenes[s];
        const size_t off = b->first[s], n = sc.n, bytes = n * sizeof(float);
        const float* src[6] = { sc.x, sc.y, sc.z, sc.vx, sc.vy, sc.vz };
        if (!sc.vx || !sc.vy || !sc.vz)
            zeros.assign(n, 0.f);
        for (int f = 0; f < 6; ++f)
            REYN_TRY(upload(ctx, b->st[0][f] + off, src[f] ? src[f] : zeros.data(), bytes));
        REYN_TRY(reyn_fill_f32(ctx, b->st[0][6] + off, sc.mass, n));
        REYN_TRY(reyn_fill_f32(ctx, b->st[0][7] + off, 0.f, n));
        REYN_TRY(reyn_fill_f32(ctx, b->st[1][7] + off, 0.f, n));
        REYN_HIP(ctx, hipStreamSynchronize(ctx->stream)); // `zeros` is reused
        const float h = sc.h, r = 2.f * h;
        // boundary: its own cell-ordered grid and Akinci masses, as BoundarySamples
        BdyDev bd {};
        if (sc.n_bdy) {
            const size_t nb = sc.n_bdy, bbytes = nb * sizeof(float);
            REYN_TRY(upload(ctx, bx[0] + boff, sc.bx, bbytes));
            REYN_TRY(upload(ctx, by[0] + boff, sc.by, bbytes));
            REYN_TRY(upload(ctx, bz[0] + boff, sc.bz, bbytes));
            reyn_float3 lo { sc.bx[0], sc.by[0], sc.bz[0] }, hi = lo;
            for (size_t k = 0; k < nb; ++k) {
                lo = { std::min(lo.x, sc.bx[k]), std::min(lo.y, sc.by[k]), std::min(lo.z, sc.bz[k]) };
                hi = { std::max(hi.x, sc.bx[k]), std::max(hi.y, sc.by[k]), std::max(hi.z, sc.bz[k]) };
            }
            reyn_grid_builder* gb = nullptr;
            REYN_TRY(reyn_grid_builder_create(ctx, lo, hi, r, r_ratio, &gb));
            b->bgrid.push_back(gb);
            float* fields[3] = { bx[0] + boff, by[0] + boff, bz[0] + boff };
            float* tmp[3] = { bx[1] + boff, by[1] + boff, bz[1] + boff };
            reyn_grid g {};
            REYN_TRY(reyn_grid_construct_and_reorder(ctx, gb, r, fields, tmp, 3, sc.n_bdy, &g));
            reyn_boundary view { fields[0], fields[1], fields[2], b->bm + boff, sc.n_bdy, g };
            REYN_TRY(reyn_boundary_masses(ctx, &view, b->bm + boff, h, prm->rho_0, 5));
            bd = to_dev(&view);
            bd.p4 = b->b4 + boff;
            {
                ProfScope _prof(ctx, "k_pack_bdy");
                k_batch_pack_bdy<<<blocks_for(sc.n_bdy), kBlock, 0, ctx->stream>>>(bd, b->b4 + boff);
            }
            REYN_LAUNCHED(ctx, "k_pack_bdy");
            boff += nb;
        }
        // fluid grid geometry in the concatenated cell space
        reyn_grid geom {};
        size_t nc = 0;
        REYN_TRY(grid_geometry(ctx, sc.bound_min, sc.bound_max, r, r_ratio, &geom, &nc));
        if (cells + nc >= (uint64_t(1) << 32) - 1)
            return set_error(ctx, REYN_ERR_CAPACITY, "batch: the scenes' grids exceed 2^32 - 1 cells");
        cell_off[s] = static_cast<uint32_t>(cells);
        cells += nc;
        reyn_sph_params sp { prm->kernel, 2.f * h, h, prm->nu, prm->rho_0, prm->g };
        SceneDev& d = b->table_host[s];
        d = SceneDev {};
        d.g = to_dev(geom);
        d.b = bd;
        d.c = make_sphk(sp);
        d.dt = prm->dt0;
        d.thr = prm->solver == REYN_SOLVER_PCISPH ? (prm->eta_rho_max + 1.f) * prm->rho_0 : prm->eta_rho_max * prm->rho_0;
        d.first = b->first[s];
        d.n = sc.n;
        d.blk_first = static_cast<uint32_t>(blk_scene.size());
        d.nblk = blocks_for(sc.n);
        d.idx = s;
        d.box_lo = { sc.bound_min.x, sc.bound_min.y, sc.bound_min.z };
        d.box_hi = { sc.bound_max.x, sc.bound_max.y, sc.bound_max.z };
        blk_scene.insert(blk_scene.end(), d.nblk, s);
        b->blk_first.push_back(d.blk_first);
        b->nblk.push_back(d.nblk);
        if (prm->solver == REYN_SOLVER_PCISPH)
            REYN_TRY(reyn_pcisph_delta(prm->kernel, 2.f * h, h, &delta[s]));
    }
    REYN_TRY(builder_alloc(ctx, static_cast<size_t>(cells), 0.f, &b->fgrid));
    for (uint32_t s = 0; s < count; ++s)
        b->table_host[s].g.prefix = b->fgrid->prefix + cell_off[s];
    REYN_TRY(upload(ctx, b->table, b->table_host.data(), sizeof(SceneDev) * count));
    REYN_TRY(upload(ctx, b->blk_scene, blk_scene.data(), sizeof(uint32_t) * blk_scene.size()));
    REYN_TRY(upload(ctx, b->cell_off, cell_off.data(), sizeof(uint32_t) * count));
    REYN_TRY(upload(ctx, b->delta, delta.data(), sizeof(double) * count));
    REYN_TRY(reyn_fill_u32(ctx, b->iters_dev, 0u, count));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->info.assign(count, reyn_batch_scene_info {});
    b->iters.assign(count, 0u);
    b->stat.assign(count, 0.0);
    b->scene_done.assign(count, 0);
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
    char* p = static_cast<char*>(b->pinned);
    Staging st;
    st.status = reinterpret_cast<uint32_t*>(p);
    st.bad = reinterpret_cast<uint32_t*>(p + 64);
    st.dt = reinterpret_cast<float*>(p + 64 + 4 * size_t(b->S));
    st.ctl = p + 128 + 8 * size_t(b->S) + (64 - (8 * size_t(b->S)) % 64) % 64;
    return st;
}

BatchTarget target(const reyn_batch* b)
{
    BatchTarget bt {};
    bt.f = batch_fluid(b);
    bt.a = batch_arg(b);
    bt.nblocks = b->nblocks;
    bt.nscenes = b->S;
    bt.rho = b->rho;
    bt.p = b->st[b->cur][kFields - 1];
    bt.a_ii = b->a_ii;
    bt.s_i = b->s_i;
    bt.ax = b->ax;
    bt.ay = b->ay;
    bt.az = b->az;
    bt.scratch = b->scratch;
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
    const reyn_batch_params& prm = b->prm;
    const SceneArg sa = batch_arg(b);
    {
        ProfScope _prof(ctx, "k_batch_vmax");
        k_batch_vmax<<<b->nblocks, kBlock, 0, ctx->stream>>>(sa, batch_fluid(b), b->vpart);
    }
    REYN_LAUNCHED(ctx, "k_batch_vmax");
    const DtParams dp { f3 { prm.g.x, prm.g.y, prm.g.z }, prm.cfl_lambda, prm.change_factor, prm.iter_target,
        prm.solver == REYN_SOLVER_PCISPH ? 1 : 0 };
    {
        ProfScope _prof(ctx, "k_batch_dt");
        k_batch_dt<<<b->S, kBlock, 0, ctx->stream>>>(b->table, b->vpart, b->iters_dev, b->delta, dp, b->dt_dev,
            b->bad);
    }
    REYN_LAUNCHED(ctx, "k_batch_dt");
    const FluidDev f = batch_fluid(b);
    REYN_TRY(sort_batch(b->fgrid, sa, b->nblocks, b->cell_off, f.x, f.y, f.z, b->N, status_host));
    const int nf = prm.solver == REYN_SOLVER_IISPH ? kFields : kFields - 1;
    REYN_TRY(launch_gather(ctx, b->fgrid->sorted, b->N, b->st[b->cur], b->st[1 - b->cur], nf));
    b->cur = 1 - b->cur;
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
    uint32_t n = 0;
    for (uint32_t s = 0; s < b->S; ++s)
        if (!b->scene_done[s])
            for (uint32_t k = 0; k < b->nblk[s]; ++k)
                b->blk_list_host[n++] = b->blk_first[s] + k;
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipMemcpyAsync(b->blk_list, b->blk_list_host, sizeof(uint32_t) * n, hipMemcpyHostToDevice,
        ctx->stream));
    BatchTarget bt = target(b);
    bt.a.blk_list = b->blk_list;
    bt.nlaunch = n;
    return batch_phase(ctx, bt, solver_of(b), BATCH_ITERATE, count);
}

int launch_end(reyn_batch* b)
{
    reyn_ctx* ctx = b->ctx;
    REYN_TRY(batch_phase(ctx, target(b), solver_of(b), BATCH_END, 0));
    {
        ProfScope _prof(ctx, "k_batch_bounds");
        k_batch_bounds<<<b->nblocks, kBlock, 0, ctx->stream>>>(batch_arg(b), batch_fluid(b));
    }
    REYN_LAUNCHED(ctx, "k_batch_bounds");
    return batch_save_iterations(ctx, target(b), b->iters_dev);
}

// Copies the loop control (and with `first` the build status, flags and time
// steps) to the host and waits; returns whether every scene's loop is done.
int fetch(reyn_batch* b, bool first, bool* done)
{
    reyn_ctx* ctx = b->ctx;
    const Staging st = staging(b);
    const uint32_t S = b->S;
    if (first) {
        REYN_HIP(ctx, hipMemcpyAsync(st.status, b->fgrid->status, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
            ctx->stream));
        REYN_HIP(ctx, hipMemcpyAsync(st.bad, b->bad, sizeof(uint32_t) * S, hipMemcpyDeviceToHost, ctx->stream));
        REYN_HIP(ctx, hipMemcpyAsync(st.dt, b->dt_dev, sizeof(float) * S, hipMemcpyDeviceToHost, ctx->stream));
    }
    REYN_HIP(ctx, hipMemcpyAsync(st.ctl, batch_loop_control(target(b)), batch_loop_control_bytes(S),
        hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (first) {
        for (uint32_t s = 0; s < S; ++s)
            if (st.bad[s])
                return set_error(ctx, REYN_ERR_NOT_FINITE, "batch: scene %u has non-finite velocities", s);
        REYN_TRY(finish_build(b->fgrid, st.status));
    }
    int all = 0;
    batch_read_control(st.ctl, S, b->prm.solver, b->iters.data(), b->stat.data(), b->scene_done.data(), &all);
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
    const int cur0 = b->cur;
    for (int k = 0; k < 2; ++k) {
        b->cur = k;
        REYN_TRY(capture(b, &b->g_begin[k], [&] { return launch_begin(b, nullptr); }));
        // launch_begin flipped cur: the chunk and end graphs of the buffer it
        // leaves the state in
        REYN_TRY(capture(b, &b->g_chunk[b->cur], [&] { return launch_chunk(b, chunk_of(b)); }));
        REYN_TRY(capture(b, &b->g_end[b->cur], [&] { return launch_end(b); }));
    }
    b->cur = cur0;
    b->graphs = true;
    return REYN_OK;
}

int step(reyn_batch* b, bool use_graph)
{
    reyn_ctx* ctx = b->ctx;
    const uint32_t chunk = chunk_of(b);
    const uint32_t min_it = std::max(b->prm.min_iter, 1u);
    const uint32_t cap = b->prm.max_iter ? b->prm.max_iter : 100000u;
    if (use_graph) {
        REYN_HIP(ctx, hipGraphLaunch(b->g_begin[b->cur], ctx->stream));
        b->cur = 1 - b->cur;
    } else {
        REYN_TRY(launch_begin(b, nullptr));
    }
    bool done = false;
    REYN_TRY(fetch(b, true, &done));
    uint32_t launched = min_it;
    while (!done) {
        if (launched >= cap)
            return set_error(ctx, REYN_ERR_NOT_CONVERGED, "batch: a scene did not converge in %u iterations", launched);
        const bool all_active = std::find(b->scene_done.begin(), b->scene_done.end(), 1) == b->scene_done.end();
        if (use_graph && all_active)
            REYN_HIP(ctx, hipGraphLaunch(b->g_chunk[b->cur], ctx->stream));
        else
            REYN_TRY(launch_chunk_active(b, chunk));
        launched += chunk;
        REYN_TRY(fetch(b, false, &done));
    }
    if (use_graph)
        REYN_HIP(ctx, hipGraphLaunch(b->g_end[b->cur], ctx->stream));
    else
        REYN_TRY(launch_end(b));
    const Staging st = staging(b);
    for (uint32_t s = 0; s < b->S; ++s) {
        reyn_batch_scene_info& in = b->info[s];
        in.steps += 1;
        in.iterations = b->iters[s];
        in.dt = st.dt[s];
        in.time += double(st.dt[s]);
        in.loop_stat = b->stat[s];
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
    if (count == 0)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: no scenes");
    if (prm->solver != REYN_SOLVER_PCISPH && prm->solver != REYN_SOLVER_IISPH)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown solver %d", prm->solver);
    if (prm->kernel < 0 || prm->kernel > 3)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown kernel %d", prm->kernel);
    if (!(prm->dt0 > 0.f) || !(prm->cfl_lambda > 0.f) || !(prm->rho_0 > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: dt0, cfl_lambda and rho_0 must be positive");
    if (prm->solver == REYN_SOLVER_IISPH && prm->stop_rule != REYN_IISPH_STOP_ANY
        && prm->stop_rule != REYN_IISPH_STOP_MEAN)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown stop rule %d", prm->stop_rule);
    reyn_batch* b = new (std::nothrow) reyn_batch();
    if (!b)
        return REYN_ERR_OUT_OF_MEMORY;
    b->ctx = ctx;
    const int rc = create(b, prm, scenes, count);
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
    for (int k = 0; k < 2; ++k)
        for (hipGraphExec_t g : { b->g_begin[k], b->g_chunk[k], b->g_end[k] })
            if (g)
                (void)hipGraphExecDestroy(g);
    for (reyn_grid_builder* g : b->bgrid)
        reyn_grid_builder_destroy(g);
    reyn_grid_builder_destroy(b->fgrid);
    for (void* p : b->allocs)
        (void)hipFree(p);
    if (b->pinned)
        (void)hipHostFree(b->pinned);
    if (b->blk_list_host)
        (void)hipHostFree(b->blk_list_host);
    delete b;
    return REYN_OK;
}

int reyn_batch_step(reyn_batch* b, uint32_t steps)
{
    if (!b)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_HIP(b->ctx, hipSetDevice(b->ctx->device));
    for (uint32_t k = 0; k < steps; ++k) {
        // the first step runs directly (it sizes every buffer), then the
        // captured graphs take over when asked for
        const bool use = b->prm.graph && b->info[0].steps > 0;
        if (use && !b->graphs)
            REYN_TRY(build_graphs(b));
        REYN_TRY(step(b, use));
    }
    REYN_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return REYN_OK;
}

int reyn_batch_info(const reyn_batch* b, uint32_t scene, reyn_batch_scene_info* out)
{
    if (!b || !out || scene >= b->S)
        return REYN_ERR_INVALID_ARGUMENT;
    *out = b->info[scene];
    return REYN_OK;
}

int reyn_batch_download(reyn_batch* b, uint32_t scene, float* x, float* y, float* z, float* vx, float* vy,
    float* vz, float* rho, float* p)
{
    if (!b || scene >= b->S)
        return set_error(b ? b->ctx : nullptr, REYN_ERR_INVALID_ARGUMENT, "batch download: bad scene");
    reyn_ctx* ctx = b->ctx;
    const size_t off = b->first[scene], bytes = size_t(b->count[scene]) * sizeof(float);
    float* const* s = b->st[b->cur];
    float* dst[8] = { x, y, z, vx, vy, vz, rho, p };
    const float* src[8] = { s[0], s[1], s[2], s[3], s[4], s[5], b->rho, s[kFields - 1] };
    for (int f = 0; f < 8; ++f)
        if (dst[f])
            REYN_HIP(ctx, hipMemcpyAsync(dst[f], src[f] + off, bytes, hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return REYN_OK;
}

} // extern "C"
