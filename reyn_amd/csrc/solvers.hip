// SPH phases and the three pressure solvers.
//
// Every neighbour phase is one launch with one thread per fluid particle over
// the cell-ordered arrays (`walk` in sph.hpp). Boundary contributions that
// depend on positions only — sum_k m_k W_ik and sum_k m_k grad W_ik — are
// formed once per step in the density launch and reused by every later phase
// and every solver iteration, so the iteration loops walk the fluid grid only.
//
// Solver loops keep their convergence test on the device: each iteration ends
// with a one-workgroup control launch that reduces per-block partials in a
// fixed order (FP64), updates the iteration count and sets `done`. Every loop
// launch returns at once when `done` is set, so the host enqueues iterations
// in chunks and synchronises once per chunk instead of once per iteration.
#include <algorithm>
#include <cmath>
#include <mutex>

#include <vector>

#include "batch.hpp"

namespace reyn {
namespace {

struct SolveCtl {
    uint32_t iter;
    uint32_t done;
    uint32_t n_violating;
    uint32_t mass_mixed; // 1: the scene's particle masses are not all equal
    double mean_rel_error;
    double avg_density;
};

// Per-step scratch carved out of the context scratch.
struct Scratch {
    float *gbx, *gby, *gbz; // sum_k m_k grad W_ik
    float* rb; // sum_k m_k W_ik
    double* part_d; // per-block FP64 partials
    uint32_t* part_u; // per-block counts
    SolveCtl* ctl;
    NbrDev L; // the step's neighbour list (cnt == NULL when not requested)
    // packed records (NULL when the step has none):
    float4* P4; // {x, y, z, m}: walked and summed by the density pass
    float4* X; // interleaved records of the later loops, see `aux4`
    float4* B4; // packed boundary records (NULL: the boundary is walked in SoA)
};

// Neighbours kept per particle; a fluid particle at rest density has about
// (4/3) pi 2^3 = 33.5 neighbours within the support 2h.
constexpr uint32_t kListCap = 64;
// Minimum waves per SIMD each neighbour kernel is compiled for (the register
// budget follows: 512 / waves VGPRs per lane).
#ifndef REYN_OCC_DENSITY
#define REYN_OCC_DENSITY 8
#endif
#ifndef REYN_OCC_LOOP
#define REYN_OCC_LOOP 4
#endif
#ifndef REYN_OCC_PACC
#define REYN_OCC_PACC 6
#endif
#define REYN_OCC(w) __attribute__((amdgpu_waves_per_eu(w)))

// LDS staging of the list rows in the density pass: entries per lane (the
// spread of neighbour counts inside one wavefront reaches about 24 at a free
// surface), and the fewest rows worth a flush.
#ifndef REYN_LIST_RING
#define REYN_LIST_RING 16
#endif
constexpr uint32_t kRing = REYN_LIST_RING, kFlush = 4;
// records of the density pass's strip staging, per wavefront
#ifndef REYN_STAGE
#define REYN_STAGE 96
#endif
constexpr uint32_t kStage = REYN_STAGE;

size_t carve_scratch(char* p, uint32_t n, Scratch* s, bool lists, uint32_t nblocks, uint32_t nscenes,
    uint32_t nbdy)
{
    const size_t nb = nblocks ? nblocks : blocks_for(n);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t f = al(size_t(n) * 4);
    const size_t head = 4 * f + al(nb * 8) + al(nb * 4) + al(sizeof(SolveCtl) * nscenes);
    // records and list entries are addressed with 32-bit byte offsets
    const bool records = lists && size_t(n) * 32 < (size_t(1) << 32);
    // 16-bit entries: rows of word pairs addressed as 32-bit words
    const bool keep_list = records && size_t(n) * kListCap * 2 < (size_t(1) << 32);
    const size_t rec_bytes = records ? al(size_t(n) * 16) + al(size_t(n) * 32) : 0;
    const size_t list_bytes = keep_list ? f + al(size_t(n) * kListCap * 2) : 0;
    const size_t bdy_bytes = (records && nbdy) ? al(size_t(nbdy) * 16) : 0;
    const size_t total = head + rec_bytes + list_bytes + bdy_bytes;
    if (!p)
        return total;
    s->gbx = reinterpret_cast<float*>(p);
    s->gby = reinterpret_cast<float*>(p + f);
    s->gbz = reinterpret_cast<float*>(p + 2 * f);
    s->rb = reinterpret_cast<float*>(p + 3 * f);
    s->part_d = reinterpret_cast<double*>(p + 4 * f);
    s->part_u = reinterpret_cast<uint32_t*>(p + 4 * f + al(nb * 8));
    s->ctl = reinterpret_cast<SolveCtl*>(p + 4 * f + al(nb * 8) + al(nb * 4));
    s->L = NbrDev { nullptr, nullptr, n, kListCap };
    s->P4 = s->X = nullptr;
    if (records) {
        s->P4 = reinterpret_cast<float4*>(p + head);
        s->X = reinterpret_cast<float4*>(p + head + al(size_t(n) * 16));
    }
    if (keep_list) {
        s->L.cnt = reinterpret_cast<uint32_t*>(p + head + rec_bytes);
        s->L.idx = reinterpret_cast<uint32_t*>(p + head + rec_bytes + f);
    }
    s->B4 = bdy_bytes ? reinterpret_cast<float4*>(p + head + rec_bytes + list_bytes) : nullptr;
    return total;
}

// lists: the solver-step scratch (packed records, and the neighbour list
// while its entries fit 32-bit byte offsets); otherwise only the boundary
// sums and the loop control. `nblocks` workgroups (0: one scene of n),
// `nscenes` loop-control blocks, `nbdy` boundary samples to pack. Carved out
// of the context scratch.
int get_scratch(reyn_ctx* ctx, uint32_t n, Scratch* s, bool lists = false, uint32_t nblocks = 0,
    uint32_t nscenes = 1, uint32_t nbdy = 0)
{
    char* p = nullptr;
    REYN_TRY(ctx_scratch(ctx, carve_scratch(nullptr, n, s, lists, nblocks, nscenes, nbdy), reinterpret_cast<void**>(&p)));
    carve_scratch(p, n, s, lists, nblocks, nscenes, nbdy);
    // the call still succeeds, through the walk form in every pass; the note
    // tells a caller why it is slower
    if (lists && !s->P4)
        set_error(ctx, REYN_OK, "note: %u particles exceed the packed records' 32-bit offsets (2^27 - 1 at most); "
            "every pass walks the grid over the SoA arrays", n);
    else if (lists && !s->L.cnt)
        set_error(ctx, REYN_OK, "note: %u particles exceed the neighbour list's 32-bit offsets (2^25 - 1 at most); "
            "every pass after the density pass walks the grid", n);
    return REYN_OK;
}

__device__ __forceinline__ bool loop_done(const SolveCtl* ctl, const SceneDev& sc)
{
    return ctl && ctl[sc.idx].done;
}

// Interleaved records X of the loops after the density pass: X[2j] holds
// {x, y, z, m} of particle j and X[2j+1] an auxiliary slot {u, w} whose
// meaning follows the phase:
//   k_pack                      u = v,  w = 0
//   k_density                   w = rho
//   k_integrate (velocity only) u = the updated velocity
//   k_iisph_pacc (loop)         u = the pressure acceleration
// A neighbour's record and auxiliary slot share one 32-byte block, so the two
// gathers of a neighbour touch one cache line. A phase writes only the
// component that no concurrent thread of the same phase reads.
//
// The dense records P4 = {x, y, z, m} serve the density pass; after it, w of
// P4 carries the pressure (IISPH: k_iisph_aii_s, k_iisph_jacobi) or the
// predicted density (PCISPH: k_predict_rho), so the pressure-acceleration
// passes read one 16-byte record per neighbour.
__device__ __forceinline__ float4 aux4(const float4* __restrict__ X, uint32_t j) { return gat(X, 2 * j + 1); }
__device__ __forceinline__ f3 aux3(const float4* __restrict__ X, uint32_t j) { return gatb<f3>(X, 32 * j + 16); }
__device__ __forceinline__ void put_aux3(float4* __restrict__ X, uint32_t i, f3 u)
{
    float* q = reinterpret_cast<float*>(X + 2 * size_t(i) + 1);
    q[0] = u.x;
    q[1] = u.y;
    q[2] = u.z;
}
__device__ __forceinline__ void put_auxw(float4* __restrict__ X, uint32_t i, float w)
{
    reinterpret_cast<float*>(X + 2 * size_t(i) + 1)[3] = w;
}

// ---- density ------------------------------------------------------------------

// Packed boundary records {x, y, z, m}.
__global__ __launch_bounds__(kBlock) void k_pack_bdy(BdyDev b, float4* __restrict__ B4)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < b.n)
        B4[k] = make_float4(b.x[k], b.y[k], b.z[k], b.m[k]);
}

// P4[i] = {x_i, y_i, z_i, m_i}, the records the density pass walks, and the
// interleaved records X of the later loops with u = v_i (see `aux4`).
// mflag[workgroup] = 1 when one of its particles has a mass other than the
// scene's first particle's.
template <bool BT>
__global__ __launch_bounds__(kBlock) void k_pack(SceneArg sa, FluidDev f, float4* __restrict__ P4,
    float4* __restrict__ X, uint32_t* __restrict__ mflag)
{
    REYN_SCENE(BT, sa);
    bool mixed = false;
    if (live) {
        const uint32_t i = t;
        // every load before the first store: the stores may alias the
        // loaded arrays as far as the compiler knows, and a load after them
        // would wait for them
        const float m = f.m[i], m0 = f.m[sc.first];
        const float4 r = make_float4(f.x[i], f.y[i], f.z[i], m);
        const float4 u = make_float4(f.vx[i], f.vy[i], f.vz[i], 0.f);
        P4[i] = r;
        X[2 * size_t(i)] = r;
        X[2 * size_t(i) + 1] = u;
        mixed = !(m == m0);
    }
    const int any = __syncthreads_or(mixed ? 1 : 0);
    if (threadIdx.x == 0)
        mflag[lb_] = any ? 1u : 0u;
}

// Densities; with GB also the boundary sums, and with a list in L the
// neighbour list of the step (every later fluid loop reads it).
template <int K, bool IND, bool GB, bool BT>
__global__ __launch_bounds__(kBlock) REYN_OCC(REYN_OCC_DENSITY) void k_density(SceneArg sa, FluidDev f,
    float* __restrict__ rho, float* __restrict__ gbx, float* __restrict__ gby,
    float* __restrict__ gbz, float* __restrict__ rb, NbrDev L, const float4* __restrict__ P4,
    float4* __restrict__ X)
{
    REYN_SCENE(BT, sa);
    const GridDev& g = sc.g;
    const BdyDev& b = sc.b;
    const KParams& kp = sc.c.kp;
    // lists serve ranges up to kListMaxRange; beyond, every later loop walks
    const bool listing = L.cnt && P4 && lists_serve(g);
    if (L.cnt && P4 && !listing && live)
        L.cnt[t] = L.cap + 1;
    if (!live && !listing)
        return;
    // lanes past the end stay for the list staging below, whose row flushes
    // take minima over the whole wavefront; they record nothing
    const uint32_t i = live ? particle_of<IND>(g, t) : 0u;
    const f3 xi = live ? ld3(f.x, f.y, f.z, i) : g.bmin;
    const uint32_t wv = threadIdx.x / kWave;
    __shared__ float4 stage[kBlock / kWave][kStage > 0 ? kStage : 1]; // strip staging, per wavefront
    ksum acc; // the walks' sum (the particle itself among the candidates)
    ksum_desc accl; // the listed sum: the particle itself first, then every
                    // neighbour, each term at most the first when masses are equal
    bool summed = false;
    if (listing) {
        // the walk only tests candidates and records neighbours; the sum runs
        // over the compact list afterwards, so the kernel body is not executed
        // once per candidate of every lane of the wavefront
        //
        // Lanes find their k-th neighbour at different times, so storing
        // entries as found would write each list row in many partial pieces.
        // Each lane instead keeps its entries in its own column of an LDS
        // ring, and after every strip the wavefront writes the rows that all
        // its lanes have filled, one contiguous row per store; the last rows
        // are written at the end. A lane whose ring is full stores its later
        // entries directly.
        __shared__ uint32_t ring[kBlock / kWave][kRing][kWave];
        const uint32_t lane = threadIdx.x % kWave;
        uint32_t cnt = 0, flushed = 0, direct = L.cap; // `flushed` is uniform
        // false once an entry's offset may not fit its 13 bits (tested per
        // neighbour in walk4, per strip in list_walk_staged)
        bool fits = true;
        // entries below `lim` = min(flushed + kRing, direct, cap) go to the
        // ring: one test on the common path
        uint32_t lim = min(kRing, L.cap);
        auto flush = [&](uint32_t end) { // rows [flushed, end) from the ring (`flushed` is even)
            const uint32_t have = min(cnt, direct);
            for (uint32_t r = flushed; r < end; r += 2) {
                // both entries of a word at once where this lane has both
                if (r + 1 < have && r + 1 < end)
                    put_pair(L, t, r, ring[wv][r % kRing][lane] | (ring[wv][(r + 1) % kRing][lane] << 16));
                else if (r < have)
                    put_entry(L, t, r, ring[wv][r % kRing][lane]);
            }
            flushed = end;
            lim = min(min(flushed + kRing, direct), L.cap);
        };
        auto record = [&](uint32_t e) {
            if (cnt < lim) {
                ring[wv][cnt % kRing][lane] = e;
            } else if (cnt < L.cap) {
                // the ring is full (or this lane already stores directly):
                // from here on the entries go straight to the list
                direct = min(direct, cnt);
                lim = 0;
                put_entry(L, t, cnt, e);
            }
            ++cnt;
        };
        auto rows_out = [&] {
            // write rows while every lane has filled the next kFlush (a lane
            // that stores directly, or one past the end, holds back no row)
            const uint32_t full = (!live || cnt >= direct) ? L.cap : min(cnt, L.cap);
            while (__ballot(full < flushed + kFlush) == 0 && flushed + kFlush <= L.cap)
                flush(flushed + kFlush);
        };
        if constexpr (IND || kStage == 0) {
            walk4<IND>(g, xi, P4,
                [&](uint32_t j, f3, float, float, const Cand& cd) {
                    if (!live || j == i)
                        return;
                    bool ok;
                    const uint32_t e = encode_entry(cd, &ok);
                    fits = fits && ok;
                    record(e);
                },
                rows_out);
        } else {
            list_walk_staged<kStage>(g, xi, P4, stage[wv], live, i, fits, record, rows_out);
        }
        flush(wave_max_u32(live ? min(cnt, direct) : 0u));
        if (!live)
            return;
        // a particle whose offsets do not all fit is not listed
        const uint32_t stored = fits ? cnt : L.cap + 1;
        L.cnt[t] = stored;
        if (stored <= L.cap) {
            accl.add(gat(f.m, i) * kW<K>(kp, 0.f)); // the particle itself
            nbrs4<IND, kNbrBatch<K>>(g, L, t, xi, P4, f.x, f.y, f.z, f.m,
                [&](uint32_t, f3, float d2, float mj) { accl.add(mj * kW<K>(kp, d2)); });
            summed = true;
        }
    }
    if (!summed) {
        if (P4)
            walk4<IND>(g, xi, P4, [&](uint32_t, f3, float d2, float mj, const Cand&) { acc.add(mj * kW<K>(kp, d2)); });
        else
            walk<IND>(g, xi, f.x, f.y, f.z, [&](uint32_t j, f3, float d2) { acc.add(gat(f.m, j) * kW<K>(kp, d2)); });
    }
    ksum accb;
    ksum3 grad;
    // a particle whose boundary cell has no sample within reach has no
    // boundary neighbour: its walk would only visit empty strips
    const bool bnear = b.present && (!b.near || b.near[cell_key(b.grid, cell_of(b.grid, xi))]);
    if (bnear) {
        auto visit = [&](f3 xik, float d2, float mk) {
            accb.add(mk * kW<K>(kp, d2));
            if (GB)
                grad.add(mk * kdW<K>(kp, xik, d2));
        };
        if (b.p4)
            walk4<false>(b.grid, xi, b.p4,
                [&](uint32_t, f3 xik, float d2, float mk, const Cand&) { visit(xik, d2, mk); });
        else
            walk<false>(b.grid, xi, b.x, b.y, b.z, [&](uint32_t k, f3 xik, float d2) { visit(xik, d2, b.m[k]); });
    }
    const float rbi = accb.value();
    const float r = (summed ? accl.value() : acc.value()) + rbi;
    rho[i] = r;
    if (X)
        put_auxw(X, i, r);
    if (GB) {
        const f3 gb = grad.value();
        gbx[i] = gb.x;
        gby[i] = gb.y;
        gbz[i] = gb.z;
        rb[i] = rbi;
    }
}

// ---- non-pressure accelerations (gravity + fluid-fluid viscosity) -----------

// PK: the interleaved records X hold {x, m} and {v, rho} of every particle
// (u = v, w = rho); otherwise the loop walks the SoA arrays.
template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) REYN_OCC(REYN_OCC_LOOP) void k_nonpressure(SceneArg sa, FluidDev f,
    const float* __restrict__ rho, float* __restrict__ ax, float* __restrict__ ay,
    float* __restrict__ az, NbrDev L, const float4* __restrict__ X)
{
    REYN_SCENE(BT, sa);
    if (!live)
        return;
    const GridDev& g = sc.g;
    const SphK& c = sc.c;
    const uint32_t i = particle_of<IND>(g, t);
    const f3 xi = ld3(f.x, f.y, f.z, i);
    const f3 vi = ld3(f.vx, f.vy, f.vz, i);
    f3 acc = c.g;
    nbrs4<IND, kNbrBatch<K>, 2>(g, L, t, xi, X, f.x, f.y, f.z, f.m, [&](uint32_t j, f3 xij, float d2, float mj) {
        float4 u;
        if constexpr (PK)
            u = aux4(X, j);
        else
            u = make_float4(gat(f.vx, j), gat(f.vy, j), gat(f.vz, j), gat(rho, j));
        const f3 vij = vi - f3 { u.x, u.y, u.z };
        const float s = c.visc_c * mj / u.w * dot(vij, xij) / (d2 + c.visc_eps);
        acc += s * kdW<K>(c.kp, xij, d2);
    });
    st3(ax, ay, az, i, acc);
}

// ---- integration --------------------------------------------------------------

template <bool BT>
__global__ __launch_bounds__(kBlock) void k_integrate(SceneArg sa, FluidDev f, const float* __restrict__ ax,
    const float* __restrict__ ay, const float* __restrict__ az, const SolveCtl* ctl, int positions,
    float4* __restrict__ X)
{
    REYN_SCENE(BT, sa);
    if (loop_done(ctl, sc) || !live)
        return;
    const uint32_t i = t;
    const float dt = sc.dt;
    // every load before the first store: the stores may alias the loaded
    // arrays as far as the compiler knows, and a load after them would wait
    const f3 v0 = ld3(f.vx, f.vy, f.vz, i), a = ld3(ax, ay, az, i);
    const f3 x0 = positions ? ld3(f.x, f.y, f.z, i) : f3 { 0.f, 0.f, 0.f };
    const f3 v = v0 + dt * a;
    st3(f.vx, f.vy, f.vz, i, v);
    if (X)
        put_aux3(X, i, v);
    if (positions)
        st3(f.x, f.y, f.z, i, x0 + dt * v);
}

// ---- SESPH ---------------------------------------------------------------------

__device__ __forceinline__ float eos_sesph(float rho, float k, float rho0)
{
    return fmaxf(0.f, k * (rho / rho0 - 1.f));
}

template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) void k_sesph_acc(SceneArg sa, FluidDev f, float k,
    const float* __restrict__ rho, const float* __restrict__ gbx,
    const float* __restrict__ gby, const float* __restrict__ gbz, float* __restrict__ ax,
    float* __restrict__ ay, float* __restrict__ az, NbrDev L, const float4* __restrict__ X)
{
    REYN_SCENE(BT, sa);
    if (!live)
        return;
    const GridDev& g = sc.g;
    const SphK& c = sc.c;
    const int bdy = sc.b.present;
    const uint32_t i = particle_of<IND>(g, t);
    const f3 xi = ld3(f.x, f.y, f.z, i);
    const f3 vi = ld3(f.vx, f.vy, f.vz, i);
    const float rhoi = rho[i];
    const float pri = eos_sesph(rhoi, k, c.rho0) / (rhoi * rhoi);
    f3 acc = c.g;
    nbrs4<IND, kNbrBatch<K>, 2>(g, L, t, xi, X, f.x, f.y, f.z, f.m, [&](uint32_t j, f3 xij, float d2, float mj) {
        float4 u; // {v_j, rho_j}
        if constexpr (PK)
            u = aux4(X, j);
        else
            u = make_float4(gat(f.vx, j), gat(f.vy, j), gat(f.vz, j), gat(rho, j));
        const float rhoj = u.w;
        const f3 vij = vi - f3 { u.x, u.y, u.z };
        const f3 dW = kdW<K>(c.kp, xij, d2);
        const float pj = eos_sesph(rhoj, k, c.rho0);
        const float visc = c.visc_c * mj / rhoj * dot(vij, xij) / (d2 + c.visc_eps);
        acc += (visc - mj * (pri + pj / (rhoj * rhoj))) * dW;
    });
    if (bdy)
        acc += -pri * ld3(gbx, gby, gbz, i);
    st3(ax, ay, az, i, acc);
}

// ---- PCISPH ----------------------------------------------------------------------

// predicted density; per-block FP64 partial sums feed the mean-density test
// PK: u of X holds the updated velocities, and w of P4 receives the
// prediction (after the density pass P4 serves the pressure accelerations
// only, which read {x, rho} from it).
template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) void k_predict_rho(SceneArg sa, FluidDev f,
    const float* __restrict__ gbx, const float* __restrict__ gby,
    const float* __restrict__ gbz, const float* __restrict__ rb, float* __restrict__ rho,
    double* __restrict__ part, const SolveCtl* ctl, NbrDev L, float4* __restrict__ X, float4* __restrict__ P4)
{
    __shared__ double lds[kBlock / kWave];
    REYN_SCENE(BT, sa);
    if (loop_done(ctl, sc))
        return;
    const GridDev& g = sc.g;
    const KParams& kp = sc.c.kp;
    const float dt = sc.dt;
    const int bdy = sc.b.present;
    double mine = 0.0;
    if (live) {
        const uint32_t i = particle_of<IND>(g, t);
        const f3 xi = ld3(f.x, f.y, f.z, i);
        const f3 vi = ld3(f.vx, f.vy, f.vz, i);
        ksum acc;
        const bool listed = nbrs4<IND, kNbrBatch<K>, 2>(g, L, t, xi, X, f.x, f.y, f.z, f.m, [&](uint32_t j, f3 xij, float d2, float mj) {
            f3 vj;
            if constexpr (PK)
                vj = aux3(X, j);
            else
                vj = gat3(f.vx, f.vy, f.vz, j);
            const f3 vij = vi - vj;
            acc.add(mj * (kW<K>(kp, d2) + dt * dot(kdW<K>(kp, xij, d2), vij)));
        });
        if (listed) // the list skips the particle itself: m_i (W(0) + dt grad W(0) . 0)
            acc.add(f.m[i] * kW<K>(kp, 0.f));
        if (bdy)
            acc.add(rb[i] + dt * dot(ld3(gbx, gby, gbz, i), vi));
        const float r = acc.value();
        rho[i] = r;
        if constexpr (PK)
            reinterpret_cast<float*>(P4 + i)[3] = r;
        mine = r;
    }
    const double s = block_reduce<0>(mine, lds);
    if (part && threadIdx.x == 0)
        part[lb_] = s;
}

// PK: w of P4 holds the predicted densities; the listed loop reads {x, rho}
// of each neighbour from one record when the scene's masses are all equal
// (m_j is then the particle's own mass) and loads m_j as well otherwise.
template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) void k_pcisph_prs_acc(SceneArg sa, FluidDev f,
    const float* __restrict__ rho, const float* __restrict__ gbx, const float* __restrict__ gby,
    const float* __restrict__ gbz, float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az,
    const SolveCtl* ctl, NbrDev L, const float4* __restrict__ P4)
{
    REYN_SCENE(BT, sa);
    if (loop_done(ctl, sc) || !live)
        return;
    const GridDev& g = sc.g;
    const KParams& kp = sc.c.kp;
    const float k = sc.kk, rho0 = sc.c.rho0, rho0_sq_inv = sc.c.rho0_sq_inv;
    const int bdy = sc.b.present;
    const uint32_t i = particle_of<IND>(g, t);
    const f3 xi = ld3(f.x, f.y, f.z, i);
    const float rhoi = rho[i];
    const float pi = fmaxf(0.f, k * (rhoi - rho0));
    const float pri = pi / (rhoi * rhoi);
    const float pr0 = pi * rho0_sq_inv;
    f3 acc { 0.f, 0.f, 0.f };
    auto term = [&](f3 xij, float d2, float mj, float rhoj) {
        const float pj = fmaxf(0.f, k * (rhoj - rho0));
        acc += (mj * (pri + pj / (rhoj * rhoj))) * kdW<K>(kp, xij, d2);
    };
    const uint32_t c = (PK && L.cnt) ? L.cnt[t] : 0xffffffffu;
    bool listed = false;
    if (PK && c <= L.cap) {
        const bool uniform = !ctl[sc.idx].mass_mixed; // the same for the whole scene
        const float mi = f.m[i];
        constexpr uint32_t B = kNbrBatch<K>;
        listed = with_list<B>(g, L, t, c, xi, [&](uint32_t kk, const uint32_t (&q)[B]) {
            uint32_t j[B];
            float4 r[B];
            float m[B];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                j[u] = IND ? gat(g.sorted, q[u]) : q[u];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                r[u] = gat(P4, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                m[u] = uniform ? mi : gat(f.m, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) {
                const f3 xij = xi - f3 { r[u].x, r[u].y, r[u].z };
                term(xij, dot(xij, xij), (kk + u < c) ? m[u] : 0.f, r[u].w);
            }
        });
    }
    if (!listed)
        walk<IND>(g, xi, f.x, f.y, f.z, [&](uint32_t j, f3 xij, float d2) { term(xij, d2, gat(f.m, j), gat(rho, j)); });
    f3 a = -acc;
    if (bdy)
        a += -(0.5f * (pri + pr0)) * ld3(gbx, gby, gbz, i);
    st3(ax, ay, az, i, a);
}

// One workgroup per scene: reduces the scene's per-block partials in a fixed
// order and advances its loop control.
template <bool BT>
__global__ __launch_bounds__(kBlock) void k_ctl_pcisph(SceneArg sa, SolveCtl* ctls, const double* __restrict__ part,
    uint32_t min_iter, uint32_t max_iter)
{
    __shared__ double lds[kBlock / kWave];
    const SceneDev& sc = scene_of_ctl<BT>(sa);
    SolveCtl* ctl = ctls + sc.idx;
    if (ctl->done)
        return;
    double s = strided_fold<8>(
        sc.nblk, 0.0, [&](uint32_t k) { return part[sc.blk_first + k]; }, [](double a, double v) { return a + v; });
    s = block_reduce<0>(s, lds);
    if (threadIdx.x == 0) {
        const uint32_t l = ctl->iter + 1;
        const double avg = s / double(sc.n);
        const bool keep = (l < min_iter) || (avg > double(sc.thr));
        ctl->iter = l;
        ctl->avg_density = avg;
        ctl->done = (!keep || (max_iter && l >= max_iter)) ? 1u : 0u;
    }
}

// ---- IISPH -------------------------------------------------------------------------

__device__ __forceinline__ float jacobi_update(float omega, float aii, float Ap, float s)
{
    // |a_ii| < 1e-6 compared in double in the reference; these float bounds
    // select exactly the same floats
    const float lim = 9.99999997475242708e-7f;
    return (aii <= lim && aii >= -lim) ? 0.f : fmaxf(0.f, omega * (s - Ap) / aii);
}

// PK: u of X holds the updated velocities, and the pressure goes to w of P4
// as well (after the density pass the records P4 serve the pressure
// accelerations only, which read {x, p} from them).
template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) REYN_OCC(REYN_OCC_LOOP) void k_iisph_aii_s(SceneArg sa, FluidDev f,
    float omega, int warm, const float* __restrict__ rho,
    const float* __restrict__ gbx, const float* __restrict__ gby, const float* __restrict__ gbz,
    float* __restrict__ aii, float* __restrict__ si, float* __restrict__ p, NbrDev L,
    float4* __restrict__ X, float4* __restrict__ P4)
{
    REYN_SCENE(BT, sa);
    if (!live)
        return;
    const GridDev& g = sc.g;
    const KParams& kp = sc.c.kp;
    const float rho0 = sc.c.rho0, dt = sc.dt;
    const int bdy = sc.b.present;
    const uint32_t i = particle_of<IND>(g, t);
    const f3 xi = ld3(f.x, f.y, f.z, i);
    const f3 vi = ld3(f.vx, f.vy, f.vz, i);
    ksum3 grad;
    ksum sq, div;
    nbrs4<IND, kNbrBatch<K>, 2>(g, L, t, xi, X, f.x, f.y, f.z, f.m, [&](uint32_t j, f3 xij, float d2, float mj) {
        f3 vj;
        if constexpr (PK)
            vj = aux3(X, j);
        else
            vj = gat3(f.vx, f.vy, f.vz, j);
        const f3 dW = kdW<K>(kp, xij, d2);
        const f3 mdW = mj * dW;
        div.add(dot(mdW, vi - vj));
        sq.add(dot(mdW, dW));
        grad.add(mdW);
    });
    f3 gb { 0.f, 0.f, 0.f };
    if (bdy)
        gb = ld3(gbx, gby, gbz, i);
    const f3 G = grad.value() + gb;
    const float a = -dt * (dot(G, G) + f.m[i] * sq.value()) / (rho0 * rho0);
    const float s = (rho0 - rho[i]) / dt - (div.value() + dot(vi, gb));
    const float pw = warm ? p[i] : 0.f; // loaded before the stores below
    aii[i] = a;
    si[i] = s;
    const float pn = warm ? pw * 0.5f : jacobi_update(omega, a, 0.f, s);
    p[i] = pn;
    if (PK)
        reinterpret_cast<float*>(P4 + i)[3] = pn;
}

// Pressure accelerations. PK: the listed loop reads {x, y, z, p} of each
// neighbour from the records P4 (w holds the pressure after the a_ii pass),
// one 16-byte load per neighbour when the scene's masses are all equal
// (`mctl`'s flag; m_j is then the particle's own mass) and a second load of
// m_j otherwise; inside the solver loop (ctl != NULL) it writes the result to
// u of X for the Jacobi pass; the final pass and the SoA mode write the work
// arrays.
template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) REYN_OCC(REYN_OCC_PACC) void k_iisph_pacc(SceneArg sa, FluidDev f,
    const float* __restrict__ p, const float* __restrict__ gbx,
    const float* __restrict__ gby, const float* __restrict__ gbz, float* __restrict__ ax,
    float* __restrict__ ay, float* __restrict__ az, const SolveCtl* ctl, const SolveCtl* mctl, NbrDev L,
    const float4* __restrict__ P4, float4* __restrict__ X)
{
    REYN_SCENE(BT, sa);
    if (loop_done(ctl, sc) || !live)
        return;
    const GridDev& g = sc.g;
    const KParams& kp = sc.c.kp;
    const float rho0_sq_inv = sc.c.rho0_sq_inv;
    const int bdy = sc.b.present;
    const uint32_t i = particle_of<IND>(g, t);
    const f3 xi = ld3(f.x, f.y, f.z, i);
    const float pi = p[i];
    f3 acc { 0.f, 0.f, 0.f };
    const uint32_t c = (PK && L.cnt) ? L.cnt[t] : 0xffffffffu;
    bool listed = false;
    if (PK && c <= L.cap) {
        const bool uniform = !mctl[sc.idx].mass_mixed; // the same for the whole scene
        const float mi = f.m[i];
        constexpr uint32_t B = kNbrBatch<K>;
        listed = with_list<B>(g, L, t, c, xi, [&](uint32_t k, const uint32_t (&q)[B]) {
            uint32_t j[B];
            float4 r[B];
            float m[B];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                j[u] = IND ? gat(g.sorted, q[u]) : q[u];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                r[u] = gat(P4, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                m[u] = uniform ? mi : gat(f.m, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) {
                const f3 xij = xi - f3 { r[u].x, r[u].y, r[u].z };
                const float mj = (k + u < c) ? m[u] : 0.f;
                acc += (mj * (pi + r[u].w)) * kdW<K>(kp, xij, dot(xij, xij));
            }
        });
    }
    if (!listed)
        walk<IND>(g, xi, f.x, f.y, f.z, [&](uint32_t j, f3 xij, float d2) {
            acc += (gat(f.m, j) * (pi + gat(p, j))) * kdW<K>(kp, xij, d2);
        });
    if (bdy)
        acc += pi * ld3(gbx, gby, gbz, i);
    const f3 a = -rho0_sq_inv * acc;
    if (PK && ctl)
        put_aux3(X, i, a);
    else
        st3(ax, ay, az, i, a);
}

// Dynamic LDS per workgroup (unused by the kernel) that caps a launch at
// `waves` waves per SIMD: the workgroups of one more wave per SIMD (4 SIMDs
// per CU) would not fit the 160 KB of LDS of a CU. 0: no cap. At most the
// 64 KB a launch may ask for (a cap below two workgroups per CU is not made).
constexpr uint32_t occupancy_lds(uint32_t waves)
{
    return waves ? std::min<uint32_t>(64u * 1024u,
                       (160u * 1024u) / (std::max<uint32_t>(1u, waves * 4u / (kBlock / kWave)) + 1u) + 256u)
                 : 0u;
}
static_assert((occupancy_lds(3) == 41216u && occupancy_lds(6) == 23661u) || kBlock != 256,
    "caps of 3 and 6 waves per SIMD with 256-lane workgroups");

// Waves per SIMD of the Jacobi pass: its two record gathers per neighbour
// keep more lines in flight than L1 and L2 serve well, so it runs fastest at
// 3 waves per SIMD (late 2M state: 231 / 204.5 / 206.7 / 210.8 us at 2 / 3 /
// 4 / 5 waves), below what its register count alone would allow. The
// one-record pressure-acceleration pass gains a little from a cap as well
// (122.4 us at its register-bound 7 waves, 119.6 us at 6).
#ifndef REYN_JACOBI_WAVES
#define REYN_JACOBI_WAVES 3
#endif
#ifndef REYN_PACC_WAVES
#define REYN_PACC_WAVES 6
#endif

// PK: the accelerations come from u of X, and the new pressure goes to w of
// P4 as well; otherwise the SoA work arrays.
template <int K, bool IND, bool PK, bool BT>
__global__ __launch_bounds__(kBlock) REYN_OCC(REYN_OCC_LOOP) void k_iisph_jacobi(SceneArg sa, FluidDev f,
    float omega, const float* __restrict__ ax, const float* __restrict__ ay,
    const float* __restrict__ az, const float* __restrict__ aii,
    const float* __restrict__ si, const float* __restrict__ gbx, const float* __restrict__ gby,
    const float* __restrict__ gbz, float* __restrict__ p, double* __restrict__ part_err,
    uint32_t* __restrict__ part_cnt, const SolveCtl* ctl, NbrDev L, float4* __restrict__ X,
    float4* __restrict__ P4)
{
    __shared__ double lds[kBlock / kWave];
    REYN_SCENE(BT, sa);
    if (loop_done(ctl, sc))
        return;
    const GridDev& g = sc.g;
    const KParams& kp = sc.c.kp;
    const float dt = sc.dt, thr = sc.thr;
    const int bdy = sc.b.present;
    double err_pos = 0.0;
    uint32_t viol = 0;
    if (live) {
        const uint32_t i = particle_of<IND>(g, t);
        const f3 xi = ld3(f.x, f.y, f.z, i);
        f3 ai;
        if constexpr (PK)
            ai = aux3(X, i);
        else
            ai = ld3(ax, ay, az, i);
        float lap = 0.f;
        nbrs4<IND, kNbrBatch<K>, 2>(g, L, t, xi, X, f.x, f.y, f.z, f.m, [&](uint32_t j, f3 xij, float d2, float mj) {
            f3 aj;
            if constexpr (PK)
                aj = aux3(X, j);
            else
                aj = gat3(ax, ay, az, j);
            lap += mj * dot(ai - aj, kdW<K>(kp, xij, d2));
        });
        if (bdy)
            lap += dot(ai, ld3(gbx, gby, gbz, i));
        const float Ap = dt * lap;
        const float s = si[i];
        const float pn = p[i] + jacobi_update(omega, aii[i], Ap, s);
        p[i] = pn;
        if (PK)
            reinterpret_cast<float*>(P4 + i)[3] = pn;
        const float err = (Ap - s) * dt;
        viol = err > thr ? 1u : 0u;
        err_pos = err > 0.f ? double(err) : 0.0;
    }
    const uint64_t vb = __ballot(viol != 0);
    const double es = block_reduce<0>(err_pos, lds);
    __shared__ uint32_t vc[kBlock / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0)
        vc[threadIdx.x / kWave] = static_cast<uint32_t>(__popcll(vb));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int w = 0; w < kBlock / kWave; ++w)
            c += vc[w];
        part_err[lb_] = es;
        part_cnt[lb_] = c;
    }
}

template <bool BT>
__global__ __launch_bounds__(kBlock) void k_ctl_iisph(SceneArg sa, SolveCtl* ctls, const double* __restrict__ part_err,
    const uint32_t* __restrict__ part_cnt, float eta, uint32_t min_iter, uint32_t max_iter, int rule)
{
    __shared__ double lds[kBlock / kWave];
    const SceneDev& sc = scene_of_ctl<BT>(sa);
    SolveCtl* ctl = ctls + sc.idx;
    if (ctl->done)
        return;
    const uint32_t n = sc.n;
    const float rho0 = sc.c.rho0;
    const double2 sc2 = strided_fold<8>(
        sc.nblk, make_double2(0.0, 0.0),
        [&](uint32_t k) { return make_double2(part_err[sc.blk_first + k], double(part_cnt[sc.blk_first + k])); },
        [](double2 a, double2 v) { return make_double2(a.x + v.x, a.y + v.y); });
    double s = sc2.x, c = sc2.y;
    s = block_reduce<0>(s, lds);
    c = block_reduce<0>(c, lds);
    if (threadIdx.x == 0) {
        const uint32_t l = ctl->iter + 1;
        const double mean = (n > 0) ? s / double(n) / double(rho0) : 0.0;
        const uint32_t viol = static_cast<uint32_t>(c);
        const bool ok = (rule == REYN_IISPH_STOP_MEAN) ? (mean <= double(eta)) : (viol == 0);
        ctl->iter = l;
        ctl->n_violating = viol;
        ctl->mean_rel_error = mean;
        ctl->done = ((l >= min_iter && ok) || (max_iter && l >= max_iter)) ? 1u : 0u;
    }
}

// One workgroup per scene: resets the scene's loop control and reduces the
// per-workgroup mass flags of the pack pass (NULL: treated as mixed).
template <bool BT>
__global__ __launch_bounds__(kBlock) void k_ctl_reset(SceneArg sa, SolveCtl* ctl, const uint32_t* __restrict__ mflag)
{
    const SceneDev& sc = scene_of_ctl<BT>(sa);
    uint32_t mixed = mflag ? 0u : 1u;
    if (mflag)
        mixed = strided_fold<8>(
            sc.nblk, mixed, [&](uint32_t k) { return mflag[sc.blk_first + k]; },
            [](uint32_t a, uint32_t v) { return a | v; });
    const int any = __syncthreads_or(mixed ? 1 : 0);
    if (threadIdx.x == 0) {
        SolveCtl& c = ctl[sc.idx];
        c.iter = 0;
        c.done = 0;
        c.n_violating = 0;
        c.mass_mixed = any ? 1u : 0u;
        c.mean_rel_error = 0.0;
        c.avg_density = 0.0;
    }
}

// ---- host-side helpers ---------------------------------------------------------------

// One launch target: the fluid arrays, the scene argument (one scene by value,
// or a batch table), whether the grid indexes an unsorted state, and the
// number of workgroups.
struct Common {
    FluidDev f;
    SceneArg a;
    bool ind;
    uint32_t nb;
};

void one_scene(Common* c, const FluidDev& f, uint32_t nb)
{
    c->f = f;
    c->nb = nb;
    c->ind = false;
    c->a = SceneArg {};
    SceneDev& s = c->a.one;
    s.first = 0;
    s.n = f.n;
    s.blk_first = 0;
    s.nblk = nb;
    s.idx = 0;
}

int prepare(reyn_ctx* ctx, const reyn_sph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, Common* c)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_TRY(check_sph(ctx, prm));
    REYN_TRY(check_fluid(ctx, fluid));
    REYN_TRY(check_grid(ctx, grid, fluid->n, "fluid"));
    REYN_TRY(check_bdy(ctx, bdy));
    if (fluid->n >= (1u << 28))
        return set_error(ctx, REYN_ERR_CAPACITY, "at most 2^28 - 1 fluid particles per call (got %u)", fluid->n);
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    one_scene(c, to_dev(*fluid), blocks_for(fluid->n));
    c->ind = grid->sorted != nullptr;
    c->a.one.g = to_dev(*grid);
    c->a.one.b = to_dev(bdy);
    c->a.one.c = make_sphk(*prm);
    return REYN_OK;
}

// Dispatches on whether the step has interleaved records.
template <class F> int with_pk(const float4* X, F&& f)
{
    return X ? f(std::true_type {}) : f(std::false_type {});
}

// A batch always reorders its state and packs records: only those variants of
// the neighbour kernels are built for it.
template <bool BT, bool IND, bool PK> constexpr bool kVariant = !BT || (!IND && PK);

int no_variant(reyn_ctx* ctx) { return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "unsupported batch layout"); }

template <bool BT, bool GB>
int launch_density(reyn_ctx* ctx, int kernel, const Common& c0, float* rho, const Scratch* s)
{
    // one scene: its boundary is packed into scratch for this pass (a batch
    // packs its boundaries once)
    Common c = c0;
    if (!BT && s && s->B4 && c.a.one.b.present && !c.a.one.b.p4) {
        {
            ProfScope _prof(ctx, "k_pack_bdy");
            k_pack_bdy<<<blocks_for(c.a.one.b.n), kBlock, 0, ctx->stream>>>(c.a.one.b, s->B4);
        }
        REYN_LAUNCHED(ctx, "k_pack_bdy");
        c.a.one.b.p4 = s->B4;
    }
    if (s && s->P4) {
        {
            ProfScope _prof(ctx, "k_pack");
            k_pack<BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, s->P4, s->X, s->part_u);
        }
        REYN_LAUNCHED(ctx, "k_pack");
    }
    return with_kernel(ctx, kernel, c.ind, [&](auto KC, auto IC) -> int {
        constexpr int K = decltype(KC)::value;
        constexpr bool I = decltype(IC)::value;
        if constexpr (!kVariant<BT, I, true>) {
            return no_variant(ctx);
        } else {
            {
                ProfScope _prof(ctx, "k_density");
                k_density<K, I, GB, BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, rho,
                    GB ? s->gbx : nullptr, GB ? s->gby : nullptr, GB ? s->gbz : nullptr, GB ? s->rb : nullptr,
                    s ? s->L : NbrDev {}, s ? s->P4 : nullptr, s ? s->X : nullptr);
            }
            REYN_LAUNCHED(ctx, "k_density");
            return REYN_OK;
        }
    });
}

template <bool BT>
int launch_nonpressure(reyn_ctx* ctx, int kernel, const Common& c, const float* rho, float* ax, float* ay,
    float* az, const NbrDev& L, const float4* X)
{
    return with_kernel(ctx, kernel, c.ind, [&](auto KC, auto IC) -> int {
        return with_pk(X, [&](auto PC) -> int {
            constexpr int K = decltype(KC)::value;
            constexpr bool I = decltype(IC)::value;
            constexpr bool PK = decltype(PC)::value;
            if constexpr (!kVariant<BT, I, PK>) {
                return no_variant(ctx);
            } else {
                {
                    ProfScope _prof(ctx, "k_nonpressure");
                    k_nonpressure<K, I, PK, BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, rho, ax, ay, az, L, X);
                }
                REYN_LAUNCHED(ctx, "k_nonpressure");
                return REYN_OK;
            }
        });
    });
}

template <bool BT>
int launch_integrate(reyn_ctx* ctx, const Common& c, const float* ax, const float* ay, const float* az,
    const SolveCtl* ctl, int positions, float4* X = nullptr)
{
    {
        ProfScope _prof(ctx, positions ? "k_integrate" : "k_integrate_v");
        k_integrate<BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, ax, ay, az, ctl, positions, X);
    }
    REYN_LAUNCHED(ctx, "k_integrate");
    return REYN_OK;
}

// Enqueue loop iterations in chunks until every scene's control block
// reports done; `nscenes` control blocks at `ctl`, copied into `out`. The
// first chunk is min_iter iterations, or for one scene with a `hint` (the
// iteration count of the previous solve of the same kind on this context)
// the hint itself: a step needs about as many iterations as the one before,
// so most solves end at their first host check, at the price of an
// iteration or two that return at once when the step needs fewer. Launches
// after convergence return at once, so the chunking never changes a result.
template <class Iter>
int run_loop(reyn_ctx* ctx, const Common& c, const Scratch& s, uint32_t nscenes, uint32_t min_iter,
    uint32_t max_iter, uint32_t check_every, Iter&& iter, SolveCtl* out, uint32_t* hint = nullptr)
{
    SolveCtl* ctl = s.ctl;
    SolveCtl* host = nullptr;
    REYN_TRY(ctx_pinned(ctx, sizeof(SolveCtl) * nscenes, reinterpret_cast<void**>(&host)));
    {
        ProfScope _prof(ctx, "k_ctl_reset");
        k_ctl_reset<false><<<1, kBlock, 0, ctx->stream>>>(c.a, ctl, s.P4 ? s.part_u : nullptr);
    }
    REYN_LAUNCHED(ctx, "k_ctl_reset");
    const uint32_t chunk = check_every ? check_every : 4;
    // without a cap the reference loops until convergence; this safety cap
    // only turns a runaway solve into an error instead of an endless loop
    const uint32_t hard_cap = max_iter ? max_iter : 100000u;
    uint32_t launched = 0;
    for (;;) {
        // the first chunk: the iterations that must run anyway, or the hint
        // (not after a solve that stopped at max_iter: its count says nothing
        // of how soon this one converges)
        const bool use_hint = hint && nscenes == 1 && *hint > 1 && !(max_iter && *hint >= max_iter);
        const uint32_t guess = use_hint ? *hint : 1u;
        uint32_t m = (launched == 0) ? std::max(min_iter, guess) : chunk;
        m = std::min(m, hard_cap - launched);
        for (uint32_t k = 0; k < m; ++k)
            REYN_TRY(iter());
        launched += m;
        REYN_HIP(ctx, hipMemcpyAsync(host, ctl, sizeof(SolveCtl) * nscenes, hipMemcpyDeviceToHost, ctx->stream));
        REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        bool all = true;
        for (uint32_t k = 0; k < nscenes; ++k)
            all = all && host[k].done;
        if (all)
            break;
        if (launched >= hard_cap)
            return set_error(ctx, REYN_ERR_NOT_CONVERGED, "solver did not converge in %u iterations", launched);
    }
    for (uint32_t k = 0; k < nscenes; ++k)
        out[k] = host[k];
    if (hint && nscenes == 1)
        *hint = host[0].iter;
    return REYN_OK;
}


// The solver steps on a launch target: one scene (BT = false) or a batch
// (BT = true, every scene advancing with its own time step and loop control).

template <bool BT>
int sesph_run(reyn_ctx* ctx, int kernel, const Common& c, float k, const reyn_sesph_work* work, const Scratch& s)
{
    REYN_TRY((launch_density<BT, true>(ctx, kernel, c, work->rho, &s)));
    REYN_TRY(with_kernel(ctx, kernel, c.ind, [&](auto KC, auto IC) -> int {
        return with_pk(s.X, [&](auto PC) -> int {
            constexpr int K = decltype(KC)::value;
            constexpr bool I = decltype(IC)::value;
            constexpr bool PK = decltype(PC)::value;
            if constexpr (!kVariant<BT, I, PK>) {
                return no_variant(ctx);
            } else {
                {
                    ProfScope _prof(ctx, "k_sesph_acc");
                    k_sesph_acc<K, I, PK, BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, k, work->rho, s.gbx,
                        s.gby, s.gbz, work->ax, work->ay, work->az, s.L, s.X);
                }
                REYN_LAUNCHED(ctx, "k_sesph_acc");
                return REYN_OK;
            }
        });
    }));
    return launch_integrate<BT>(ctx, c, work->ax, work->ay, work->az, nullptr, 1);
}

template <bool BT>
int pcisph_begin(reyn_ctx* ctx, int kernel, const Common& c, const reyn_pcisph_work* work, const Scratch& s)
{
    REYN_TRY((launch_density<BT, true>(ctx, kernel, c, work->rho, &s)));
    return launch_nonpressure<BT>(ctx, kernel, c, work->rho, work->ax, work->ay, work->az, s.L, s.X);
}

// One loop iteration: velocities with the current accelerations, predicted
// densities, pressure accelerations, then the loop control of every scene.
template <bool BT>
int pcisph_iteration(reyn_ctx* ctx, int kernel, const Common& c, const reyn_pcisph_work* work, const Scratch& s,
    uint32_t nscenes, uint32_t min_iter, uint32_t max_iter)
{
    REYN_TRY(launch_integrate<BT>(ctx, c, work->ax, work->ay, work->az, s.ctl, 0, s.X));
    REYN_TRY(with_kernel(ctx, kernel, c.ind, [&](auto KC, auto IC) -> int {
        return with_pk(s.X, [&](auto PC) -> int {
            constexpr int K = decltype(KC)::value;
            constexpr bool I = decltype(IC)::value;
            constexpr bool PK = decltype(PC)::value;
            if constexpr (!kVariant<BT, I, PK>) {
                return no_variant(ctx);
            } else {
                {
                    ProfScope _prof(ctx, "k_predict_rho");
                    k_predict_rho<K, I, PK, BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, s.gbx, s.gby, s.gbz,
                        s.rb, work->rho, s.part_d, s.ctl, s.L, s.X, s.P4);
                }
                REYN_LAUNCHED(ctx, "k_predict_rho");
                {
                    ProfScope _prof(ctx, "k_pcisph_prs_acc");
                    k_pcisph_prs_acc<K, I, PK, BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, work->rho, s.gbx,
                        s.gby, s.gbz, work->ax, work->ay, work->az, s.ctl, s.L, s.P4);
                }
                REYN_LAUNCHED(ctx, "k_pcisph_prs_acc");
                return REYN_OK;
            }
        });
    }));
    {
        ProfScope _prof(ctx, "k_ctl_pcisph");
        k_ctl_pcisph<BT><<<nscenes, kBlock, 0, ctx->stream>>>(c.a, s.ctl, s.part_d, min_iter, max_iter);
    }
    REYN_LAUNCHED(ctx, "k_ctl_pcisph");
    return REYN_OK;
}

template <bool BT>
int pcisph_end(reyn_ctx* ctx, const Common& c, const reyn_pcisph_work* work)
{
    return launch_integrate<BT>(ctx, c, work->ax, work->ay, work->az, nullptr, 1);
}

template <bool BT>
int pcisph_run(reyn_ctx* ctx, int kernel, const Common& c, const reyn_pcisph_work* work, const Scratch& s,
    uint32_t nscenes, uint32_t min_iter, uint32_t max_iter, uint32_t check_every, SolveCtl* res)
{
    REYN_TRY(pcisph_begin<BT>(ctx, kernel, c, work, s));
    REYN_TRY(run_loop(
        ctx, c, s, nscenes, min_iter, max_iter, check_every,
        [&]() -> int { return pcisph_iteration<BT>(ctx, kernel, c, work, s, nscenes, min_iter, max_iter); },
        res, BT ? nullptr : &ctx->loop_hint[0]));
    return pcisph_end<BT>(ctx, c, work);
}

struct IisphLoop {
    float omega, eta;
    uint32_t min_iter, max_iter, check_every;
    int rule, warm;
};

template <bool BT>
int iisph_pacc(reyn_ctx* ctx, int kind, const Common& c, const reyn_iisph_work* work, const Scratch& s,
    const SolveCtl* ctl)
{
    return with_kernel(ctx, kind, c.ind, [&](auto KC, auto IC) -> int {
        return with_pk(s.X, [&](auto PC) -> int {
            constexpr int K = decltype(KC)::value;
            constexpr bool I = decltype(IC)::value;
            constexpr bool PK = decltype(PC)::value;
            if constexpr (!kVariant<BT, I, PK>) {
                return no_variant(ctx);
            } else {
                {
                    ProfScope _prof(ctx, "k_iisph_pacc");
                    k_iisph_pacc<K, I, PK, BT><<<c.nb, kBlock, occupancy_lds(REYN_PACC_WAVES), ctx->stream>>>(c.a, c.f, work->p, s.gbx, s.gby,
                        s.gbz, work->ax, work->ay, work->az, ctl, s.ctl, s.L, s.P4, s.X);
                }
                REYN_LAUNCHED(ctx, "k_iisph_pacc");
                return REYN_OK;
            }
        });
    });
}

template <bool BT>
int iisph_begin(reyn_ctx* ctx, int kind, const Common& c, const IisphLoop& lp, const reyn_iisph_work* work,
    const Scratch& s)
{
    REYN_TRY((launch_density<BT, true>(ctx, kind, c, work->rho, &s)));
    REYN_TRY(launch_nonpressure<BT>(ctx, kind, c, work->rho, work->ax, work->ay, work->az, s.L, s.X));
    REYN_TRY(launch_integrate<BT>(ctx, c, work->ax, work->ay, work->az, nullptr, 0, s.X));
    return with_kernel(ctx, kind, c.ind, [&](auto KC, auto IC) -> int {
        return with_pk(s.X, [&](auto PC) -> int {
            constexpr int K = decltype(KC)::value;
            constexpr bool I = decltype(IC)::value;
            constexpr bool PK = decltype(PC)::value;
            if constexpr (!kVariant<BT, I, PK>) {
                return no_variant(ctx);
            } else {
                {
                    ProfScope _prof(ctx, "k_iisph_aii_s");
                    k_iisph_aii_s<K, I, PK, BT><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, lp.omega, lp.warm,
                        work->rho, s.gbx, s.gby, s.gbz, work->a_ii, work->s_i, work->p, s.L, s.X, s.P4);
                }
                REYN_LAUNCHED(ctx, "k_iisph_aii_s");
                return REYN_OK;
            }
        });
    });
}

// One loop iteration: pressure accelerations, the relaxed Jacobi update, then
// the loop control of every scene.
template <bool BT>
int iisph_iteration(reyn_ctx* ctx, int kind, const Common& c, const IisphLoop& lp, const reyn_iisph_work* work,
    const Scratch& s, uint32_t nscenes)
{
    REYN_TRY(iisph_pacc<BT>(ctx, kind, c, work, s, s.ctl));
    REYN_TRY(with_kernel(ctx, kind, c.ind, [&](auto KC, auto IC) -> int {
        return with_pk(s.X, [&](auto PC) -> int {
            constexpr int K = decltype(KC)::value;
            constexpr bool I = decltype(IC)::value;
            constexpr bool PK = decltype(PC)::value;
            if constexpr (!kVariant<BT, I, PK>) {
                return no_variant(ctx);
            } else {
                {
                    ProfScope _prof(ctx, "k_iisph_jacobi");
                    k_iisph_jacobi<K, I, PK, BT><<<c.nb, kBlock, occupancy_lds(REYN_JACOBI_WAVES), ctx->stream>>>(c.a, c.f, lp.omega, work->ax,
                        work->ay, work->az, work->a_ii, work->s_i, s.gbx, s.gby, s.gbz, work->p, s.part_d,
                        s.part_u, s.ctl, s.L, s.X, s.P4);
                }
                REYN_LAUNCHED(ctx, "k_iisph_jacobi");
                return REYN_OK;
            }
        });
    }));
    {
        ProfScope _prof(ctx, "k_ctl_iisph");
        k_ctl_iisph<BT><<<nscenes, kBlock, 0, ctx->stream>>>(c.a, s.ctl, s.part_d, s.part_u, lp.eta, lp.min_iter,
            lp.max_iter, lp.rule);
    }
    REYN_LAUNCHED(ctx, "k_ctl_iisph");
    return REYN_OK;
}

template <bool BT>
int iisph_end(reyn_ctx* ctx, int kind, const Common& c, const reyn_iisph_work* work, const Scratch& s)
{
    REYN_TRY(iisph_pacc<BT>(ctx, kind, c, work, s, nullptr));
    return launch_integrate<BT>(ctx, c, work->ax, work->ay, work->az, nullptr, 1);
}

template <bool BT>
int iisph_run(reyn_ctx* ctx, int kind, const Common& c, const IisphLoop& lp, const reyn_iisph_work* work,
    const Scratch& s, uint32_t nscenes, SolveCtl* res)
{
    REYN_TRY(iisph_begin<BT>(ctx, kind, c, lp, work, s));
    REYN_TRY(run_loop(
        ctx, c, s, nscenes, lp.min_iter, lp.max_iter, lp.check_every,
        [&]() -> int { return iisph_iteration<BT>(ctx, kind, c, lp, work, s, nscenes); }, res,
        BT ? nullptr : &ctx->loop_hint[1]));
    return iisph_end<BT>(ctx, kind, c, work, s);
}

// The device state of a batch as a launch target with its own scratch.
void batch_common(const BatchTarget& bt, Common* c, Scratch* s)
{
    c->f = bt.f;
    c->a = bt.a;
    c->ind = false;
    c->nb = bt.a.blk_list ? bt.nlaunch : bt.nblocks;
    carve_scratch(static_cast<char*>(bt.scratch), bt.f.n, s, true, bt.nblocks, bt.nscenes, 0);
}

} // namespace

size_t batch_scratch_bytes(uint32_t n, uint32_t nblocks, uint32_t nscenes)
{
    Scratch s;
    return carve_scratch(nullptr, n, &s, true, nblocks, nscenes, 0);
}

const void* batch_loop_control(const BatchTarget& bt)
{
    Common c;
    Scratch s;
    batch_common(bt, &c, &s);
    return s.ctl;
}

size_t batch_loop_control_bytes(uint32_t nscenes) { return sizeof(SolveCtl) * nscenes; }

namespace {
__global__ void k_save_iterations(const SolveCtl* __restrict__ ctl, uint32_t* __restrict__ out, uint32_t n)
{
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x)
        out[k] = ctl[k].iter;
}
} // namespace

int batch_save_iterations(reyn_ctx* ctx, const BatchTarget& bt, uint32_t* iters_dev)
{
    Common c;
    Scratch s;
    batch_common(bt, &c, &s);
    {
        ProfScope _prof(ctx, "k_save_iterations");
        k_save_iterations<<<1, kBlock, 0, ctx->stream>>>(s.ctl, iters_dev, bt.nscenes);
    }
    REYN_LAUNCHED(ctx, "k_save_iterations");
    return REYN_OK;
}

void batch_read_control(const void* host, uint32_t nscenes, int solver, uint32_t* iters, double* stats,
    uint8_t* scene_done, int* done)
{
    const SolveCtl* c = static_cast<const SolveCtl*>(host);
    *done = 1;
    for (uint32_t k = 0; k < nscenes; ++k) {
        iters[k] = c[k].iter;
        stats[k] = solver == REYN_SOLVER_PCISPH ? c[k].avg_density : c[k].mean_rel_error;
        scene_done[k] = c[k].done ? 1 : 0;
        *done = *done && c[k].done;
    }
}

int batch_phase(reyn_ctx* ctx, const BatchTarget& bt, const BatchSolver& bs, int phase, uint32_t count)
{
    Common c;
    Scratch s;
    batch_common(bt, &c, &s);
    if (!s.X)
        return set_error(ctx, REYN_ERR_CAPACITY, "batch: %u particles exceed the packed-record limit", bt.f.n);
    const IisphLoop lp { bs.omega, bs.eta, bs.min_iter, bs.max_iter, bs.check_every, bs.stop_rule, bs.warm };
    const reyn_pcisph_work pw { bt.rho, bt.ax, bt.ay, bt.az };
    const reyn_iisph_work iw { bt.rho, bt.p, bt.a_ii, bt.s_i, bt.ax, bt.ay, bt.az };
    const bool pc = bs.solver == REYN_SOLVER_PCISPH;
    if (!pc && bs.solver != REYN_SOLVER_IISPH)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown solver %d", bs.solver);
    switch (phase) {
    case BATCH_BEGIN:
        REYN_TRY(pc ? pcisph_begin<true>(ctx, bs.kernel, c, &pw, s) : iisph_begin<true>(ctx, bs.kernel, c, lp, &iw, s));
        {
            ProfScope _prof(ctx, "k_ctl_reset");
            k_ctl_reset<true><<<bt.nscenes, kBlock, 0, ctx->stream>>>(c.a, s.ctl, s.part_u);
        }
        REYN_LAUNCHED(ctx, "k_ctl_reset");
        [[fallthrough]];
    case BATCH_ITERATE:
        for (uint32_t k = 0; k < count; ++k)
            REYN_TRY(pc ? pcisph_iteration<true>(ctx, bs.kernel, c, &pw, s, bt.nscenes, bs.min_iter, bs.max_iter)
                        : iisph_iteration<true>(ctx, bs.kernel, c, lp, &iw, s, bt.nscenes));
        return REYN_OK;
    case BATCH_END:
        return pc ? pcisph_end<true>(ctx, c, &pw) : iisph_end<true>(ctx, bs.kernel, c, &iw, s);
    default:
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "batch: unknown phase %d", phase);
    }
}

// ---- static scratch of the entry point that allocates nothing --------------
// reyn_sesph_step_noalloc keeps its scratch (the boundary sums of the density
// pass, 16 bytes per fluid particle, and the per-workgroup partials and loop
// control) in this file-scope array instead of the context's grown scratch.
// Every GPU the library runs on holds its own copy of the array until the
// process ends. In host code `scratch` only names the symbol: its device
// address on a device comes from hipGetSymbolAddress with that device
// current, taken once per device and kept in `noalloc_ptr`.
constexpr uint32_t kNoAllocMaxFluid = 1u << 20;
constexpr size_t kNoAllocScratchBytes = 17u << 20; // 16 MiB of sums + partials and control
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
std::mutex noalloc_mu;
char* noalloc_ptr[kNoAllocDevices] = {};
} // namespace

static int noalloc_scratch(reyn_ctx* ctx, char** out)
{
    if (ctx->device < 0 || ctx->device >= kNoAllocDevices)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "no-alloc scratch: device %d out of range", ctx->device);
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[ctx->device]) {
        REYN_HIP(ctx, hipSetDevice(ctx->device));
        void* p = nullptr;
        REYN_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        noalloc_ptr[ctx->device] = static_cast<char*>(p);
    }
    *out = noalloc_ptr[ctx->device];
    return REYN_OK;
}

} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_density(reyn_ctx* ctx, const reyn_sph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, float* rho)
{
    Common c;
    REYN_TRY(prepare(ctx, prm, fluid, grid, bdy, &c));
    if (!rho && fluid->n)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "density: rho is NULL");
    if (fluid->n == 0)
        return REYN_OK;
    return launch_density<false, false>(ctx, prm->kernel, c, rho, nullptr);
}

int reyn_nonpressure_acc(reyn_ctx* ctx, const reyn_sph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const float* rho, float* ax, float* ay, float* az)
{
    Common c;
    REYN_TRY(prepare(ctx, prm, fluid, grid, nullptr, &c));
    if (fluid->n == 0)
        return REYN_OK;
    if (!rho || !ax || !ay || !az)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "nonpressure_acc: NULL buffer");
    return launch_nonpressure<false>(ctx, prm->kernel, c, rho, ax, ay, az, NbrDev {}, nullptr);
}

int reyn_predict_rho(reyn_ctx* ctx, const reyn_sph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, float dt, float* rho)
{
    Common c;
    REYN_TRY(prepare(ctx, prm, fluid, grid, bdy, &c));
    if (fluid->n == 0)
        return REYN_OK;
    if (!rho)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "predict_rho: rho is NULL");
    c.a.one.dt = dt;
    Scratch s;
    REYN_TRY(get_scratch(ctx, fluid->n, &s));
    // the boundary sums come from a density pass into scratch; its density
    // output lands in `rho` and is overwritten by the prediction
    if (c.a.one.b.present)
        REYN_TRY((launch_density<false, true>(ctx, prm->kernel, c, rho, &s)));
    return with_kernel(ctx, prm->kernel, c.ind, [&](auto KC, auto IC) -> int {
        constexpr int K = decltype(KC)::value;
        constexpr bool I = decltype(IC)::value;
        {
            ProfScope _prof(ctx, "k_predict_rho");
            k_predict_rho<K, I, false, false><<<c.nb, kBlock, 0, ctx->stream>>>(c.a, c.f, s.gbx, s.gby, s.gbz,
                s.rb, rho, nullptr, nullptr, NbrDev {}, nullptr, nullptr);
        }
        REYN_LAUNCHED(ctx, "k_predict_rho");
        return REYN_OK;
    });
}

int reyn_integrate(reyn_ctx* ctx, const reyn_particles* fluid, const float* ax, const float* ay,
    const float* az, float dt)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_TRY(check_fluid(ctx, fluid));
    if (fluid->n == 0)
        return REYN_OK;
    Common c;
    one_scene(&c, to_dev(*fluid), blocks_for(fluid->n));
    c.a.one.dt = dt;
    return launch_integrate<false>(ctx, c, ax, ay, az, nullptr, 1);
}

int reyn_integrate_a_v_only(reyn_ctx* ctx, const reyn_particles* fluid, const float* ax,
    const float* ay, const float* az, float dt)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_TRY(check_fluid(ctx, fluid));
    if (fluid->n == 0)
        return REYN_OK;
    Common c;
    one_scene(&c, to_dev(*fluid), blocks_for(fluid->n));
    c.a.one.dt = dt;
    return launch_integrate<false>(ctx, c, ax, ay, az, nullptr, 0);
}

int reyn_sesph_step(reyn_ctx* ctx, const reyn_sesph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, const reyn_sesph_work* work, float dt)
{
    if (!prm || !work)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "sesph: NULL parameters or work buffers");
    Common c;
    REYN_TRY(prepare(ctx, &prm->sph, fluid, grid, bdy, &c));
    if (fluid->n == 0)
        return REYN_OK;
    if (!work->rho || !work->ax || !work->ay || !work->az)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "sesph: NULL work buffer");
    c.a.one.dt = dt;
    Scratch s;
    REYN_TRY(get_scratch(ctx, fluid->n, &s, true, 0, 1, c.a.one.b.present ? c.a.one.b.n : 0));
    return sesph_run<false>(ctx, prm->sph.kernel, c, prm->k, work, s);
}

int reyn_sesph_step_noalloc(reyn_ctx* ctx, const reyn_sesph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, const reyn_sesph_work* work, float dt)
{
    if (!prm || !work)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "sesph (no alloc): NULL parameters or work buffers");
    Common c;
    REYN_TRY(prepare(ctx, &prm->sph, fluid, grid, bdy, &c));
    if (fluid->n > kNoAllocMaxFluid)
        return set_error(ctx, REYN_ERR_CAPACITY,
            "sesph (no alloc): %u fluid particles exceed the %u its static scratch holds", fluid->n,
            kNoAllocMaxFluid);
    if (fluid->n == 0)
        return REYN_OK;
    if (!work->rho || !work->ax || !work->ay || !work->az)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "sesph (no alloc): NULL work buffer");
    c.a.one.dt = dt;
    char* p = nullptr;
    REYN_TRY(noalloc_scratch(ctx, &p));
    // the walk form: boundary sums and loop control only (no list, no packed
    // records), within the static scratch
    Scratch s;
    if (carve_scratch(nullptr, fluid->n, &s, false, 0, 1, 0) > kNoAllocScratchBytes)
        return set_error(ctx, REYN_ERR_CAPACITY, "sesph (no alloc): scratch layout exceeds the static scratch");
    carve_scratch(p, fluid->n, &s, false, 0, 1, 0);
    return sesph_run<false>(ctx, prm->sph.kernel, c, prm->k, work, s);
}

int reyn_pcisph_delta(int32_t kernel, float h_bar, float h, double* delta)
{
    if (!delta || kernel < 0 || kernel > 3 || !(h_bar > 0.f) || !(h > 0.f))
        return REYN_ERR_INVALID_ARGUMENT;
    // template particle on a 5^3 lattice of spacing h (PCISPH.cuh:49-77); the
    // gradient is evaluated in float on the host as in the reference
    const KParams kp = make_kparams(kernel, h_bar);
    float sx = 0.f, sy = 0.f, sz = 0.f;
    double sum_sq = 0.0;
    const int kappa = 2;
    for (int x = -kappa; x <= kappa; ++x)
        for (int y = -kappa; y <= kappa; ++y)
            for (int z = -kappa; z <= kappa; ++z) {
                const float dx = -(static_cast<float>(x) * h), dy = -(static_cast<float>(y) * h),
                            dz = -(static_cast<float>(z) * h);
                const float d2 = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
                const float len = std::sqrt(d2);
                const float q = len * kp.hinv;
                float dw;
                const float t1 = std::fmax(0.f, 1.f - q);
                switch (kernel) {
                case KC2:
                    dw = t1 * t1 * t1 * (-20.f * q);
                    break;
                case KB3: {
                    const float t2 = std::fmin(std::fmax(0.5f - q, 0.f), 1.f);
                    dw = -3.f * t1 * t1 + 12.f * t2 * t2;
                    break;
                }
                case KW6: {
                    const float t2 = t1 * t1;
                    const float t7 = t2 * t2 * t2 * t1;
                    dw = t7 * ((16.f * q * q + 7.f * q + 1.f) * (-22.f * q));
                    break;
                }
                default: {
                    const float pis = 3.14159265358979323846f * std::fmin(1.f, q);
                    const float two_pi = 6.28318530717958647692f;
                    dw = -2.f * two_pi * std::sin(pis) - std::sin(2.f * pis) * two_pi;
                    break;
                }
                }
                float scale = kp.dws * dw / len;
                if (!std::isfinite(scale))
                    scale = 0.f;
                const float gx = scale * dx, gy = scale * dy, gz = scale * dz;
                sx += gx;
                sy += gy;
                sz += gz;
                sum_sq += static_cast<double>(std::fma(gx, gx, std::fma(gy, gy, gz * gz)));
            }
    const double sq_of_sum = static_cast<double>(std::fma(sx, sx, std::fma(sy, sy, sz * sz)));
    const double V0 = double(h) * double(h) * double(h);
    const double beta = 2.0 * V0 * V0;
    *delta = static_cast<double>(static_cast<float>(1.0 / (beta * (sq_of_sum + sum_sq))));
    return REYN_OK;
}

int reyn_pcisph_step(reyn_ctx* ctx, const reyn_pcisph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, const reyn_pcisph_work* work, float dt,
    reyn_solve_info* info)
{
    if (!prm || !work)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "pcisph: NULL parameters or work buffers");
    Common c;
    REYN_TRY(prepare(ctx, &prm->sph, fluid, grid, bdy, &c));
    if (info)
        *info = reyn_solve_info {};
    if (fluid->n == 0)
        return REYN_OK;
    if (!work->rho || !work->ax || !work->ay || !work->az)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "pcisph: NULL work buffer");
    if (!(dt > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "pcisph: dt must be positive");
    c.a.one.dt = dt;
    c.a.one.kk = static_cast<float>((prm->delta / dt) / dt);
    c.a.one.thr = (prm->eta_rho_max + 1.f) * prm->sph.rho_0;
    Scratch s;
    REYN_TRY(get_scratch(ctx, fluid->n, &s, true, 0, 1, c.a.one.b.present ? c.a.one.b.n : 0));
    SolveCtl res;
    REYN_TRY(pcisph_run<false>(ctx, prm->sph.kernel, c, work, s, 1, prm->min_iter, prm->max_iter, prm->check_every,
        &res));
    if (info) {
        info->iterations = res.iter;
        info->avg_density = res.avg_density;
    }
    return REYN_OK;
}

int reyn_iisph_step(reyn_ctx* ctx, const reyn_iisph_params* prm, const reyn_particles* fluid,
    const reyn_grid* grid, const reyn_boundary* bdy, const reyn_iisph_work* work, float dt,
    reyn_solve_info* info)
{
    if (!prm || !work)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "iisph: NULL parameters or work buffers");
    Common c;
    REYN_TRY(prepare(ctx, &prm->sph, fluid, grid, bdy, &c));
    if (info)
        *info = reyn_solve_info {};
    if (fluid->n == 0)
        return REYN_OK;
    if (!work->rho || !work->p || !work->a_ii || !work->s_i || !work->ax || !work->ay || !work->az)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "iisph: NULL work buffer");
    if (!(dt > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "iisph: dt must be positive");
    if (prm->stop_rule != REYN_IISPH_STOP_ANY && prm->stop_rule != REYN_IISPH_STOP_MEAN)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "iisph: unknown stop rule %d", prm->stop_rule);
    c.a.one.dt = dt;
    c.a.one.thr = prm->eta_rho_max * prm->sph.rho_0;
    Scratch s;
    REYN_TRY(get_scratch(ctx, fluid->n, &s, true, 0, 1, c.a.one.b.present ? c.a.one.b.n : 0));
    const IisphLoop lp { prm->omega, prm->eta_rho_max, prm->min_iter, prm->max_iter, prm->check_every,
        prm->stop_rule, prm->warm_start };
    SolveCtl res;
    REYN_TRY(iisph_run<false>(ctx, prm->sph.kernel, c, lp, work, s, 1, &res));
    if (info) {
        info->iterations = res.iter;
        info->n_violating = res.n_violating;
        info->mean_rel_error = res.mean_rel_error;
    }
    return REYN_OK;
}

} // extern "C"
