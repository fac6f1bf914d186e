// Device-side building blocks of every SPH phase: smoothing kernels, the
// grid view and the fixed-radius neighbour walk.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace reyn {

// ---- smoothing kernels (src/kernels.cuh:19-174) ---------------------------
// W(x) = alpha/h^3 w(sat(|x|/h)), grad W(x) = alpha/h^4 dw(|x|/h)/|x| x, with
// the gradient forced to 0 when the scale is not finite (|x| = 0). alpha is
// formed in double and rounded to float as in src/kernels.cuh:32-37.

enum : int { KC2 = REYN_KERNEL_C2, KB3 = REYN_KERNEL_B3, KW6 = REYN_KERNEL_W6, KCOS = REYN_KERNEL_COS };

struct KParams {
    float hinv; // 1/h_bar
    float ws; // alpha / h_bar^3
    float dws; // alpha / h_bar^4
};

inline double kernel_alpha(int kind)
{
    constexpr double one_over_pi = 0.318309886183790671537767526745028724;
    constexpr double pi = 3.14159265358979323846264338327950288;
    switch (kind) {
    case KC2:
        return 10.5 * one_over_pi; // 21/(2 pi)
    case KB3:
        return 16.0 * one_over_pi; // 16/pi
    case KW6:
        return 21.328125 * one_over_pi; // 1365/(64 pi)
    case KCOS:
        return pi / (4.0 * pi * pi - 30.0);
    default:
        return 0.0;
    }
}

inline KParams make_kparams(int kind, float h_bar)
{
    const float alpha = static_cast<float>(kernel_alpha(kind));
    KParams p;
    p.hinv = 1.0f / h_bar;
    p.ws = alpha / (h_bar * h_bar * h_bar);
    p.dws = alpha / (h_bar * h_bar * h_bar * h_bar);
    return p;
}

__device__ __forceinline__ float sat(float a) { return fminf(fmaxf(a, 0.f), 1.f); }

template <int K> struct Shape;

template <> struct Shape<KC2> {
    __device__ __forceinline__ static float w(float q)
    {
        const float t1 = fmaxf(0.f, 1.f - q);
        const float t12 = t1 * t1;
        return t12 * t12 * (4.f * q + 1.f);
    }
    __device__ __forceinline__ static float dw(float q)
    {
        const float t1 = fmaxf(0.f, 1.f - q);
        return t1 * t1 * t1 * (-20.f * q);
    }
};
template <> struct Shape<KB3> {
    __device__ __forceinline__ static float w(float q)
    {
        const float t1 = fmaxf(0.f, 1.f - q);
        const float t2 = sat(0.5f - q);
        return t1 * t1 * t1 - 4.f * t2 * t2 * t2;
    }
    __device__ __forceinline__ static float dw(float q)
    {
        const float t1 = fmaxf(0.f, 1.f - q);
        const float t2 = sat(0.5f - q);
        return -3.f * t1 * t1 + 12.f * t2 * t2;
    }
};
template <> struct Shape<KW6> {
    __device__ __forceinline__ static float w(float q)
    {
        const float t1 = fmaxf(0.f, 1.f - q);
        const float t2 = t1 * t1;
        const float t4 = t2 * t2;
        const float q2 = q * q;
        return (t4 * t4) * (32.f * q2 * q + 25.f * q2 + 8.f * q + 1.f);
    }
    __device__ __forceinline__ static float dw(float q)
    {
        const float t1 = fmaxf(0.f, 1.f - q);
        const float t2 = t1 * t1;
        const float t7 = t2 * t2 * t2 * t1;
        return t7 * ((16.f * q * q + 7.f * q + 1.f) * (-22.f * q));
    }
};
template <> struct Shape<KCOS> {
    __device__ __forceinline__ static float w(float q)
    {
        const float pis = 3.14159265358979323846f * fminf(1.f, q);
        return 4.f * cosf(pis) + cosf(2.f * pis) + 3.f;
    }
    __device__ __forceinline__ static float dw(float q)
    {
        const float pis = 3.14159265358979323846f * fminf(1.f, q);
        const float two_pi = 6.28318530717958647692f;
        return -2.f * two_pi * sinf(pis) - sinf(2.f * pis) * two_pi;
    }
};

// W from the squared distance (the neighbour walk already formed it). |x| and
// 1/|x| come from one hardware reciprocal square root (v_rsq_f32, about 1 ulp)
// instead of a correctly rounded sqrt and divide; the results stay within the
// FP32 rounding the solver parity bar allows.
template <int K> __device__ __forceinline__ float kW(const KParams& p, float d2)
{
    float len = d2 * __builtin_amdgcn_rsqf(d2);
    // not finite: d2 = 0 or below the normal range (the hardware rsqrt flushes
    // it to 0) gives |x| ~ 0; d2 = inf gives |x| = inf
    if (!(len <= 3.0e38f))
        len = d2 > 1.f ? d2 : 0.f;
    return p.ws * Shape<K>::w(sat(len * p.hinv));
}

// grad W from the offset and its squared length; the zero vector at x = 0
// (and wherever the scale is not finite), as in the reference.
template <int K> __device__ __forceinline__ f3 kdW(const KParams& p, f3 x, float d2)
{
    const float inv = __builtin_amdgcn_rsqf(d2);
    const float scale = p.dws * Shape<K>::dw(d2 * inv * p.hinv) * inv;
    return ((d2 > 0.f && __builtin_isfinite(scale)) ? scale : 0.f) * x;
}

// ---- grid view and neighbour walk -----------------------------------------
// (src/datastructure/uniformgrid.cuh:17-64, 113-178)

struct GridDev {
    f3 bmin;
    float cell;
    float r2;
    int nx, ny, nz, nxny;
    int range;
    int order; // REYN_CELL_ORDER_*
    int nbx, nbxy; // bricks per row and per plane (REYN_CELL_ORDER_BRICK)
    const uint32_t* __restrict__ prefix;
    const uint32_t* __restrict__ sorted;
};

inline GridDev to_dev(const reyn_grid& g)
{
    GridDev d;
    d.bmin = { g.bound_min.x, g.bound_min.y, g.bound_min.z };
    d.cell = g.cell_size;
    d.r2 = g.r_c_2;
    d.nx = g.nx;
    d.ny = g.ny;
    d.nz = g.nz;
    d.nxny = g.nx * g.ny;
    d.range = g.range;
    d.order = g.order;
    d.nbx = g.nx / 2;
    d.nbxy = (g.nx / 2) * (g.ny / 2);
    d.prefix = g.prefix;
    d.sorted = g.sorted;
    return d;
}

// floor((x - bound_min) / cell) per axis with IEEE division, clamped to
// [0, n-1]; NaN maps to 0. Inside the grid this is the reference's `_index3`
// (uniformgrid.cuh:17-22, vector.cuh:137-141); the clamp only changes points
// the reference would index out of bounds.
__device__ __forceinline__ int cell_axis(float x, float lo, float cell, int n)
{
    float f = floorf((x - lo) / cell);
    f = fmaxf(f, 0.f); // NaN -> 0
    f = fminf(f, static_cast<float>(n - 1));
    return static_cast<int>(f);
}
struct int3c {
    int x, y, z;
};
__device__ __forceinline__ int3c cell_of(const GridDev& g, f3 x)
{
    return { cell_axis(x.x, g.bmin.x, g.cell, g.nx), cell_axis(x.y, g.bmin.y, g.cell, g.ny),
        cell_axis(x.z, g.bmin.z, g.cell, g.nz) };
}

// The first key of brick (bx, by, bz) (REYN_CELL_ORDER_BRICK).
__device__ __forceinline__ uint32_t brick_base(const GridDev& g, int bx, int by, int bz)
{
    return (static_cast<uint32_t>(bz) * static_cast<uint32_t>(g.nbxy) + static_cast<uint32_t>(by * g.nbx + bx)) << 3;
}
__device__ __forceinline__ uint32_t cell_key(const GridDev& g, int3c c)
{
    if (g.order == REYN_CELL_ORDER_BRICK)
        return brick_base(g, c.x >> 1, c.y >> 1, c.z >> 1) + static_cast<uint32_t>(((c.z & 1) << 2) | ((c.y & 1) << 1) | (c.x & 1));
    return static_cast<uint32_t>(c.z) * static_cast<uint32_t>(g.nxny) + static_cast<uint32_t>(c.y * g.nx + c.x);
}

// Candidates one walk visits at most. A grid build fails with
// REYN_ERR_CAPACITY when a cell holds more than walk_cell_cap(range) points,
// so a walk over a grid that built without error (at most (2 range + 1)^3
// cells of at most that many points) never reaches the budget; the clip in
// `strips` only keeps walks over a hand-made grid bounded.
constexpr uint32_t kWalkBudget = 1u << 15;
__host__ __device__ constexpr uint32_t walk_cell_cap(int range)
{
    return kWalkBudget / uint32_t((2 * range + 1) * (2 * range + 1) * (2 * range + 1));
}

// Calls r(s, e, plane, base) for the candidate range [s, e) (cell order) of
// each of the (2R+1)^2 x-strips of 2R+1 cells around the cell of xi, clipped to
// the grid, in the order z, then y ascending; `plane` is the strip's z-plane
// among the 2R+1 (0 for z - R) and, for R <= 2, `base` the plane's first
// candidate position, prefix[clamped z, max(y - R, 0), max(x - R, 0)] (no
// candidate of the plane lies below it; 0 for larger R). Every lane makes all
// (2R+1)^2 calls, with an empty range for a strip outside the grid, so a call
// is a point the whole wavefront reaches together. Clipping leaves every point
// whose cell lies R or more cells inside the grid unchanged. For R <= 2 the
// bounds of the strips of one z-plane are loaded together (from clamped, valid
// rows), so a walk waits for 2R+1 round trips of cell-start loads instead of
// (2R+1)^2.
template <class Rg> __device__ __forceinline__ void strips(const GridDev& g, f3 xi, Rg&& r)
{
    const int3c c = cell_of(g, xi);
    const int R = g.range;
    const int x0 = max(c.x - R, 0), x1 = min(c.x + R, g.nx - 1);
    // candidate budget: unreachable over a grid that built without error
    // (walk_cell_cap); it keeps a walk over any other grid bounded
    uint32_t budget = kWalkBudget;
    auto emit = [&](uint32_t s, uint32_t e, uint32_t plane, uint32_t base) {
        if (e - s > budget)
            e = s + budget;
        budget -= e - s;
        r(s, e, plane, base);
    };
    if (g.order == REYN_CELL_ORDER_BRICK) {
        // A strip's 2R+1 cells lie in R+1 bricks; the cells of one strip in
        // one brick have consecutive keys, so each brick is one range (empty
        // where clipped): (2R+1)^2 (R+1) calls. For R = 1 the strips fall in
        // 4 segments, (brick plane, brick row) with z, y each the lower or
        // upper of the two bricks the 3 cells span; `plane` is the segment and
        // `base` the first candidate of its lower brick in x, below every
        // candidate of the segment.
        const int bz0 = max(c.z - R, 0) >> 1, by0 = max(c.y - R, 0) >> 1, bx0 = x0 >> 1;
        for (int dz = -R; dz <= R; ++dz) {
            const int z = c.z + dz;
            const int zc = min(max(z, 0), g.nz - 1);
            for (int dy = -R; dy <= R; ++dy) {
                const int y = c.y + dy;
                const int yc = min(max(y, 0), g.ny - 1);
                const bool in = z >= 0 && z < g.nz && y >= 0 && y < g.ny;
                uint32_t seg = 0, sbase = 0;
                if (R == 1) {
                    seg = static_cast<uint32_t>((((zc >> 1) - bz0) << 1) | ((yc >> 1) - by0));
                    sbase = gat(g.prefix, brick_base(g, bx0, yc >> 1, zc >> 1));
                }
                int xs = x0;
                for (int k = 0; k <= R; ++k) {
                    const int xe = min(xs | 1, x1);
                    uint32_t s0 = 0, e0 = 0;
                    if (in && xs <= x1) {
                        s0 = gat(g.prefix, cell_key(g, int3c { xs, yc, zc }));
                        e0 = gat(g.prefix, cell_key(g, int3c { xe, yc, zc }) + 1);
                    }
                    emit(s0, e0, seg, sbase);
                    xs = (xs | 1) + 1;
                }
            }
        }
        return;
    }
    if (R <= 2) {
        for (int dz = -R; dz <= R; ++dz) {
            const int z = c.z + dz;
            const int zc = min(max(z, 0), g.nz - 1);
            uint32_t s[5], e[5];
#pragma unroll
            for (int u = 0; u < 5; ++u) {
                if (u <= 2 * R) {
                    const int yc = min(max(c.y - R + u, 0), g.ny - 1);
                    const uint32_t base = static_cast<uint32_t>(zc * g.nxny + yc * g.nx);
                    s[u] = gat(g.prefix, base + x0);
                    e[u] = gat(g.prefix, base + x1 + 1);
                }
            }
#pragma unroll
            for (int u = 0; u < 5; ++u) {
                if (u <= 2 * R) {
                    const int y = c.y - R + u;
                    const bool in = z >= 0 && z < g.nz && y >= 0 && y < g.ny;
                    emit(in ? s[u] : 0u, in ? e[u] : 0u, uint32_t(dz + R), s[0]);
                }
            }
        }
        return;
    }
    for (int dz = -R; dz <= R; ++dz) {
        for (int dy = -R; dy <= R; ++dy) {
            const int z = c.z + dz, y = c.y + dy;
            uint32_t s = 0, e = 0;
            if (z >= 0 && z < g.nz && y >= 0 && y < g.ny) {
                const uint32_t base = static_cast<uint32_t>(z * g.nxny + y * g.nx);
                s = gat(g.prefix, base + x0);
                e = gat(g.prefix, base + x1 + 1);
            }
            emit(s, e, uint32_t(dz + R), 0u);
        }
    }
}

// Candidates of one strip are loaded kWalkBatch at a time, all in flight
// before the first test (past the end of the strip the last candidate is
// loaded again and not tested).
constexpr uint32_t kWalkBatch = 4;

// Calls f(j, x_ij, |x_ij|^2) for every indexed point j with
// |x_ij|^2 = fmaf-dot <= r_c_2 (the point itself included), in walk order
// (src/datastructure/uniformgrid.cuh:113-178).
template <bool IND, class F>
__device__ __forceinline__ void walk(const GridDev& g, f3 xi, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, F&& f)
{
    strips(g, xi, [&](uint32_t s, uint32_t e, uint32_t, uint32_t) {
        constexpr uint32_t B = kWalkBatch;
        for (uint32_t q = s; q < e; q += B) {
            uint32_t j[B];
            f3 pj[B];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) {
                const uint32_t qq = min(q + u, e - 1);
                j[u] = IND ? gat(g.sorted, qq) : qq;
            }
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                pj[u] = gat3(xs, ys, zs, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) {
                if (q + u < e) {
                    const f3 xij = xi - pj[u];
                    const float d2 = dot(xij, xij);
                    if (d2 <= g.r2)
                        f(j[u], xij, d2);
                }
            }
        }
    });
}

// Where a candidate sits: its position q in cell order, the z-plane of its
// strip and that plane's base (see `strips`).
struct Cand {
    uint32_t q, plane, base;
};

// The same walk over packed records P[j] = {x, y, z, m}; f(j, x_ij, d2, m_j,
// cand). `strip_end()` runs after each strip, at a point every lane of the
// wavefront that entered the walk reaches together.
template <bool IND, class F, class E>
__device__ __forceinline__ void walk4(const GridDev& g, f3 xi, const float4* __restrict__ P, F&& f,
    E&& strip_end)
{
    strips(g, xi, [&](uint32_t s, uint32_t e, uint32_t plane, uint32_t base) {
        constexpr uint32_t B = kWalkBatch;
        for (uint32_t q = s; q < e; q += B) {
            uint32_t j[B];
            float4 pj[B];
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) {
                const uint32_t qq = min(q + u, e - 1);
                j[u] = IND ? gat(g.sorted, qq) : qq;
            }
#pragma unroll
            for (uint32_t u = 0; u < B; ++u)
                pj[u] = gat(P, j[u]);
#pragma unroll
            for (uint32_t u = 0; u < B; ++u) {
                if (q + u < e) {
                    const f3 xij = xi - f3 { pj[u].x, pj[u].y, pj[u].z };
                    const float d2 = dot(xij, xij);
                    if (d2 <= g.r2)
                        f(j[u], xij, d2, pj[u].w, Cand { q + u, plane, base });
                }
            }
        }
        strip_end();
    });
}
template <bool IND, class F>
__device__ __forceinline__ void walk4(const GridDev& g, f3 xi, const float4* __restrict__ P, F&& f)
{
    walk4<IND>(g, xi, P, f, [] {});
}

// LDS stores of some lanes made visible to the other lanes of the wavefront.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Strip staging for walks over a cell-ordered state (no index indirection).
// The lanes of a wavefront sit in neighbouring cells, so for one strip offset
// their candidate ranges [s, e) fall into one short index window, or into two
// when the wavefront spans two rows of cells. With WIN windows of up to CAP
// records each that hold every lane's range, the wavefront loads the windows
// once, with coalesced loads, into its own WIN * CAP slots of LDS `st`, and
// each lane reads candidate q from st[q - base]; otherwise (`staged` false,
// the same for the whole wavefront) every lane loads from memory as before.
// One window costs fewer reductions and less LDS; two also cover the
// wavefronts that cross a row.
struct StageWin {
    uint32_t base;
    bool staged;
};
template <uint32_t CAP, int WIN, class Ld>
__device__ __forceinline__ StageWin stage_strip(uint32_t s, uint32_t e, float4* __restrict__ st, Ld&& load)
{
    const uint32_t lane = threadIdx.x % kWave;
    const bool ne = e > s;
    const uint32_t lo1 = wave_min_u32(ne ? s : 0xffffffffu);
    if constexpr (WIN == 1) {
        const uint32_t hi = wave_max_u32(ne ? e : 0u);
        if (hi <= lo1 || hi - lo1 > CAP) // uniform
            return { 0u, false };
        for (uint32_t q = lo1 + lane; q < hi; q += kWave)
            st[q - lo1] = load(q);
        wave_lds_sync();
        return { lo1, true };
    }
    const bool in1 = ne && e - lo1 <= CAP;
    const uint32_t hi1 = wave_max_u32(in1 ? e : 0u);
    const uint32_t lo2 = wave_min_u32(ne && !in1 ? s : 0xffffffffu);
    const bool in2 = ne && !in1 && e - lo2 <= CAP;
    const uint32_t hi2 = wave_max_u32(in2 ? e : 0u);
    const bool all = __ballot(ne && !in1 && !in2) == 0;
    if (!all || hi1 <= lo1) // uniform
        return { 0u, false };
    for (uint32_t q = lo1 + lane; q < hi1; q += kWave)
        st[q - lo1] = load(q);
    for (uint32_t q = lo2 + lane; q < hi2; q += kWave)
        st[CAP + (q - lo2)] = load(q);
    wave_lds_sync();
    // q - base in unsigned arithmetic: q - lo1, or CAP + q - lo2
    return { in1 ? lo1 : lo2 - CAP, true };
}

// Calls f(q, x_ij, |x_ij|^2, p_q) for the candidates q in [s, e) within
// reach, p_q = get(q) loaded kWalkBatch ahead (past the end the last one again).
template <class G, class F>
__device__ __forceinline__ void strip_loop(const GridDev& g, f3 xi, uint32_t s, uint32_t e, G&& get, F&& f)
{
    constexpr uint32_t B = kWalkBatch;
    for (uint32_t q = s; q < e; q += B) {
        float4 pj[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
            pj[u] = get(min(q + u, e - 1));
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            if (q + u < e) {
                const f3 xij = xi - f3 { pj[u].x, pj[u].y, pj[u].z };
                const float d2 = dot(xij, xij);
                if (d2 <= g.r2)
                    f(q + u, xij, d2, pj[u]);
            }
        }
    }
}

// strip_loop over staged records, p = &st[s - base]: whole batches of
// kWalkBatch with no bounds tests and LDS addresses as immediate offsets from
// one pointer, then the tail one by one (the same candidates in the same
// order).
template <class F>
__device__ __forceinline__ void strip_loop_lds(const GridDev& g, f3 xi, uint32_t s, uint32_t e,
    const float4* __restrict__ p, F&& f)
{
    constexpr uint32_t B = kWalkBatch;
    uint32_t q = s;
    for (; q + B <= e; q += B, p += B) {
        float4 pj[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
            pj[u] = p[u];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            const f3 xij = xi - f3 { pj[u].x, pj[u].y, pj[u].z };
            const float d2 = dot(xij, xij);
            if (d2 <= g.r2)
                f(q + u, xij, d2, pj[u]);
        }
    }
    for (; q < e; ++q, ++p) {
        const float4 pj = *p;
        const f3 xij = xi - f3 { pj.x, pj.y, pj.z };
        const float d2 = dot(xij, xij);
        if (d2 <= g.r2)
            f(q, xij, d2, pj);
    }
}

// walk over SoA positions of a cell-ordered state with every strip staged
// (stage_strip; `st` holds WIN * CAP records of this wavefront), staged as
// {x, y, z, 0}; f(j, x_ij, |x_ij|^2, x_j). Every lane of the wavefront must
// take part; lanes with `active` false visit nothing. The candidates and their
// order are walk's.
template <uint32_t CAP, int WIN, class F>
__device__ __forceinline__ void walk_staged(const GridDev& g, f3 xi, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, float4* __restrict__ st, bool active, F&& f)
{
    strips(g, xi, [&](uint32_t s, uint32_t e, uint32_t, uint32_t) {
        if (!active)
            e = s;
        const StageWin w
            = stage_strip<CAP, WIN>(s, e, st, [&](uint32_t q) { return make_float4(xs[q], ys[q], zs[q], 0.f); });
        auto body = [&](uint32_t q, f3 xij, float d2, float4 p) { f(q, xij, d2, f3 { p.x, p.y, p.z }); };
        if (w.staged) {
            strip_loop_lds(g, xi, s, e, st + (s - w.base), body);
            wave_lds_sync();
        } else {
            strip_loop(g, xi, s, e,
                [&](uint32_t q) {
                    const f3 p = gat3(xs, ys, zs, q);
                    return make_float4(p.x, p.y, p.z, 0.f);
                },
                body);
        }
    });
}

// ---- per-step neighbour lists ------------------------------------------------
// The density pass of a solver step walks the grid once and records, for the
// particle in cell-order slot t, its fluid neighbours j != self in walk order,
// at most `cap` of them. An entry takes 16 bits: the z-plane u of the
// neighbour's strip among the walk's 2R+1 planes (3 bits) and the neighbour's
// cell-order position q minus the plane's base (13 bits; `strips`). A loop
// that reads the list recomputes the bases from the particle's cell, so an
// entry e decodes as q = e + B_u with B_u = base_u - (u << 13) (unsigned).
// Entries k and k+1 (k even) share the 32-bit word [(k/2) n + t]: the rows
// are column-major, so a wavefront reads one contiguous 256-byte row per two
// entries. A particle with more than `cap` neighbours, or with a neighbour
// 2^13 or more candidates above its plane's base, is stored with a count above
// `cap` and walks the grid in every later loop; lists are kept for ranges up
// to kListMaxRange (cell ratios from 0.5 up).

constexpr uint32_t kPlaneShift = 13;
constexpr uint32_t kPlaneSpan = 1u << kPlaneShift; // offsets from a plane's base below this
constexpr int kListMaxRange = 2;

struct NbrDev {
    uint32_t* cnt; // NULL: no list, every loop walks the grid
    uint32_t* idx; // 16-bit entries in word pairs, see above
    uint32_t n, cap;
};

// Stores entry k of slot t.
__device__ __forceinline__ void put_entry(const NbrDev& L, uint32_t t, uint32_t k, uint32_t e)
{
    reinterpret_cast<uint16_t*>(L.idx)[(size_t(k >> 1) * L.n + t) * 2 + (k & 1)] = static_cast<uint16_t>(e);
}

// Stores entries k and k+1 (k even) of slot t: e01 = e_k | e_{k+1} << 16.
__device__ __forceinline__ void put_pair(const NbrDev& L, uint32_t t, uint32_t k, uint32_t e01)
{
    L.idx[size_t(k >> 1) * L.n + t] = e01;
}

// The entry of a candidate of the walk, and whether its offset fits.
__device__ __forceinline__ uint32_t encode_entry(const Cand& c, bool* fits)
{
    const uint32_t off = c.q - c.base;
    *fits = off < kPlaneSpan;
    return off + (c.plane << kPlaneShift);
}

// The decode bases B_u = base_u - (u << kPlaneShift) of the particle at xi
// for its NP = 2R+1 planes (XYZ order) or its 4 segments (brick order, range
// 1), from the same cell-start entries the walk loads.
template <int NP> __device__ __forceinline__ void plane_bases(const GridDev& g, f3 xi, uint32_t (&B)[NP])
{
    if constexpr (NP == 4) {
        const int3c c = cell_of(g, xi);
        const int bz0 = max(c.z - 1, 0) >> 1, by0 = max(c.y - 1, 0) >> 1, bx0 = max(c.x - 1, 0) >> 1;
        const int bz1 = min(c.z + 1, g.nz - 1) >> 1, by1 = min(c.y + 1, g.ny - 1) >> 1;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int bz = (u >> 1) ? bz1 : bz0, by = (u & 1) ? by1 : by0;
            B[u] = gat(g.prefix, brick_base(g, bx0, by, bz)) - (uint32_t(u) << kPlaneShift);
        }
        return;
    }
    constexpr int R = (NP - 1) / 2;
    const int3c c = cell_of(g, xi);
    const int x0 = max(c.x - R, 0);
    const int yc = min(max(c.y - R, 0), g.ny - 1);
#pragma unroll
    for (int u = 0; u < NP; ++u) {
        const int zc = min(max(c.z - R + u, 0), g.nz - 1);
        const uint32_t row = static_cast<uint32_t>(zc * g.nxny + yc * g.nx);
        B[u] = gat(g.prefix, row + x0) - (uint32_t(u) << kPlaneShift);
    }
}

// q of entry e: e + B_u for the plane u in e's top bits. B_u is formed as B_0
// plus the steps B_v - B_{v-1} of every plane v <= u (unsigned, so the sum
// is B_u exactly): written as a chain of selects the compiler turns the
// choice into an indexed read of B, which it then keeps in LDS, adding an
// LDS round trip between every list word and the record it names.
template <int NP> __device__ __forceinline__ uint32_t decode_entry(const uint32_t (&B)[NP], uint32_t e)
{
    uint32_t b = B[0];
#pragma unroll
    for (int u = 1; u < NP; ++u)
        b += (e >= (uint32_t(u) << kPlaneShift)) ? B[u] - B[u - 1] : 0u;
    return e + b;
}

// Candidates of a staged strip, p = &st[s - base]: calls hit(q + eoff) for
// every candidate q within reach, except the lane's own point `self` when
// SELF (the strip may hold it); whole batches of kWalkBatch from one LDS
// pointer, then the tail one by one (the candidates and their order are
// strip_loop_lds's).
template <bool SELF, class H>
__device__ __forceinline__ void strip_hits_lds(const GridDev& g, f3 xi, uint32_t s, uint32_t e,
    const float4* __restrict__ p, uint32_t self, uint32_t eoff, H&& hit)
{
    constexpr uint32_t B = kWalkBatch;
    uint32_t q = s;
    for (; q + B <= e; q += B, p += B) {
        float4 pj[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
            pj[u] = p[u];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            const f3 xij = xi - f3 { pj[u].x, pj[u].y, pj[u].z };
            if (dot(xij, xij) <= g.r2 && (!SELF || q + u != self))
                hit(q + u + eoff);
        }
    }
    for (; q < e; ++q, ++p) {
        const float4 pj = *p;
        const f3 xij = xi - f3 { pj.x, pj.y, pj.z };
        if (dot(xij, xij) <= g.r2 && (!SELF || q != self))
            hit(q + eoff);
    }
}

// The listing walk of the density pass over a cell-ordered state (either order),
// every strip staged (stage_strip; `st` holds CAP records of this
// wavefront): calls hit(e) with the list entry e (`encode_entry`) of every
// candidate within reach other than the lane's own point `self`, in walk
// order, and strip_end() after each strip, a point every lane of the
// wavefront reaches together. Lanes with `active` false visit nothing.
// `fits` turns false when a strip's last candidate lies kPlaneSpan or more
// above its plane's base: a test once per strip instead of once per
// neighbour, on a candidate at or above every neighbour of the strip (a
// particle is then at worst left unlisted where only a candidate out of reach
// would not fit). The self test runs only in the strip that holds a lane's
// own point (the strip order is the same for every lane, so the choice is
// uniform). The neighbours and their order are walk4's.
template <uint32_t CAP, class H, class E>
__device__ __forceinline__ void list_walk_staged(const GridDev& g, f3 xi, const float4* __restrict__ P,
    float4* __restrict__ st, bool active, uint32_t self, bool& fits, H&& hit, E&& strip_end)
{
    strips(g, xi, [&](uint32_t s, uint32_t e, uint32_t plane, uint32_t base) {
        if (!active)
            e = s;
        fits = fits && (e <= s || e - 1u - base < kPlaneSpan);
        const uint32_t eoff = (plane << kPlaneShift) - base;
        const bool mine = self - s < e - s;
        const StageWin w = stage_strip<CAP, 1>(s, e, st, [&](uint32_t q) { return P[q]; });
        if (w.staged) {
            const float4* p = st + (s - w.base);
            if (__ballot(mine) != 0) // uniform
                strip_hits_lds<true>(g, xi, s, e, p, self, eoff, hit);
            else
                strip_hits_lds<false>(g, xi, s, e, p, self, eoff, hit);
            wave_lds_sync(); // every lane has read before the next strip is staged
        } else {
            strip_loop(g, xi, s, e, [&](uint32_t q) { return gat(P, q); },
                [&](uint32_t q, f3, float, float4) {
                    if (q != self)
                        hit(q + eoff);
                });
        }
        strip_end();
    });
}

// Reads the list of slot t (count c <= cap, particle at xi) in rounds of B
// entries (B even) and calls body(k, q) with the round's first entry k and the
// cell-order positions q[B]. Rounds run past the count with entry k again
// (always written), so every q is valid; entry k + u is real when k + u < c.
// The index words, the decodes and whatever the body gathers are all in
// flight together within a round.
template <int NP, uint32_t B, class Body>
__device__ __forceinline__ void list_rounds(const GridDev& g, const NbrDev& L, uint32_t t, uint32_t c, f3 xi,
    Body&& body)
{
    static_assert(B % 2 == 0, "list rounds read entries in word pairs");
    uint32_t bases[NP];
    plane_bases<NP>(g, xi, bases);
    uint32_t off = t;
    for (uint32_t k = 0; k < c; k += B, off += (B / 2) * L.n) {
        uint32_t w[B / 2];
        // the words are streamed once per pass: non-temporal loads keep them
        // from pushing the gathered neighbour records out of the caches
        // (about 5% faster Jacobi and pressure-acceleration passes at 2M)
#pragma unroll
        for (uint32_t v = 0; v < B / 2; ++v) // past the count: the row of entry k (valid)
            w[v] = gat_nt(L.idx, (k + 2 * v < c) ? off + v * L.n : off);
        uint32_t q[B];
#pragma unroll
        for (uint32_t u = 0; u < B; ++u) {
            const uint32_t e = (k + u < c) ? ((u & 1) ? (w[u / 2] >> 16) : (w[u / 2] & 0xffffu)) : (w[0] & 0xffffu);
            q[u] = decode_entry<NP>(bases, e);
        }
        body(k, q);
        // Keeps the next round's index loads below this round's bodies. Left
        // free, the compiler may issue them right after this round's gathers
        // (a software-pipelined loop), which measured 10-15% slower for the
        // two-record loops (a_ii/source and Jacobi passes) than issuing each
        // gather as soon as its own index returns.
        asm volatile("" ::: "memory");
    }
}

// list_rounds for the grid's range and order (3 or 5 planes in XYZ order, 4
// segments in brick order at range 1); false if the list cannot serve this
// grid.
template <uint32_t B, class Body>
__device__ __forceinline__ bool with_list(const GridDev& g, const NbrDev& L, uint32_t t, uint32_t c, f3 xi,
    Body&& body)
{
    if (g.order == REYN_CELL_ORDER_BRICK) {
        if (g.range != 1)
            return false;
        list_rounds<4, B>(g, L, t, c, xi, body);
    } else if (g.range == 1) {
        list_rounds<3, B>(g, L, t, c, xi, body);
    } else if (g.range == 2) {
        list_rounds<5, B>(g, L, t, c, xi, body);
    } else {
        return false;
    }
    return true;
}

// Whether the density pass keeps a list for this grid (as with_list serves).
__device__ __forceinline__ bool lists_serve(const GridDev& g)
{
    return g.order == REYN_CELL_ORDER_BRICK ? g.range == 1 : g.range <= kListMaxRange;
}

// Calls f(j, x_ij, d2, m_j) for the neighbours of the particle in slot t at
// xi, reading the packed record {x, y, z, m} of j at P[S j]: one 16-byte load
// per listed neighbour instead of one load per SoA field. Through the list
// the particle itself is skipped (returns true); through the grid (no list,
// or a count above the cap) it is included (returns false). The walk reads
// the SoA arrays, which hold the same values, and every list entry gives the
// same x_ij and d2 bits as the walk.
//
// Rounds are padded past the count with the round's first entry passed as
// m_j = 0, so every body must be linear in m_j (each SPH pair term carries
// the factor m_j); a padding call then adds an exact zero for finite particle
// states.
#ifndef REYN_NBR_BATCH
#define REYN_NBR_BATCH 8
#endif
// entries per round: the double-cosine kernel's sin/cos bodies need more
// registers per entry, so it reads fewer entries per round
template <int K> constexpr uint32_t kNbrBatch = (K == KCOS) ? 2u : uint32_t(REYN_NBR_BATCH);
template <bool IND, uint32_t B, uint32_t S = 1, class F>
__device__ __forceinline__ bool nbrs4(const GridDev& g, const NbrDev& L, uint32_t t, f3 xi,
    const float4* __restrict__ P, const float* __restrict__ xs, const float* __restrict__ ys,
    const float* __restrict__ zs, const float* __restrict__ ms, F&& f)
{
    if (L.cnt) {
        const uint32_t c = L.cnt[t];
        if (c <= L.cap
            && with_list<B>(g, L, t, c, xi, [&](uint32_t k, const uint32_t (&q)[B]) {
                   uint32_t j[B];
                   float4 r[B];
#pragma unroll
                   for (uint32_t u = 0; u < B; ++u)
                       j[u] = IND ? gat(g.sorted, q[u]) : q[u];
#pragma unroll
                   for (uint32_t u = 0; u < B; ++u)
                       r[u] = gat(P, S * j[u]);
#pragma unroll
                   for (uint32_t u = 0; u < B; ++u) {
                       const f3 xij = xi - f3 { r[u].x, r[u].y, r[u].z };
                       f(j[u], xij, dot(xij, xij), (k + u < c) ? r[u].w : 0.f);
                   }
               }))
            return true;
    }
    walk<IND>(g, xi, xs, ys, zs, [&](uint32_t j, f3 xij, float d2) { f(j, xij, d2, gat(ms, j)); });
    return false;
}

// ---- particle views ----------------------------------------------------------

struct FluidDev {
    float *x, *y, *z, *vx, *vy, *vz, *m;
    uint32_t n;
};
inline FluidDev to_dev(const reyn_particles& p)
{
    return { p.xx, p.xy, p.xz, p.vx, p.vy, p.vz, p.m, p.n };
}

struct BdyDev {
    const float *x, *y, *z, *m;
    const float4* p4; // packed {x, y, z, m} of the same samples, or NULL
    const uint8_t* near; // per grid cell: samples within reach, or NULL
    uint32_t n;
    GridDev grid;
    int present;
};
inline BdyDev to_dev(const reyn_boundary* b)
{
    BdyDev d {};
    if (b && b->n > 0) {
        d.x = b->xx;
        d.y = b->xy;
        d.z = b->xz;
        d.m = b->m;
        d.n = b->n;
        d.grid = to_dev(b->grid);
        d.near = b->near;
        d.present = 1;
    }
    return d;
}

// ---- scenes ---------------------------------------------------------------------
// The constants of one scene: its fluid grid, boundary, kernel, material
// constants, time step and particle range. A call on one scene passes it by
// value; a batch passes a device table and the scene of every workgroup.
// Workgroups are aligned to scene starts, so a workgroup sees one scene, its
// particles map to threads as in a call on that scene alone, and its partial
// sums are the ones that call would form.

struct SphK {
    KParams kp;
    float h, nu, rho0;
    f3 g;
    float visc_c; // 10 nu
    float visc_eps; // 0.01 h^2, formed in double
    float rho0_sq_inv; // 1 / rho0^2
};

inline SphK make_sphk(const reyn_sph_params& p)
{
    SphK c;
    c.kp = make_kparams(p.kernel, p.h_bar);
    c.h = p.h;
    c.nu = p.nu;
    c.rho0 = p.rho_0;
    c.g = { p.g.x, p.g.y, p.g.z };
    c.visc_c = 10.f * p.nu;
    c.visc_eps = static_cast<float>(0.01 * double(p.h) * double(p.h));
    c.rho0_sq_inv = 1.f / (p.rho_0 * p.rho_0);
    return c;
}

struct SceneDev {
    GridDev g;
    BdyDev b;
    SphK c;
    float dt; // time step of the step in progress
    float kk; // PCISPH stiffness (delta / dt) / dt
    float thr; // the solver loop's error threshold for this scene
    uint32_t first, n; // particle range [first, first + n)
    uint32_t blk_first, nblk; // its workgroups (after the XCD remap)
    uint32_t idx; // scene index: its loop-control block
    f3 box_lo, box_hi; // bounds of the scene's domain (batch bounds pass)
};

struct SceneArg {
    SceneDev one;
    const SceneDev* table; // batch: per scene; unused for one scene
    const uint32_t* blk_scene; // batch: the scene of every workgroup
    const uint32_t* blk_list; // batch: the workgroups launched (NULL: all)
};

template <bool BT> __device__ __forceinline__ const SceneDev& scene_of(const SceneArg& a, uint32_t lb)
{
    if constexpr (BT)
        return a.table[a.blk_scene[lb]];
    else
        return a.one;
}

// The scene of a loop-control workgroup: workgroup s of a batch reduces
// scene s.
template <bool BT> __device__ __forceinline__ const SceneDev& scene_of_ctl(const SceneArg& a)
{
    if constexpr (BT)
        return a.table[blockIdx.x];
    else
        return a.one;
}

// The workgroup's scene, the thread's particle slot and whether it holds one.
#define REYN_SCENE(BT, sa)                                                                                 \
    const uint32_t lb_ = (BT && (sa).blk_list) ? (sa).blk_list[xcd_block(blockIdx.x, gridDim.x)]             \
                                               : xcd_block(blockIdx.x, gridDim.x);                           \
    const SceneDev& sc = scene_of<BT>((sa), lb_);                                                          \
    const uint32_t t = sc.first + (lb_ - sc.blk_first) * kBlock + threadIdx.x;                            \
    const bool live = t < sc.first + sc.n

// Thread -> particle map: in Resort::no mode threads follow the cell order
// through `sorted`, so a wavefront's neighbourhoods overlap in cache.
template <bool IND> __device__ __forceinline__ uint32_t particle_of(const GridDev& g, uint32_t t)
{
    return IND ? g.sorted[t] : t;
}

// Dispatch a generic launcher over the four kernels and the two Resort modes.
template <class F> int with_kernel(reyn_ctx* ctx, int kind, bool ind, F&& f)
{
    switch (kind) {
    case KC2:
        return ind ? f(std::integral_constant<int, KC2> {}, std::true_type {})
                   : f(std::integral_constant<int, KC2> {}, std::false_type {});
    case KB3:
        return ind ? f(std::integral_constant<int, KB3> {}, std::true_type {})
                   : f(std::integral_constant<int, KB3> {}, std::false_type {});
    case KW6:
        return ind ? f(std::integral_constant<int, KW6> {}, std::true_type {})
                   : f(std::integral_constant<int, KW6> {}, std::false_type {});
    case KCOS:
        return ind ? f(std::integral_constant<int, KCOS> {}, std::true_type {})
                   : f(std::integral_constant<int, KCOS> {}, std::false_type {});
    default:
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "unknown kernel kind %d", kind);
    }
}

// Validation helpers shared by the entry points.
int check_grid(reyn_ctx* ctx, const reyn_grid* g, uint32_t n_expected, const char* what);
int check_fluid(reyn_ctx* ctx, const reyn_particles* p);
int check_bdy(reyn_ctx* ctx, const reyn_boundary* b);
int check_sph(reyn_ctx* ctx, const reyn_sph_params* prm);

} // namespace reyn
