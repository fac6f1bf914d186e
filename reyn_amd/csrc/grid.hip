// Uniform grid: cell keys, counting sort with a deterministic in-cell order,
// prefix scan, reorder gather and the reference's query test kernels.
//
// Build pipeline (one launch each, all on the context stream):
//   k_count     key_i = cell(x_i); run-aggregated atomicAdd on counts[key]
//               gives every point a slot inside its cell
//   k_scan_*    prefix = exclusive_scan(counts) (reduce / top / down), with
//               prefix[C] = n and counts reset to 0 for the next build
//   k_scatter   tmp[prefix[key_i] + slot_i] = i
//   k_stabilize sorted[prefix[key] + rank] = i with rank = #{j in cell, j < i}
// The atomic arrival order decides the slots, k_stabilize removes that
// dependence so `sorted` is bit-reproducible (ascending index per cell).
#include <cmath>
#include <new>

#include "grid.hpp"

namespace reyn {
namespace {

constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kBlock * kScanItems; // 4096 cells per scan block
// cells holding more points than this are left in slot order (and reported):
// their O(c^2) ranking would be unbounded for degenerate inputs
constexpr uint32_t kMaxStableCell = 4096;
// records per staging window (two per wavefront) of the query walks
#ifndef REYN_QSTAGE
#define REYN_QSTAGE 96
#endif
constexpr uint32_t kQueryStage = REYN_QSTAGE;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= o)
            v += t;
    }
    return v;
}

// Exclusive scan across the block; returns the block total in *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t* total)
{
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == kWave - 1)
        lds[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        const uint32_t s = lds[k];
        if (k < w)
            off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

__global__ void k_zero_status(uint32_t* status)
{
    if (threadIdx.x < 4)
        status[threadIdx.x] = 0u;
}

// Keys and in-cell slots; for a batch the key of a point of scene s is offset
// by cell_off[s], the first cell of that scene in the concatenated tables.
template <bool BT>
__global__ __launch_bounds__(kBlock) void k_count(SceneArg sa, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, const uint32_t* __restrict__ cell_off,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ keys, uint32_t* __restrict__ slots,
    uint32_t* __restrict__ status)
{
    REYN_SCENE(BT, sa);
    const GridDev& g = sc.g;
    const uint32_t i = t;
    const int lane = threadIdx.x & (kWave - 1);
    const bool valid = live;
    uint32_t key = 0xFFFFFFFFu;
    bool outside = false;
    if (valid) {
        const f3 p = ld3(xs, ys, zs, i);
        const int3c c = cell_of(g, p);
        key = cell_key(g, c) + (BT ? cell_off[sc.idx] : 0u);
        const int R = g.range;
        outside = c.x < R || c.y < R || c.z < R || c.x > g.nx - 1 - R || c.y > g.ny - 1 - R
            || c.z > g.nz - 1 - R;
    }
    // runs of equal keys in consecutive lanes share one atomic
    const uint32_t prev = __shfl_up(key, 1, kWave);
    const bool head = valid && (lane == 0 || prev != key);
    const uint64_t V = __ballot(valid);
    const uint64_t H = __ballot(head);
    const uint64_t le = (lane == kWave - 1) ? ~0ull : ((2ull << lane) - 1ull);
    const uint64_t hb = H & le;
    const int h = hb ? 63 - __clzll(static_cast<long long>(hb)) : lane;
    uint32_t base = 0;
    if (head) {
        const uint64_t ends = (H | ~V) & ~le; // first later head or invalid lane
        const int end = ends ? (__ffsll(static_cast<long long>(ends)) - 1) : kWave;
        base = atomicAdd(&counts[key], static_cast<uint32_t>(end - lane));
    }
    base = __shfl(base, h, kWave);
    if (valid) {
        keys[i] = key;
        slots[i] = base + static_cast<uint32_t>(lane - h);
        if (outside)
            atomicAdd(&status[0], 1u);
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ counts,
    uint32_t ncells, uint32_t* __restrict__ block_sums)
{
    __shared__ double lds[kBlock / kWave];
    const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t s = 0;
    if (base + kScanItems <= ncells) {
        const uint4* p = reinterpret_cast<const uint4*>(counts + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const uint4 v = p[k];
            s += v.x + v.y + v.z + v.w;
        }
    } else {
        for (size_t k = base; k < ncells && k < base + kScanItems; ++k)
            s += counts[k];
    }
    const double tot = block_reduce<0>(static_cast<double>(s), lds);
    if (threadIdx.x == 0)
        block_sums[blockIdx.x] = static_cast<uint32_t>(tot);
}

// Single workgroup: exclusive scan of the block sums; prefix[C] = total.
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ block_sums,
    uint32_t nb, uint32_t* __restrict__ prefix, uint32_t ncells)
{
    __shared__ uint32_t lds[1024 / kWave];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += blockDim.x) {
        const uint32_t k = c0 + threadIdx.x;
        const uint32_t v = k < nb ? block_sums[k] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, lds, &tot);
        if (k < nb)
            block_sums[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0)
        prefix[ncells] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(uint32_t* __restrict__ counts,
    uint32_t ncells, const uint32_t* __restrict__ block_sums, uint32_t* __restrict__ prefix,
    uint32_t* __restrict__ status)
{
    __shared__ uint32_t lds[kBlock / kWave];
    const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    const bool full = base + kScanItems <= ncells;
    if (full) {
        const uint4* p = reinterpret_cast<const uint4*>(counts + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const uint4 q = p[k];
            v[4 * k + 0] = q.x;
            v[4 * k + 1] = q.y;
            v[4 * k + 2] = q.z;
            v[4 * k + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            v[k] = (base + k < ncells) ? counts[base + k] : 0u;
    }
    uint32_t s = 0, mx = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t c = v[k];
        v[k] = s;
        s += c;
        mx = max(mx, c);
    }
    uint32_t tot;
    const uint32_t off = block_excl_scan(s, lds, &tot) + block_sums[blockIdx.x];
    if (full) {
        uint4* p = reinterpret_cast<uint4*>(prefix + base);
        uint4* c = reinterpret_cast<uint4*>(counts + base);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            p[k] = make_uint4(off + v[4 * k], off + v[4 * k + 1], off + v[4 * k + 2], off + v[4 * k + 3]);
            c[k] = make_uint4(0u, 0u, 0u, 0u);
        }
    } else {
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < ncells) {
                prefix[base + k] = off + v[k];
                counts[base + k] = 0u;
            }
        }
    }
    // largest cell count of the build (wave max, one atomic per wave)
    uint32_t wm = mx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        wm = max(wm, __shfl_xor(wm, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0 && wm > 0)
        atomicMax(&status[1], wm);
}

__global__ __launch_bounds__(kBlock) void k_scatter(uint32_t n, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ slots, const uint32_t* __restrict__ prefix,
    uint32_t* __restrict__ tmp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    tmp[prefix[keys[i]] + slots[i]] = i;
}

__global__ __launch_bounds__(kBlock) void k_stabilize(uint32_t n, const uint32_t* __restrict__ tmp,
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ prefix,
    uint32_t* __restrict__ sorted, uint32_t* __restrict__ status)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const uint32_t v = tmp[p];
    const uint32_t k = keys[v];
    const uint32_t s = prefix[k], e = prefix[k + 1];
    if (e - s > kMaxStableCell) {
        sorted[p] = v;
        if (p == s)
            atomicAdd(&status[2], 1u);
        return;
    }
    uint32_t r = 0;
    for (uint32_t q = s; q < e; ++q)
        r += (tmp[q] < v) ? 1u : 0u;
    sorted[s + r] = v;
}

struct GatherArgs {
    const float* src[16];
    float* dst[16];
};

__global__ __launch_bounds__(kBlock) void k_gather(uint32_t n, const uint32_t* __restrict__ sorted,
    GatherArgs a, uint32_t nfields)
{
    const uint32_t i = global_index();
    if (i >= n)
        return;
    const uint32_t s = sorted[i];
    // every field's load before the first store: the stores may alias the
    // sources as far as the compiler knows, so a load after a store would
    // wait for it and each field would take its own round trip
    float v[16];
#pragma unroll
    for (uint32_t f = 0; f < 16; ++f)
        if (f < nfields)
            v[f] = a.src[f][s];
#pragma unroll
    for (uint32_t f = 0; f < 16; ++f)
        if (f < nfields)
            a.dst[f][i] = v[f];
}

// near[c] = 1 when a point lies in the (2R+1)^3 cells around cell c: the
// strips a walk from cell c would visit, each tested with two cell starts.
__global__ __launch_bounds__(kBlock) void k_occupancy(GridDev g, uint64_t ncells, uint8_t* __restrict__ near)
{
    const uint64_t c = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (c >= ncells)
        return;
    const int R = g.range;
    if (g.order == REYN_CELL_ORDER_BRICK) {
        // key -> cell, then every cell in reach tested by its own count
        const uint32_t br = static_cast<uint32_t>(c >> 3), loc = static_cast<uint32_t>(c & 7);
        const int bx = static_cast<int>(br % uint32_t(g.nbx));
        const int by = static_cast<int>((br / uint32_t(g.nbx)) % uint32_t(g.ny / 2));
        const int bz = static_cast<int>(br / uint32_t(g.nbxy));
        const int x = 2 * bx + int(loc & 1), y = 2 * by + int((loc >> 1) & 1), z = 2 * bz + int(loc >> 2);
        uint8_t any = 0;
        for (int zz = max(z - R, 0); zz <= min(z + R, g.nz - 1) && !any; ++zz)
            for (int yy = max(y - R, 0); yy <= min(y + R, g.ny - 1) && !any; ++yy)
                for (int xx = max(x - R, 0); xx <= min(x + R, g.nx - 1) && !any; ++xx) {
                    const uint32_t k = cell_key(g, int3c { xx, yy, zz });
                    any = g.prefix[k + 1] > g.prefix[k] ? 1 : 0;
                }
        near[c] = any;
        return;
    }
    const int x = static_cast<int>(c % uint64_t(g.nx));
    const int y = static_cast<int>((c / uint64_t(g.nx)) % uint64_t(g.ny));
    const int z = static_cast<int>(c / uint64_t(g.nxny));
    const int x0 = max(x - R, 0), x1 = min(x + R, g.nx - 1);
    uint8_t any = 0;
    for (int zz = max(z - R, 0); zz <= min(z + R, g.nz - 1) && !any; ++zz)
        for (int yy = max(y - R, 0); yy <= min(y + R, g.ny - 1) && !any; ++yy) {
            const uint64_t base = uint64_t(zz) * uint64_t(g.nxny) + uint64_t(yy) * uint64_t(g.nx);
            any = g.prefix[base + x1 + 1] > g.prefix[base + x0] ? 1 : 0;
        }
    near[c] = any;
}

template <bool IND>
__global__ __launch_bounds__(kBlock) void k_query_stats(GridDev g, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n,
    uint32_t* __restrict__ count, float* __restrict__ len2, float* __restrict__ vx,
    float* __restrict__ vy, float* __restrict__ vz)
{
    // every lane stays until the walk is done (the staged walk reduces over
    // the wavefront); lanes past the end visit nothing
    const uint32_t t = global_index();
    const bool live = t < n;
    const uint32_t i = live ? particle_of<IND>(g, t) : 0u;
    const f3 xi = live ? ld3(xs, ys, zs, i) : g.bmin;
    uint32_t c = 0;
    float l2 = 0.f;
    f3 v { 0.f, 0.f, 0.f };
    auto body = [&](uint32_t j, f3 xij, float d2) {
        c += (j == i) ? 0u : 1u;
        l2 += d2;
        v += xij;
    };
    if constexpr (IND) {
        if (live)
            walk<IND>(g, xi, xs, ys, zs, body);
    } else {
        __shared__ float4 stage[kBlock / kWave][2 * kQueryStage];
        walk_staged<kQueryStage, 2>(g, xi, xs, ys, zs, stage[threadIdx.x / kWave], live,
            [&](uint32_t j, f3 xij, float d2, f3) { body(j, xij, d2); });
    }
    if (!live)
        return;
    if (count)
        count[i] = c;
    if (len2)
        len2[i] = l2;
    if (vx)
        vx[i] = v.x;
    if (vy)
        vy[i] = v.y;
    if (vz)
        vz[i] = v.z;
}

template <int K, bool IND>
__global__ __launch_bounds__(kBlock) void k_query_bench(GridDev g, KParams kp,
    const float* __restrict__ xs, const float* __restrict__ ys, const float* __restrict__ zs,
    uint32_t n, float* __restrict__ res)
{
    const uint32_t t = global_index();
    const bool live = t < n;
    const uint32_t i = live ? particle_of<IND>(g, t) : 0u;
    const f3 xi = live ? ld3(xs, ys, zs, i) : g.bmin;
    float acc = 0.f;
    if constexpr (IND) {
        if (live)
            walk<IND>(g, xi, xs, ys, zs, [&](uint32_t j, f3, float d2) { acc += xi.x * xs[j] * kW<K>(kp, d2); });
    } else {
        __shared__ float4 stage[kBlock / kWave][2 * kQueryStage];
        walk_staged<kQueryStage, 2>(g, xi, xs, ys, zs, stage[threadIdx.x / kWave], live,
            [&](uint32_t, f3, float d2, f3 xj) { acc += xi.x * xj.x * kW<K>(kp, d2); });
    }
    if (live)
        res[i] = acc;
}

int ensure_capacity(reyn_grid_builder* b, uint32_t n)
{
    reyn_ctx* ctx = b->ctx;
    if (n <= b->cap)
        return REYN_OK;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint32_t** arrs[] = { &b->keys, &b->slots, &b->tmp, &b->sorted };
    for (uint32_t** a : arrs) {
        if (*a) {
            REYN_HIP(ctx, hipFree(*a));
            *a = nullptr;
        }
    }
    b->cap = 0;
    const size_t bytes = size_t(n) * sizeof(uint32_t);
    for (uint32_t** a : arrs)
        REYN_HIP(ctx, hipMalloc(reinterpret_cast<void**>(a), bytes));
    b->cap = n;
    return REYN_OK;
}

// count -> scan -> scatter -> stabilize over the builder's cells for n
// points; `count` launches the key/slot kernel.
template <class Count>
int sort_points(reyn_grid_builder* b, uint32_t n, uint32_t* status_host, Count&& count)
{
    reyn_ctx* ctx = b->ctx;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    REYN_TRY(ensure_capacity(b, n > 0 ? n : 1));
    hipStream_t st = ctx->stream;
    {
        ProfScope _prof(ctx, "k_zero_status");
        k_zero_status<<<1, kWave, 0, st>>>(b->status);
    }
    REYN_LAUNCHED(ctx, "k_zero_status");
    const uint32_t nb = blocks_for(n);
    if (n > 0) {
        {
            ProfScope _prof(ctx, "k_count");
            count(st);
        }
        REYN_LAUNCHED(ctx, "k_count");
    }
    const uint32_t C = static_cast<uint32_t>(b->ncells);
    const uint32_t nsb = static_cast<uint32_t>(b->nblocks_scan);
    {
        ProfScope _prof(ctx, "k_scan_reduce");
        k_scan_reduce<<<nsb, kBlock, 0, st>>>(b->counts, C, b->block_sums);
    }
    REYN_LAUNCHED(ctx, "k_scan_reduce");
    {
        ProfScope _prof(ctx, "k_scan_top");
        k_scan_top<<<1, 1024, 0, st>>>(b->block_sums, nsb, b->prefix, C);
    }
    REYN_LAUNCHED(ctx, "k_scan_top");
    {
        ProfScope _prof(ctx, "k_scan_down");
        k_scan_down<<<nsb, kBlock, 0, st>>>(b->counts, C, b->block_sums, b->prefix, b->status);
    }
    REYN_LAUNCHED(ctx, "k_scan_down");
    if (n > 0) {
        {
            ProfScope _prof(ctx, "k_scatter");
            k_scatter<<<nb, kBlock, 0, st>>>(n, b->keys, b->slots, b->prefix, b->tmp);
        }
        REYN_LAUNCHED(ctx, "k_scatter");
        {
            ProfScope _prof(ctx, "k_stabilize");
            k_stabilize<<<nb, kBlock, 0, st>>>(n, b->tmp, b->keys, b->prefix, b->sorted, b->status);
        }
        REYN_LAUNCHED(ctx, "k_stabilize");
    }
    if (status_host)
        REYN_HIP(ctx, hipMemcpyAsync(status_host, b->status, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    return REYN_OK;
}

// One scene: fills b->geom.prefix/sorted/n.
int build(reyn_grid_builder* b, float search_radius, const float* xx, const float* xy,
    const float* xz, uint32_t n, uint32_t* status_host)
{
    reyn_ctx* ctx = b->ctx;
    if (!(search_radius > 0.f) || search_radius > b->search_radius * 1.0000001f)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
            "construct: search radius %g must be positive and at most the builder's %g",
            static_cast<double>(search_radius), static_cast<double>(b->search_radius));
    if (n > 0 && (!xx || !xy || !xz))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "construct: NULL position pointer");
    b->geom.r_c_2 = search_radius * search_radius;
    SceneArg sa {};
    sa.one.g = to_dev(b->geom);
    sa.one.n = n;
    sa.one.nblk = blocks_for(n);
    REYN_TRY(sort_points(b, n, status_host, [&](hipStream_t st) {
        k_count<false><<<blocks_for(n), kBlock, 0, st>>>(sa, xx, xy, xz, nullptr, b->counts, b->keys, b->slots,
            b->status);
    }));
    b->geom.prefix = b->prefix;
    b->geom.sorted = b->sorted;
    b->geom.n = n;
    return REYN_OK;
}

} // namespace

int sort_batch(reyn_grid_builder* b, const SceneArg& sa, uint32_t nblocks, const uint32_t* cell_off,
    const float* xx, const float* xy, const float* xz, uint32_t n, uint32_t* status_host)
{
    return sort_points(b, n, status_host, [&](hipStream_t st) {
        k_count<true><<<nblocks, kBlock, 0, st>>>(sa, xx, xy, xz, cell_off, b->counts, b->keys, b->slots,
            b->status);
    });
}

int finish_build(reyn_grid_builder* b, const uint32_t* status_host)
{
    reyn_ctx* ctx = b->ctx;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (status_host[2] > 0)
        return set_error(ctx, REYN_ERR_CAPACITY,
            "construct: %u cells hold more than %u points (largest %u); the scene is degenerate",
            status_host[2], kMaxStableCell, status_host[1]);
    if (b->cell_cap && status_host[1] > b->cell_cap)
        return set_error(ctx, REYN_ERR_CAPACITY,
            "construct: a cell holds %u points, more than the %u a neighbour walk over the cells in reach "
            "may visit within its %u-candidate budget; the scene is degenerate",
            status_host[1], b->cell_cap, kWalkBudget);
    if (status_host[0] > 0)
        return set_error(ctx, REYN_ERR_OUT_OF_BOUNDS,
            "construct: %u points lie outside the grid's query-safe cells and were clamped",
            status_host[0]);
    return REYN_OK;
}

int launch_gather(reyn_ctx* ctx, const uint32_t* sorted, uint32_t n, const float* const* src,
    float* const* dst, uint32_t nfields)
{
    if (nfields == 0 || n == 0)
        return REYN_OK;
    for (uint32_t f0 = 0; f0 < nfields; f0 += 16) {
        GatherArgs a {};
        const uint32_t nf = std::min<uint32_t>(16, nfields - f0);
        for (uint32_t f = 0; f < nf; ++f) {
            if (!src[f0 + f] || !dst[f0 + f])
                return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "gather: NULL field %u", f0 + f);
            a.src[f] = src[f0 + f];
            a.dst[f] = dst[f0 + f];
        }
        {
            ProfScope _prof(ctx, "k_gather");
            k_gather<<<blocks_for(n), kBlock, 0, ctx->stream>>>(n, sorted, a, nf);
        }
        REYN_LAUNCHED(ctx, "k_gather");
    }
    return REYN_OK;
}

} // namespace reyn

namespace reyn {

int grid_geometry(reyn_ctx* ctx, reyn_float3 bmin, reyn_float3 bmax, float search_radius, float cell_ratio,
    reyn_grid* out, size_t* ncells, int order)
{
    if (order != REYN_CELL_ORDER_XYZ && order != REYN_CELL_ORDER_BRICK)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "grid: unknown cell order %d", order);
    if (!(search_radius > 0.f) || !std::isfinite(search_radius))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "grid: search radius must be positive");
    if (!(cell_ratio >= 0.25f && cell_ratio <= 4.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "grid: cell ratio must lie in [0.25, 4]");
    // geometry as in uniformgrid.cu:65-96, with the cell size scaled by the ratio
    const float cell = search_radius * cell_ratio;
    const float pad = static_cast<float>(search_radius * 2.0);
    const float lo[3] = { bmin.x - pad, bmin.y - pad, bmin.z - pad };
    const float hi[3] = { bmax.x + pad, bmax.y + pad, bmax.z + pad };
    int n[3];
    for (int a = 0; a < 3; ++a) {
        const float e = std::ceil((hi[a] - lo[a]) / cell);
        if (!(e > 0.f) || !(e < 2.0e9f))
            return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
                "Negative number of grid cells encountered in construction of uniform grid.");
        n[a] = static_cast<int>(e);
        // bricks of 2 x 2 x 2 cells tile the grid: the counts are even (the
        // extra layer of cells lies beyond the padded bound and stays empty)
        if (order == REYN_CELL_ORDER_BRICK)
            n[a] += n[a] & 1;
    }
    const double cells = double(n[0]) * double(n[1]) * double(n[2]);
    if (cells >= 1073741824.0)
        return set_error(ctx, REYN_ERR_CAPACITY, "grid: %.0f cells exceed the 2^30 cell limit", cells);
    *out = reyn_grid {};
    out->bound_min = { lo[0], lo[1], lo[2] };
    out->cell_size = cell;
    out->r_c_2 = search_radius * search_radius;
    out->nx = n[0];
    out->ny = n[1];
    out->nz = n[2];
    // cells searched on each side: ceil(radius / cell)
    out->range = static_cast<int>(std::ceil(1.0 / double(cell_ratio) - 1e-6));
    out->order = order;
    *ncells = size_t(n[0]) * size_t(n[1]) * size_t(n[2]);
    return REYN_OK;
}

int builder_alloc(reyn_ctx* ctx, size_t ncells, float search_radius, reyn_grid_builder** out)
{
    *out = nullptr;
    if (ncells >= (size_t(1) << 32) - 1)
        return set_error(ctx, REYN_ERR_CAPACITY, "grid: %zu cells exceed the 2^32 - 1 cell limit", ncells);
    reyn_grid_builder* b = new (std::nothrow) reyn_grid_builder();
    if (!b)
        return REYN_ERR_OUT_OF_MEMORY;
    b->ctx = ctx;
    b->search_radius = search_radius;
    b->ncells = ncells;
    b->nblocks_scan = (b->ncells + kScanTile - 1) / kScanTile;
    int rc = check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&b->counts), b->ncells * 4), "hipMalloc counts");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&b->prefix), (b->ncells + 1) * 4), "hipMalloc prefix");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&b->block_sums), b->nblocks_scan * 4), "hipMalloc block sums");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&b->status), 4 * 4), "hipMalloc status");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMemsetAsync(b->counts, 0, b->ncells * 4, ctx->stream), "hipMemset counts");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc != REYN_OK) {
        reyn_grid_builder_destroy(b);
        return rc;
    }
    *out = b;
    return REYN_OK;
}

} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_grid_builder_create(reyn_ctx* ctx, reyn_float3 bmin, reyn_float3 bmax,
    float search_radius, float cell_ratio, reyn_grid_builder** out)
{
    if (!ctx || !out)
        return REYN_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    return reyn_grid_builder_create_ordered(ctx, bmin, bmax, search_radius, cell_ratio, REYN_CELL_ORDER_XYZ, out);
}

int reyn_grid_builder_create_ordered(reyn_ctx* ctx, reyn_float3 bmin, reyn_float3 bmax, float search_radius,
    float cell_ratio, int32_t order, reyn_grid_builder** out)
{
    if (!ctx || !out)
        return REYN_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    reyn_grid geom {};
    size_t ncells = 0;
    REYN_TRY(grid_geometry(ctx, bmin, bmax, search_radius, cell_ratio, &geom, &ncells, order));
    reyn_grid_builder* b = nullptr;
    REYN_TRY(builder_alloc(ctx, ncells, search_radius, &b));
    b->geom = geom;
    b->cell_cap = walk_cell_cap(geom.range);
    *out = b;
    return REYN_OK;
}

int reyn_grid_cell_capacity(int32_t range, uint32_t* cap)
{
    if (!cap || range < 1 || range > 16)
        return REYN_ERR_INVALID_ARGUMENT;
    *cap = walk_cell_cap(range);
    return REYN_OK;
}

int reyn_grid_builder_destroy(reyn_grid_builder* b)
{
    if (!b)
        return REYN_OK;
    if (b->ctx) {
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
    }
    uint32_t* arrs[] = { b->counts, b->prefix, b->block_sums, b->status, b->keys, b->slots, b->tmp, b->sorted };
    for (uint32_t* a : arrs)
        if (a)
            (void)hipFree(a);
    if (b->status_pinned)
        (void)hipHostFree(b->status_pinned);
    delete b;
    return REYN_OK;
}

int reyn_grid_builder_geometry(const reyn_grid_builder* b, reyn_grid* out)
{
    if (!b || !out)
        return REYN_ERR_INVALID_ARGUMENT;
    *out = b->geom;
    return REYN_OK;
}

int reyn_grid_builder_defer_check(reyn_grid_builder* b, int32_t on)
{
    if (!b || !b->ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    reyn_ctx* ctx = b->ctx;
    if (on && !b->status_pinned) {
        REYN_HIP(ctx, hipSetDevice(ctx->device));
        REYN_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&b->status_pinned), 4 * sizeof(uint32_t),
                          hipHostMallocDefault));
    }
    b->defer = on != 0;
    return on ? REYN_OK : reyn_grid_builder_check(ctx, b);
}

int reyn_grid_builder_check(reyn_ctx* ctx, reyn_grid_builder* b)
{
    if (!ctx || !b || b->ctx != ctx)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "check: bad context or builder");
    if (!b->pending)
        return REYN_OK;
    b->pending = false;
    return finish_build(b, b->status_pinned);
}

} // extern "C"

namespace {

// The status buffer of a construct call: the builder's own pinned words when
// its checks are deferred (after reading what an unchecked earlier build
// left there), else the context's staging memory.
int status_target(reyn_ctx* ctx, reyn_grid_builder* b, uint32_t** st)
{
    if (b->defer) {
        REYN_TRY(reyn_grid_builder_check(ctx, b));
        *st = b->status_pinned;
        return REYN_OK;
    }
    return ctx_pinned(ctx, 64, reinterpret_cast<void**>(st));
}

// Deferred: the build is in flight and its status unread; else wait and check.
int end_build(reyn_grid_builder* b, const uint32_t* st)
{
    if (b->defer) {
        b->pending = true;
        return REYN_OK;
    }
    return finish_build(b, st);
}

} // namespace

extern "C" {

int reyn_grid_construct(reyn_ctx* ctx, reyn_grid_builder* b, float search_radius,
    const float* xx, const float* xy, const float* xz, uint32_t n, reyn_grid* out)
{
    if (!ctx || !b || !out || b->ctx != ctx)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "construct: bad context, builder or output");
    uint32_t* st = nullptr;
    REYN_TRY(status_target(ctx, b, &st));
    REYN_TRY(build(b, search_radius, xx, xy, xz, n, st));
    *out = b->geom;
    return end_build(b, st);
}

int reyn_grid_construct_and_reorder(reyn_ctx* ctx, reyn_grid_builder* b, float search_radius,
    float** fields, float** tmp, uint32_t nfields, uint32_t n, reyn_grid* out)
{
    if (!ctx || !b || !out || b->ctx != ctx || !fields || !tmp || nfields < 3)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
            "construct_and_reorder: bad arguments (x, y, z must be the first three fields)");
    uint32_t* st = nullptr;
    REYN_TRY(status_target(ctx, b, &st));
    REYN_TRY(build(b, search_radius, fields[0], fields[1], fields[2], n, st));
    REYN_TRY(launch_gather(ctx, b->sorted, n, fields, tmp, nfields));
    for (uint32_t f = 0; f < nfields; ++f) {
        float* t = fields[f];
        fields[f] = tmp[f];
        tmp[f] = t;
    }
    *out = b->geom;
    out->sorted = nullptr; // Resort::yes
    return end_build(b, st);
}

int reyn_grid_debug_tables(reyn_grid_builder* b, const uint32_t** counts, const uint32_t** keys)
{
    if (!b)
        return REYN_ERR_INVALID_ARGUMENT;
    if (counts)
        *counts = b->counts;
    if (keys)
        *keys = b->keys;
    return REYN_OK;
}

int reyn_gather_f32(reyn_ctx* ctx, const uint32_t* sorted, uint32_t n, const float* const* src,
    float* const* dst, uint32_t nfields)
{
    if (!ctx || (n && (!sorted || !src || !dst)))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "gather: NULL argument");
    return launch_gather(ctx, sorted, n, src, dst, nfields);
}

int reyn_grid_query_stats(reyn_ctx* ctx, const reyn_grid* grid, const float* xx,
    const float* xy, const float* xz, uint32_t n, uint32_t* count, float* len2, float* vx,
    float* vy, float* vz)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_TRY(check_grid(ctx, grid, n, "query_stats"));
    if (n == 0)
        return REYN_OK;
    const GridDev g = to_dev(*grid);
    {
        ProfScope _prof(ctx, "k_query_stats");
        if (grid->sorted)
            k_query_stats<true><<<blocks_for(n), kBlock, 0, ctx->stream>>>(g, xx, xy, xz, n, count, len2, vx, vy, vz);
        else
            k_query_stats<false><<<blocks_for(n), kBlock, 0, ctx->stream>>>(g, xx, xy, xz, n, count, len2, vx, vy, vz);
    }
    REYN_LAUNCHED(ctx, "k_query_stats");
    return REYN_OK;
}

int reyn_grid_occupancy(reyn_ctx* ctx, const reyn_grid* grid, uint8_t* near)
{
    if (!ctx || !grid || !near || !grid->prefix)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "grid occupancy: NULL argument");
    if (grid->nx <= 0 || grid->ny <= 0 || grid->nz <= 0 || grid->range < 0)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "grid occupancy: bad grid");
    const uint64_t nc = uint64_t(grid->nx) * uint64_t(grid->ny) * uint64_t(grid->nz);
    const uint64_t nb = (nc + kBlock - 1) / kBlock;
    if (nb >= (uint64_t(1) << 31))
        return set_error(ctx, REYN_ERR_CAPACITY, "grid occupancy: too many cells");
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    {
        ProfScope _prof(ctx, "k_occupancy");
        k_occupancy<<<static_cast<uint32_t>(nb), kBlock, 0, ctx->stream>>>(to_dev(*grid), nc, near);
    }
    REYN_LAUNCHED(ctx, "k_occupancy");
    return REYN_OK;
}

int reyn_grid_query_bench(reyn_ctx* ctx, const reyn_grid* grid, int32_t kernel, float h_bar,
    const float* xx, const float* xy, const float* xz, uint32_t n, float* res)
{
    if (!ctx || !res)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "query_bench: NULL argument");
    REYN_TRY(check_grid(ctx, grid, n, "query_bench"));
    if (n == 0)
        return REYN_OK;
    const GridDev g = to_dev(*grid);
    const KParams kp = make_kparams(kernel, h_bar);
    return with_kernel(ctx, kernel, grid->sorted != nullptr, [&](auto KC, auto IC) -> int {
        constexpr int K = decltype(KC)::value;
        constexpr bool I = decltype(IC)::value;
        {
            ProfScope _prof(ctx, "k_query_bench");
            k_query_bench<K, I><<<blocks_for(n), kBlock, 0, ctx->stream>>>(g, kp, xx, xy, xz, n, res);
        }
        REYN_LAUNCHED(ctx, "k_query_bench");
        return REYN_OK;
    });
}

} // extern "C"
