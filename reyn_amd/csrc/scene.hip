// Device-side lattice scenes: the dam-break fluid block, the wall shell of its
// box and a sphere obstacle, generated on the GPU in the same double-precision
// operations as the host generator (reyn_amd/scenes.py), so both give the same
// float bits.
//
// The reference initialises its fluid on the device from a box of lattice
// points with curand jitter (src/scene/scene.cu:14-54, 239-253) and samples
// boundaries from meshes (sample_boundary.cu); here the boundary is the
// lattice shell of the box plus Fibonacci-lattice sphere shells, and the
// jitter comes from a counter-based hash of (seed, point, axis) that numpy
// reproduces exactly.
#include <cmath>

#include "batch.hpp"
#include "common.hpp"
#include "sph.hpp"

namespace reyn {
namespace {

#pragma clang fp contract(off)

__host__ __device__ inline uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// U[0, 1) with 24 bits from (seed, point, axis)
__device__ inline double hash_uniform(uint32_t seed, uint32_t idx, uint32_t axis)
{
    const uint32_t h = mix32(mix32(seed * 0x9e3779b9u + 0x7f4a7c15u) ^ (idx * 3u + axis));
    return double(h >> 8) * (1.0 / 16777216.0);
}

struct DamDev {
    uint32_t nx, ny, nz;
    double s, jit;
    uint32_t seed;
    int sphere;
    double cx, cy, cz, cull; // cull radius: radius + s / 2
};

// Lattice point `idx` (x fastest): (i + 1/2) s + jitter, and whether it lies
// outside the sphere's cull radius.
__device__ inline bool fluid_point(const DamDev& d, uint32_t idx, float* out)
{
    const uint32_t ix = idx % d.nx, iy = (idx / d.nx) % d.ny, iz = idx / (d.nx * d.ny);
    const uint32_t ii[3] = { ix, iy, iz };
    double p[3];
    for (int a = 0; a < 3; ++a) {
        double v = (double(ii[a]) + 0.5) * d.s;
        v = v + 0.0; // origin
        if (d.jit != 0.0) {
            const double lo = -(d.jit * d.s), hi = d.jit * d.s;
            v = v + (lo + (hi - lo) * hash_uniform(d.seed, idx, a));
        }
        out[a] = static_cast<float>(v);
        p[a] = double(out[a]);
    }
    if (!d.sphere)
        return true;
    const double dx = p[0] - d.cx, dy = p[1] - d.cy, dz = p[2] - d.cz;
    return sqrt((dx * dx + dy * dy) + dz * dz) > d.cull;
}

// Per-workgroup counts of kept points.
__global__ __launch_bounds__(kBlock) void k_fluid_count(DamDev d, uint32_t n, uint32_t* __restrict__ counts)
{
    __shared__ uint32_t wc[kBlock / kWave];
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    float q[3];
    const bool keep = idx < n && fluid_point(d, idx, q);
    const uint64_t m = __ballot(keep);
    if ((threadIdx.x & (kWave - 1)) == 0)
        wc[threadIdx.x / kWave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int w = 0; w < kBlock / kWave; ++w)
            c += wc[w];
        counts[blockIdx.x] = c;
    }
}

// Exclusive scan of the workgroup counts in one workgroup; the total goes to
// counts[nb].
__global__ __launch_bounds__(1024) void k_fluid_scan(uint32_t* __restrict__ counts, uint32_t nb)
{
    __shared__ uint32_t part[1024];
    const uint32_t per = (nb + 1023) / 1024, b0 = threadIdx.x * per;
    uint32_t s = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < nb)
            s += counts[b0 + k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int k = 0; k < 1024; ++k) {
            const uint32_t v = part[k];
            part[k] = run;
            run += v;
        }
        counts[nb] = run;
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (uint32_t k = 0; k < per; ++k)
        if (b0 + k < nb) {
            const uint32_t v = counts[b0 + k];
            counts[b0 + k] = run;
            run += v;
        }
}

// Writes the kept points in lattice order.
__global__ __launch_bounds__(kBlock) void k_fluid_write(DamDev d, uint32_t n, const uint32_t* __restrict__ offs,
    float* __restrict__ x, float* __restrict__ y, float* __restrict__ z)
{
    __shared__ uint32_t wc[kBlock / kWave];
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    float q[3];
    const bool keep = idx < n && fluid_point(d, idx, q);
    const uint64_t m = __ballot(keep);
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    if (lane == 0)
        wc[wv] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t base = offs[blockIdx.x];
    for (uint32_t w = 0; w < wv; ++w)
        base += wc[w];
    if (keep) {
        const uint32_t r = base + static_cast<uint32_t>(__popcll(m & ((uint64_t(1) << lane) - 1)));
        x[r] = q[0];
        y[r] = q[1];
        z[r] = q[2];
    }
}

struct WallDev {
    int nx, ny, nz, L;
    double sx, sy, sz, s; // start = lo - L s per axis
};

// Shell point k in lattice order (x fastest): the points of the (nx, ny, nz)
// lattice outside the inner index box [L, n - L)^3.
__global__ __launch_bounds__(kBlock) void k_walls(WallDev w, uint32_t count, float* __restrict__ bx,
    float* __restrict__ by, float* __restrict__ bz)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count)
        return;
    const uint64_t nx = w.nx, ny = w.ny, L = w.L;
    const bool hollow = w.nx > 2 * w.L && w.ny > 2 * w.L && w.nz > 2 * w.L;
    uint64_t ix, iy, iz;
    if (!hollow) {
        ix = k % nx;
        iy = (k / nx) % ny;
        iz = k / (nx * ny);
    } else {
        const uint64_t layer = nx * ny, mid = layer - (nx - 2 * L) * (ny - 2 * L);
        const uint64_t nmid = uint64_t(w.nz) - 2 * L;
        uint64_t r = k;
        if (r < L * layer) {
            iz = r / layer;
            r %= layer;
            iy = r / nx;
            ix = r % nx;
        } else if (r < L * layer + nmid * mid) {
            r -= L * layer;
            iz = L + r / mid;
            r %= mid;
            if (r < L * nx) {
                iy = r / nx;
                ix = r % nx;
            } else if (r < L * nx + (ny - 2 * L) * 2 * L) {
                r -= L * nx;
                iy = L + r / (2 * L);
                const uint64_t q = r % (2 * L);
                ix = q < L ? q : nx - 2 * L + q;
            } else {
                r -= L * nx + (ny - 2 * L) * 2 * L;
                iy = ny - L + r / nx;
                ix = r % nx;
            }
        } else {
            r -= L * layer + nmid * mid;
            iz = L + nmid + r / layer;
            r %= layer;
            iy = r / nx;
            ix = r % nx;
        }
    }
    bx[k] = static_cast<float>(w.sx + (double(ix) + 0.5) * w.s);
    by[k] = static_cast<float>(w.sy + (double(iy) + 0.5) * w.s);
    bz[k] = static_cast<float>(w.sz + (double(iz) + 0.5) * w.s);
}

// Point i of a Fibonacci sphere of n points and radius r around c.
__global__ __launch_bounds__(kBlock) void k_sphere(double cx, double cy, double cz, double r, uint32_t n,
    double golden, float* __restrict__ bx, float* __restrict__ by, float* __restrict__ bz)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const double y = 1.0 - 2.0 * (double(i) + 0.5) / double(n);
    const double rad = sqrt(fmax(0.0, 1.0 - y * y));
    const double th = golden * double(i);
    bx[i] = static_cast<float>(cx + r * (cos(th) * rad));
    by[i] = static_cast<float>(cy + r * y);
    bz[i] = static_cast<float>(cz + r * (sin(th) * rad));
}

#pragma clang fp contract(on)

struct Plan {
    uint32_t n_lattice = 0, n_walls = 0, n_bdy = 0;
    WallDev w {};
    uint32_t shell_n[8] = {};
    double shell_r[8] = {};
    int shells = 0;
};

int make_plan(reyn_ctx* ctx, const reyn_dam_break* d, Plan* p)
{
    if (!d || !(d->spacing > 0.0) || d->counts[0] == 0 || d->counts[1] == 0 || d->counts[2] == 0)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "dam break: empty lattice or non-positive spacing");
    if (d->layers > 8)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "dam break: at most 8 wall layers");
    const uint64_t nl = uint64_t(d->counts[0]) * d->counts[1] * d->counts[2];
    if (nl >= (uint64_t(1) << 31))
        return set_error(ctx, REYN_ERR_CAPACITY, "dam break: lattice of %llu points is too large",
            static_cast<unsigned long long>(nl));
    p->n_lattice = static_cast<uint32_t>(nl);
    const double s = d->spacing;
    const int L = static_cast<int>(d->layers);
    int n[3];
    for (int a = 0; a < 3; ++a) {
        const double e = std::rint((d->box[a] - 0.0) / s);
        if (!(e >= 1.0) || e > 1.0e6)
            return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "dam break: box side %d spans %g spacings", a, e);
        n[a] = static_cast<int>(e) + 2 * L;
    }
    p->w = WallDev { n[0], n[1], n[2], L, 0.0 - L * s, 0.0 - L * s, 0.0 - L * s, s };
    const uint64_t all = uint64_t(n[0]) * n[1] * n[2];
    const uint64_t inner = (n[0] > 2 * L && n[1] > 2 * L && n[2] > 2 * L)
        ? uint64_t(n[0] - 2 * L) * (n[1] - 2 * L) * (n[2] - 2 * L)
        : 0;
    uint64_t nb = L > 0 ? all - inner : 0;
    p->n_walls = static_cast<uint32_t>(nb);
    if (d->sphere) {
        for (int k = 0; k < L; ++k) {
            const double r = d->radius - (k + 0.5) * s;
            if (r <= 0.0)
                break;
            const double cnt = std::nearbyint(4.0 * M_PI * r * r / (s * s));
            p->shell_r[p->shells] = r;
            p->shell_n[p->shells] = static_cast<uint32_t>(std::max(4.0, cnt));
            nb += p->shell_n[p->shells];
            ++p->shells;
        }
    }
    if (nb >= (uint64_t(1) << 31))
        return set_error(ctx, REYN_ERR_CAPACITY, "dam break: too many boundary samples");
    p->n_bdy = static_cast<uint32_t>(nb);
    return REYN_OK;
}

} // namespace

// Generates one dam break into device arrays (fluid capacity n_lattice).
int dam_break_generate(reyn_ctx* ctx, const reyn_dam_break* d, float* x, float* y, float* z, uint32_t* n_fluid,
    float* bx, float* by, float* bz)
{
    Plan p;
    REYN_TRY(make_plan(ctx, d, &p));
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    const DamDev dd { d->counts[0], d->counts[1], d->counts[2], d->spacing, d->jitter, d->seed, d->sphere ? 1 : 0,
        d->center[0], d->center[1], d->center[2], d->radius + 0.5 * d->spacing };
    const uint32_t nbk = blocks_for(p.n_lattice);
    uint32_t* counts = nullptr;
    REYN_TRY(ctx_scratch(ctx, sizeof(uint32_t) * (size_t(nbk) + 1), reinterpret_cast<void**>(&counts)));
    k_fluid_count<<<nbk, kBlock, 0, ctx->stream>>>(dd, p.n_lattice, counts);
    REYN_LAUNCHED(ctx, "k_fluid_count");
    k_fluid_scan<<<1, 1024, 0, ctx->stream>>>(counts, nbk);
    REYN_LAUNCHED(ctx, "k_fluid_scan");
    k_fluid_write<<<nbk, kBlock, 0, ctx->stream>>>(dd, p.n_lattice, counts, x, y, z);
    REYN_LAUNCHED(ctx, "k_fluid_write");
    if (p.n_walls) {
        k_walls<<<blocks_for(p.n_walls), kBlock, 0, ctx->stream>>>(p.w, p.n_walls, bx, by, bz);
        REYN_LAUNCHED(ctx, "k_walls");
    }
    const double golden = M_PI * (3.0 - std::sqrt(5.0));
    uint32_t off = p.n_walls;
    for (int k = 0; k < p.shells; ++k) {
        k_sphere<<<blocks_for(p.shell_n[k]), kBlock, 0, ctx->stream>>>(d->center[0], d->center[1], d->center[2],
            p.shell_r[k], p.shell_n[k], golden, bx + off, by + off, bz + off);
        REYN_LAUNCHED(ctx, "k_sphere");
        off += p.shell_n[k];
    }
    uint32_t nf = 0;
    REYN_HIP(ctx, hipMemcpyAsync(&nf, counts + nbk, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *n_fluid = nf;
    return REYN_OK;
}

int dam_break_sizes(reyn_ctx* ctx, const reyn_dam_break* d, uint32_t* n_lattice, uint32_t* n_bdy)
{
    Plan p;
    REYN_TRY(make_plan(ctx, d, &p));
    *n_lattice = p.n_lattice;
    *n_bdy = p.n_bdy;
    return REYN_OK;
}

} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_dam_break_sizes(const reyn_dam_break* d, uint32_t* n_lattice, uint32_t* n_bdy)
{
    if (!d || !n_lattice || !n_bdy)
        return REYN_ERR_INVALID_ARGUMENT;
    return dam_break_sizes(nullptr, d, n_lattice, n_bdy);
}

int reyn_dam_break_generate(reyn_ctx* ctx, const reyn_dam_break* d, float* x, float* y, float* z,
    uint32_t* n_fluid, float* bx, float* by, float* bz)
{
    if (!ctx || !d || !x || !y || !z || !n_fluid || ((!bx || !by || !bz) && d->layers))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "dam break: NULL argument");
    return dam_break_generate(ctx, d, x, y, z, n_fluid, bx, by, bz);
}

} // extern "C"
