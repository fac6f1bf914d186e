// Scene set-up on the device: the lattice box of fluid positions, the
// normally distributed jitter, the cull of fluid points too close to the
// boundary samples, and the stable compaction that removes them. The C++
// `reyn::Scene` builds the reference's two scene kinds from these
// (src/scene/scene.cu:14-54 `_init_box_kernel`, `GenJitter`, 55-83
// `CullPredicate`, 183-206 `cull_particles_by_predicate`, 239-464 `Scene`).
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "sph.hpp"

namespace reyn {
namespace {

#pragma clang fp contract(off)

__device__ inline uint32_t smix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// U(0, 1] with 24 bits from (seed, point, draw)
__device__ inline float open_uniform(uint32_t seed, uint32_t idx, uint32_t draw)
{
    const uint32_t h = smix32(smix32(seed * 0x9e3779b9u + 0x7f4a7c15u) ^ (idx * 6u + draw));
    return float((h >> 8) + 1u) * (1.f / 16777216.f);
}

// Point i of an nx x ny x nz lattice: origin + (ix h, iy h, iz h), x fastest
// (the index order and float operations of `_init_box_kernel`).
__global__ __launch_bounds__(kBlock) void k_lattice(float ox, float oy, float oz, uint32_t nx, uint32_t ny,
    uint32_t n, float h, float* __restrict__ xs, float* __restrict__ ys, float* __restrict__ zs)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    uint32_t ix = i;
    const uint32_t iz = ix / (nx * ny);
    ix -= iz * nx * ny;
    const uint32_t iy = ix / nx;
    ix -= iy * nx;
    xs[i] = ox + float(ix) * h;
    ys[i] = oy + float(iy) * h;
    zs[i] = oz + float(iz) * h;
}

// x += sigma N(0, 1) per axis, the normal deviates by the Box-Muller transform
// of two hashed uniforms per axis.
__global__ __launch_bounds__(kBlock) void k_jitter(float* __restrict__ xs, float* __restrict__ ys,
    float* __restrict__ zs, uint32_t n, float sigma, uint32_t seed)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    float* a[3] = { xs, ys, zs };
    for (uint32_t k = 0; k < 3; ++k) {
        const float u1 = open_uniform(seed, i, 2 * k), u2 = open_uniform(seed, i, 2 * k + 1);
        const float z = sqrtf(-2.f * logf(u1)) * cosf(6.283185307179586f * u2);
        a[k][i] = a[k][i] + sigma * z;
    }
}

// keep[k] = 0 when a boundary sample lies within `r2` (squared) of point k.
__global__ __launch_bounds__(kBlock) void k_cull(BdyDev b, float r2, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n, uint8_t* __restrict__ keep)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n)
        return;
    bool near = false;
    walk<false>(b.grid, ld3(xs, ys, zs, k), b.x, b.y, b.z, [&](uint32_t, f3, float d2) { near |= d2 <= r2; });
    keep[k] = near ? 0 : 1;
}

} // namespace
} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_lattice_box(reyn_ctx* ctx, reyn_float3 origin, uint32_t nx, uint32_t ny, uint32_t nz, float h,
    float* xs, float* ys, float* zs)
{
    const uint64_t n = uint64_t(nx) * ny * nz;
    if (!ctx || (n && (!xs || !ys || !zs)) || n > 0xffffffffull)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lattice box: bad argument");
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    k_lattice<<<blocks_for(uint32_t(n)), kBlock, 0, ctx->stream>>>(origin.x, origin.y, origin.z, nx, ny,
        uint32_t(n), h, xs, ys, zs);
    REYN_HIP(ctx, hipGetLastError());
    return REYN_OK;
}

int reyn_jitter_normal(reyn_ctx* ctx, float* xs, float* ys, float* zs, uint32_t n, float sigma, uint32_t seed)
{
    if (!ctx || (n && (!xs || !ys || !zs)))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "jitter: NULL argument");
    if (n == 0 || sigma == 0.f)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    k_jitter<<<blocks_for(n), kBlock, 0, ctx->stream>>>(xs, ys, zs, n, sigma, seed);
    REYN_HIP(ctx, hipGetLastError());
    return REYN_OK;
}

int reyn_boundary_cull(reyn_ctx* ctx, const reyn_boundary* b, float radius, const float* xs, const float* ys,
    const float* zs, uint32_t n, uint8_t* keep)
{
    if (!ctx || !b || (n && (!xs || !ys || !zs || !keep)))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "boundary cull: NULL argument");
    if (!(radius >= 0.f) || radius * radius > b->grid.r_c_2)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
            "boundary cull: radius %g beyond the boundary grid's search radius", double(radius));
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    const BdyDev d = to_dev(b);
    if (!d.present) { // no samples: nothing is culled
        REYN_HIP(ctx, hipMemsetAsync(keep, 1, n, ctx->stream));
        return REYN_OK;
    }
    k_cull<<<blocks_for(n), kBlock, 0, ctx->stream>>>(d, radius * radius, xs, ys, zs, n, keep);
    REYN_HIP(ctx, hipGetLastError());
    return REYN_OK;
}

int reyn_compact_f32(reyn_ctx* ctx, const uint8_t* keep, uint32_t n, float* const* fields, uint32_t nfields,
    uint32_t* count)
{
    if (!ctx || !count || nfields == 0 || !fields || (n && !keep))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "compact: NULL argument or no field");
    *count = 0;
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    size_t tbytes = 0;
    REYN_HIP(ctx, hipcub::DeviceSelect::Flagged(nullptr, tbytes, static_cast<const float*>(nullptr), keep,
                      static_cast<float*>(nullptr), static_cast<uint32_t*>(nullptr), int(n), ctx->stream));
    const size_t al = (tbytes + 255) & ~size_t(255);
    char* mem = nullptr;
    REYN_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&mem), al + 256 + sizeof(float) * size_t(n)));
    uint32_t* dcount = reinterpret_cast<uint32_t*>(mem + al);
    float* out = reinterpret_cast<float*>(mem + al + 256);
    int rc = REYN_OK;
    uint32_t kept = n;
    for (uint32_t f = 0; f < nfields && rc == REYN_OK; ++f) {
        rc = check_hip(ctx,
            hipcub::DeviceSelect::Flagged(mem, tbytes, fields[f], keep, out, dcount, int(n), ctx->stream),
            "select");
        if (rc == REYN_OK)
            rc = check_hip(ctx,
                hipMemcpyAsync(fields[f], out, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream), "copy");
    }
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMemcpyAsync(&kept, dcount, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
            "copy");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(mem);
    if (rc == REYN_OK)
        *count = kept;
    return rc;
}

} // extern "C"
