// Kernel evaluation, CFL time step, bounds enforcement, Akinci boundary
// masses and FP64 reductions.
#include <cmath>

#include "sph.hpp"

namespace reyn {
namespace {

template <int K>
__global__ __launch_bounds__(kBlock) void k_kernel_eval(KParams kp, const float* __restrict__ dx,
    const float* __restrict__ dy, const float* __restrict__ dz, uint32_t n, float* __restrict__ w,
    float* __restrict__ gx, float* __restrict__ gy, float* __restrict__ gz)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const f3 x { dx[i], dy[i], dz[i] };
    const float d2 = dot(x, x);
    if (w)
        w[i] = kW<K>(kp, d2);
    const f3 g = kdW<K>(kp, x, d2);
    if (gx)
        gx[i] = g.x;
    if (gy)
        gy[i] = g.y;
    if (gz)
        gz[i] = g.z;
}

// per-block max |v| in FP64
__global__ __launch_bounds__(kBlock) void k_vmax(const float* __restrict__ vx,
    const float* __restrict__ vy, const float* __restrict__ vz, uint32_t n, double* __restrict__ part)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double a = vx[i], b = vy[i], c = vz[i];
        v = sqrt(fma(a, a, fma(b, b, c * c)));
        if (v != v) // a NaN velocity makes the time step non-finite
            v = INFINITY;
    }
    v = block_reduce<2>(v, lds);
    if (threadIdx.x == 0)
        part[blockIdx.x] = v;
}

// generic FP64 reduction of a float array: per-block partials
template <int OP>
__global__ __launch_bounds__(kBlock) void k_reduce_part(const float* __restrict__ x, uint32_t n,
    double* __restrict__ part)
{
    __shared__ double lds[kBlock / kWave];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double neutral = OP == 0 ? 0.0 : (OP == 1 ? INFINITY : -INFINITY);
    double v = neutral;
    if (i < n) {
        v = x[i];
        if (v != v) // NaN poisons every reduction
            v = NAN;
    }
    v = block_reduce<OP>(v, lds);
    if (threadIdx.x == 0)
        part[blockIdx.x] = v;
}

template <int OP>
__global__ __launch_bounds__(kBlock) void k_reduce_final(const double* __restrict__ part, uint32_t nb,
    double* __restrict__ out)
{
    __shared__ double lds[kBlock / kWave];
    const double neutral = OP == 0 ? 0.0 : (OP == 1 ? INFINITY : -INFINITY);
    // x: the fold, y: 1 once a NaN was seen
    const double2 r = strided_fold<8>(
        nb, make_double2(neutral, 0.0), [&](uint32_t k) { return part[k]; },
        [](double2 a, double p) {
            const double f = OP == 0 ? a.x + p : (OP == 1 ? fmin(a.x, p) : fmax(a.x, p));
            return make_double2(f, (p != p) ? 1.0 : a.y);
        });
    double v = block_reduce<OP>(r.x, lds);
    const double anynan = block_reduce<2>(r.y, lds);
    if (threadIdx.x == 0)
        out[0] = anynan > 0.0 ? NAN : v;
}

__global__ __launch_bounds__(kBlock) void k_bounds(f3 lo, f3 hi, FluidDev f)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= f.n)
        return;
    const f3 x = ld3(f.x, f.y, f.z, i);
    // any component on or beyond a face reflects and damps the velocity
    // (`operator<=` is an OR over components, src/utils/vector.cuh:332-342)
    if (x.x <= lo.x || x.y <= lo.y || x.z <= lo.z || hi.x <= x.x || hi.y <= x.y || hi.z <= x.z) {
        const float c = static_cast<float>(-0.1);
        st3(f.vx, f.vy, f.vz, i, c * ld3(f.vx, f.vy, f.vz, i));
    }
    const f3 cl { fmaxf(fminf(x.x, hi.x), lo.x), fmaxf(fminf(x.y, hi.y), lo.y), fmaxf(fminf(x.z, hi.z), lo.z) };
    // only clamped positions are written back (the same values; inside the
    // box nothing changes and nothing is stored)
    if (!(cl.x == x.x && cl.y == x.y && cl.z == x.z))
        st3(f.x, f.y, f.z, i, cl);
}

// one Jacobi refinement of Akinci masses: m_new = rho0 m_old / sum_j m_old_j W_bj
__global__ __launch_bounds__(kBlock) void k_bdy_mass(GridDev g, KParams kp, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n,
    const float* __restrict__ m_old, float* __restrict__ m_new, float rho0)
{
    const uint32_t i = global_index();
    if (i >= n)
        return;
    const f3 xi = ld3(xs, ys, zs, i);
    ksum s;
    walk<false>(g, xi, xs, ys, zs, [&](uint32_t j, f3, float d2) { s.add(m_old[j] * kW<KB3>(kp, d2)); });
    m_new[i] = rho0 * m_old[i] / s.value();
}

int reduce(reyn_ctx* ctx, const float* x, uint32_t n, int op, double* out_dev, double* part)
{
    const uint32_t nb = blocks_for(n);
    if (op < 0 || op > 3)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "reduce: unknown op %d", op);
    ProfScope _prof(ctx, "k_reduce");
    if (op == 0 || op == 3) {
        k_reduce_part<0><<<nb, kBlock, 0, ctx->stream>>>(x, n, part);
        k_reduce_final<0><<<1, kBlock, 0, ctx->stream>>>(part, nb, out_dev);
    } else if (op == 1) {
        k_reduce_part<1><<<nb, kBlock, 0, ctx->stream>>>(x, n, part);
        k_reduce_final<1><<<1, kBlock, 0, ctx->stream>>>(part, nb, out_dev);
    } else {
        k_reduce_part<2><<<nb, kBlock, 0, ctx->stream>>>(x, n, part);
        k_reduce_final<2><<<1, kBlock, 0, ctx->stream>>>(part, nb, out_dev);
    }
    REYN_LAUNCHED(ctx, "k_reduce");
    return REYN_OK;
}

} // namespace
} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_kernel_eval(reyn_ctx* ctx, int32_t kernel, float h_bar, const float* dx, const float* dy,
    const float* dz, uint32_t n, float* w, float* gx, float* gy, float* gz)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    if (n && (!dx || !dy || !dz))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "kernel_eval: NULL offsets");
    if (!(h_bar > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "kernel_eval: h_bar must be positive");
    if (n == 0)
        return REYN_OK;
    const KParams kp = make_kparams(kernel, h_bar);
    return with_kernel(ctx, kernel, false, [&](auto KC, auto) -> int {
        constexpr int K = decltype(KC)::value;
        {
            ProfScope _prof(ctx, "k_kernel_eval");
            k_kernel_eval<K><<<blocks_for(n), kBlock, 0, ctx->stream>>>(kp, dx, dy, dz, n, w, gx, gy, gz);
        }
        REYN_LAUNCHED(ctx, "k_kernel_eval");
        return REYN_OK;
    });
}

int reyn_cfl_time_step(reyn_ctx* ctx, float lambda, float h, const reyn_particles* fluid,
    reyn_float3 g, float* dt, double* v_max)
{
    if (!ctx || !dt)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "cfl: NULL output");
    REYN_TRY(check_fluid(ctx, fluid));
    const uint32_t n = fluid->n;
    const uint32_t nb = blocks_for(n);
    double* scratch = nullptr;
    REYN_TRY(ctx_scratch(ctx, (size_t(nb) + 8) * sizeof(double), reinterpret_cast<void**>(&scratch)));
    double* host = nullptr;
    REYN_TRY(ctx_pinned(ctx, 64, reinterpret_cast<void**>(&host)));
    double vmax = 0.0;
    if (n > 0) {
        {
            ProfScope _prof(ctx, "k_vmax");
            k_vmax<<<nb, kBlock, 0, ctx->stream>>>(fluid->vx, fluid->vy, fluid->vz, n, scratch + 8);
        }
        REYN_LAUNCHED(ctx, "k_vmax");
        {
            ProfScope _prof(ctx, "k_reduce_final");
            k_reduce_final<2><<<1, kBlock, 0, ctx->stream>>>(scratch + 8, nb, scratch);
        }
        REYN_LAUNCHED(ctx, "k_reduce_final");
        REYN_HIP(ctx, hipMemcpyAsync(host, scratch, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
        vmax = host[0];
    }
    // (src/timestep/cfl.cu:5-37)
    const float dt_vel = static_cast<float>(double(lambda * h) / vmax);
    const double gn = std::sqrt(double(std::fma(g.x, g.x, std::fma(g.y, g.y, g.z * g.z))));
    const float dt_g = std::sqrt(static_cast<float>(2.0 * double(lambda) * double(h) / gn));
    *dt = std::fmin(dt_vel, dt_g);
    if (v_max)
        *v_max = vmax;
    if (!std::isfinite(vmax))
        return set_error(ctx, REYN_ERR_NOT_FINITE, "cfl: velocity field is not finite");
    return REYN_OK;
}

int reyn_simple_dt_controller(reyn_ctx* ctx, uint32_t iter_last, float dt_last, float lambda,
    float h, const reyn_particles* fluid, reyn_float3 g, uint32_t iter_target, float change_factor,
    float* dt)
{
    if (!dt)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "dt controller: NULL output");
    // (src/timestep/cfl.cu:39-52)
    const float factor = iter_last < iter_target ? 1.f + change_factor : 1.f - change_factor;
    const float suggested = dt_last * factor;
    float cfl = 0.f;
    REYN_TRY(reyn_cfl_time_step(ctx, lambda, h, fluid, g, &cfl, nullptr));
    *dt = std::fmin(cfl, suggested);
    return REYN_OK;
}

int reyn_hard_enforce_bounds(reyn_ctx* ctx, reyn_float3 bmin, reyn_float3 bmax, const reyn_particles* fluid)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_TRY(check_fluid(ctx, fluid));
    if (fluid->n == 0)
        return REYN_OK;
    {
        ProfScope _prof(ctx, "k_bounds");
        k_bounds<<<blocks_for(fluid->n), kBlock, 0, ctx->stream>>>(f3 { bmin.x, bmin.y, bmin.z },
            f3 { bmax.x, bmax.y, bmax.z }, to_dev(*fluid));
    }
    REYN_LAUNCHED(ctx, "k_bounds");
    return REYN_OK;
}

int reyn_boundary_masses(reyn_ctx* ctx, const reyn_boundary* bdy, float* m, float h, float rho_0,
    uint32_t iterations)
{
    if (!ctx || !bdy || !m)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "boundary_masses: NULL argument");
    if (iterations == 0)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
            "At least one iteration of mass refinements must be run");
    if (bdy->grid.sorted)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "boundary grid must be cell-ordered");
    REYN_TRY(check_grid(ctx, &bdy->grid, bdy->n, "boundary"));
    const uint32_t n = bdy->n;
    if (n == 0)
        return REYN_OK;
    float* tmp = nullptr;
    REYN_TRY(ctx_scratch(ctx, size_t(n) * sizeof(float), reinterpret_cast<void**>(&tmp)));
    const KParams kp = make_kparams(KB3, 2.f * h);
    const GridDev g = to_dev(bdy->grid);
    REYN_TRY(reyn_fill_f32(ctx, m, 1.f, n));
    // ping-pong between m and scratch; an odd count ends in scratch and is copied back
    float* src = m;
    float* dst = tmp;
    for (uint32_t it = 0; it < iterations; ++it) {
        {
            ProfScope _prof(ctx, "k_bdy_mass");
            k_bdy_mass<<<blocks_for(n), kBlock, 0, ctx->stream>>>(g, kp, bdy->xx, bdy->xy, bdy->xz, n, src, dst, rho_0);
        }
        REYN_LAUNCHED(ctx, "k_bdy_mass");
        std::swap(src, dst);
    }
    if (src != m)
        REYN_HIP(ctx, hipMemcpyAsync(m, src, size_t(n) * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return REYN_OK;
}

int reyn_reduce_f32(reyn_ctx* ctx, const float* x, uint32_t n, int32_t op, double* out)
{
    if (!ctx || !out || (n && !x))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "reduce: NULL argument");
    if (op < 0 || op > 3)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "reduce: unknown op %d", op);
    if (n == 0) {
        *out = (op == 0) ? 0.0 : NAN;
        return REYN_OK;
    }
    const uint32_t nb = blocks_for(n);
    double* scratch = nullptr;
    REYN_TRY(ctx_scratch(ctx, (size_t(nb) + 8) * sizeof(double), reinterpret_cast<void**>(&scratch)));
    double* host = nullptr;
    REYN_TRY(ctx_pinned(ctx, 64, reinterpret_cast<void**>(&host)));
    REYN_TRY(reduce(ctx, x, n, op, scratch, scratch + 8));
    REYN_HIP(ctx, hipMemcpyAsync(host, scratch, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = (op == 3) ? host[0] / double(n) : host[0];
    return REYN_OK;
}

} // extern "C"
