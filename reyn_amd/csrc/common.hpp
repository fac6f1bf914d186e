// Shared device/host definitions of the gfx950 SPH library: vector algebra,
// the context object, error handling and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/reyn_amd.h"

namespace reyn {

// ---- small vector algebra -------------------------------------------------
// Component-wise float3 algebra with the same operation order as the
// reference's `vector.cuh` where the order decides bits: the dot product is
// fmaf(a.x,b.x, fmaf(a.y,b.y, a.z*b.z)) (src/utils/vector.cuh:152-155).

struct f3 {
    float x, y, z;
};

__host__ __device__ __forceinline__ f3 mk(float x, float y, float z) { return f3 { x, y, z }; }
__host__ __device__ __forceinline__ f3 operator+(f3 a, f3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__host__ __device__ __forceinline__ f3 operator-(f3 a, f3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__host__ __device__ __forceinline__ f3 operator-(f3 a) { return { -a.x, -a.y, -a.z }; }
__host__ __device__ __forceinline__ f3 operator*(float s, f3 a) { return { s * a.x, s * a.y, s * a.z }; }
__host__ __device__ __forceinline__ f3 operator*(f3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
__host__ __device__ __forceinline__ f3& operator+=(f3& a, f3 b)
{
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
__host__ __device__ __forceinline__ float dot(f3 a, f3 b)
{
    return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
}
__device__ __forceinline__ f3 ld3(const float* __restrict__ x, const float* __restrict__ y,
    const float* __restrict__ z, uint32_t i)
{
    return { x[i], y[i], z[i] };
}

// Gathers through a 32-bit byte offset from a uniform base: the compiler
// emits the scalar-base form of global loads (one offset VGPR, no 64-bit
// address arithmetic). Valid while the byte offset fits in 32 bits, which the
// entry points check (arrays below 4 GiB).
template <class T> __device__ __forceinline__ T gat(const T* __restrict__ base, uint32_t i)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + static_cast<uint32_t>(i * sizeof(T)));
}
// gat with the non-temporal hint: data streamed once per pass, so it does not
// push the gathered records out of the caches.
template <class T> __device__ __forceinline__ T gat_nt(const T* __restrict__ base, uint32_t i)
{
    return __builtin_nontemporal_load(
        reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + static_cast<uint32_t>(i * sizeof(T))));
}
// The same through an explicit byte offset (for records of mixed layout).
template <class T> __device__ __forceinline__ T gatb(const void* __restrict__ base, uint32_t byte_off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ f3 gat3(const float* __restrict__ x, const float* __restrict__ y,
    const float* __restrict__ z, uint32_t i)
{
    return { gat(x, i), gat(y, i), gat(z, i) };
}
__device__ __forceinline__ void st3(float* __restrict__ x, float* __restrict__ y,
    float* __restrict__ z, uint32_t i, f3 v)
{
    x[i] = v.x;
    y[i] = v.y;
    z[i] = v.z;
}

// Neumaier-compensated float accumulator: the running sum plus a correction
// term that collects the low-order bits lost by each addition.
struct ksum {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v)
    {
        const float t = s + v;
        c += (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};
struct ksum3 {
    ksum x, y, z;
    __device__ __forceinline__ void add(f3 v)
    {
        x.add(v.x);
        y.add(v.y);
        z.add(v.z);
    }
    __device__ __forceinline__ f3 value() const { return { x.value(), y.value(), z.value() }; }
};

// Neumaier's accumulator for sums whose running value is never smaller in
// magnitude than the next term (non-negative terms, the largest first): ksum's
// branch then always takes (s - t) + v, the exact rounding error of s + v
// (Fast2Sum), and for the first term, added to zero, both forms give zero; so
// this gives ksum's bits with half the instructions. Where a later term does
// exceed the running sum the correction is no longer exact (the sum stays
// compensated, to a looser bound).
struct ksum_desc {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v)
    {
        const float t = s + v;
        c += (s - t) + v;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

// ---- launch geometry --------------------------------------------------------

// lanes per workgroup: 4 wave64 (builds for measurement may set REYN_BLOCK
// to another multiple of 64)
#ifndef REYN_BLOCK
#define REYN_BLOCK 256
#endif
constexpr int kBlock = REYN_BLOCK;
constexpr int kWave = 64;
constexpr uint32_t kXcds = 8; // MI355X: 8 XCDs, workgroups dispatched round-robin

__host__ __device__ __forceinline__ uint32_t ceil_div_u(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Workgroup b of nb is dispatched to XCD b % 8. Remap so that each XCD works
// on one contiguous range of logical blocks: particles stored in cell order
// then share their neighbours' cache lines inside one XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb / kXcds, r = nb % kXcds;
    const uint32_t xcd = b % kXcds, k = b / kXcds;
    const uint32_t base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
    return base + k;
}
__device__ __forceinline__ uint32_t global_index()
{
    return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// ---- wave64 and block reductions -------------------------------------------

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kWave)));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kWave)));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, kWave);
    return v;
}

// acc = op(acc, get(k)) for k = threadIdx.x, + blockDim.x, ... below n, in
// that order, with U loads in flight at a time (a one-workgroup reduction of
// thousands of partials otherwise waits for each load in turn).
template <int U, class T, class G, class Op>
__device__ __forceinline__ T strided_fold(uint32_t n, T acc, G&& get, Op&& op)
{
    const uint32_t stride = blockDim.x;
    uint32_t k = threadIdx.x;
    for (; k + (U - 1) * stride < n; k += U * stride) {
        decltype(get(k)) v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[u] = get(k + u * stride);
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc = op(acc, v[u]);
    }
    for (; k < n; k += stride)
        acc = op(acc, get(k));
    return acc;
}

// Block-wide reduction in a fixed order (waves combined by wave 0), so that a
// block's partial result does not depend on scheduling. op: 0 sum, 1 min, 2 max
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* lds)
{
    if (OP == 0)
        v = wave_sum(v);
    else if (OP == 1)
        v = wave_min(v);
    else
        v = wave_max(v);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0)
        lds[w] = v;
    __syncthreads();
    double r = lds[0];
    const int nw = blockDim.x / kWave;
    for (int k = 1; k < nw; ++k) {
        if (OP == 0)
            r += lds[k];
        else if (OP == 1)
            r = fmin(r, lds[k]);
        else
            r = fmax(r, lds[k]);
    }
    __syncthreads();
    return r;
}

} // namespace reyn

// ---- the context ------------------------------------------------------------

struct reyn_prof_rec {
    const char* name;
    uint32_t a, b; // indices into the event pool
};

struct reyn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::string last_error;
    // per-launch event timing, enabled by reyn_profile_enable
    bool prof = false;
    std::string prof_only; // record only this kernel's launches (empty: all)
    std::vector<hipEvent_t> prof_pool;
    uint32_t prof_used = 0;
    std::vector<reyn_prof_rec> prof_recs;
    // iterations of the last single-scene PCISPH [0] and IISPH [1] solve on
    // this context: the first chunk of the next solve's loop
    uint32_t loop_hint[2] = { 0, 0 };
    // scratch owned by the context, grown on demand and freed with it
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // pinned host staging for device -> host results
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
};

namespace reyn {

int set_error(reyn_ctx* ctx, int code, const char* fmt, ...);
int check_hip(reyn_ctx* ctx, hipError_t e, const char* what);
// device scratch of at least `bytes`, aligned to 256 bytes
int ctx_scratch(reyn_ctx* ctx, size_t bytes, void** out);
int ctx_pinned(reyn_ctx* ctx, size_t bytes, void** out);

#define REYN_TRY(expr)                                                          \
    do {                                                                        \
        const int _rc = (expr);                                                 \
        if (_rc != REYN_OK)                                                     \
            return _rc;                                                         \
    } while (0)

#define REYN_HIP(ctx, expr) REYN_TRY(::reyn::check_hip((ctx), (expr), #expr))

#define REYN_LAUNCHED(ctx, what) REYN_TRY(::reyn::check_hip((ctx), hipGetLastError(), (what)))

inline uint32_t blocks_for(uint32_t n, uint32_t bs = kBlock)
{
    const uint32_t b = ceil_div_u(n, bs);
    return b == 0 ? 1 : b;
}

// Records a start event before and a stop event after the launches in its
// scope when profiling is on (a no-op otherwise).
uint32_t prof_event(reyn_ctx* ctx);
struct ProfScope {
    reyn_ctx* ctx;
    const char* name;
    uint32_t a = 0;
    ProfScope(reyn_ctx* c, const char* n)
        : ctx(c)
        , name(n)
    {
        if (ctx->prof && (ctx->prof_only.empty() || ctx->prof_only == name))
            a = prof_event(ctx);
        else
            name = nullptr;
    }
    ~ProfScope()
    {
        if (name && ctx->prof) {
            const uint32_t b = prof_event(ctx);
            ctx->prof_recs.push_back({ name, a, b });
        }
    }
};

} // namespace reyn
