// Context, error handling, memory helpers, timers and argument checks.
#include <cstdarg>
#include <cstring>
#include <new>

#include "sph.hpp"

namespace reyn {

int set_error(reyn_ctx* ctx, int code, const char* fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

int check_hip(reyn_ctx* ctx, hipError_t e, const char* what)
{
    if (e == hipSuccess)
        return REYN_OK;
    const int code = (e == hipErrorOutOfMemory) ? REYN_ERR_OUT_OF_MEMORY : REYN_ERR_HIP;
    return set_error(ctx, code, "HIP error %d (%s) in %s", static_cast<int>(e),
        hipGetErrorString(e), what);
}

int ctx_scratch(reyn_ctx* ctx, size_t bytes, void** out)
{
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) {
            REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
            REYN_HIP(ctx, hipFree(ctx->scratch));
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        REYN_HIP(ctx, hipMalloc(&ctx->scratch, bytes));
        ctx->scratch_bytes = bytes;
    }
    *out = ctx->scratch;
    return REYN_OK;
}

int ctx_pinned(reyn_ctx* ctx, size_t bytes, void** out)
{
    if (bytes > ctx->pinned_bytes) {
        if (ctx->pinned) {
            REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
            REYN_HIP(ctx, hipHostFree(ctx->pinned));
            ctx->pinned = nullptr;
            ctx->pinned_bytes = 0;
        }
        REYN_HIP(ctx, hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
        ctx->pinned_bytes = bytes;
    }
    *out = ctx->pinned;
    return REYN_OK;
}

uint32_t prof_event(reyn_ctx* ctx)
{
    if (ctx->prof_used == ctx->prof_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) {
            ctx->prof = false; // out of events: stop profiling rather than fail the launch
            return 0;
        }
        ctx->prof_pool.push_back(e);
    }
    const uint32_t k = ctx->prof_used++;
    (void)hipEventRecord(ctx->prof_pool[k], ctx->stream);
    return k;
}

int check_grid(reyn_ctx* ctx, const reyn_grid* g, uint32_t n_expected, const char* what)
{
    if (!g)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "%s: grid is NULL", what);
    if (!g->prefix)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "%s: grid has no prefix table", what);
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || g->range < 1 || !(g->cell_size > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "%s: invalid grid geometry", what);
    if (n_expected != UINT32_MAX && g->n != n_expected)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
            "%s: grid indexes %u points but %u were given", what, g->n, n_expected);
    return REYN_OK;
}

int check_fluid(reyn_ctx* ctx, const reyn_particles* p)
{
    if (!p)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "particles are NULL");
    if (p->n > 0 && (!p->xx || !p->xy || !p->xz || !p->vx || !p->vy || !p->vz || !p->m))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "particle field pointer is NULL");
    return REYN_OK;
}

int check_bdy(reyn_ctx* ctx, const reyn_boundary* b)
{
    if (!b || b->n == 0)
        return REYN_OK;
    if (!b->xx || !b->xy || !b->xz || !b->m)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "boundary field pointer is NULL");
    if (b->grid.sorted)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT,
            "boundary grid must be cell-ordered (Resort::yes)");
    // packed 16-byte boundary records are addressed with 32-bit byte offsets
    if (b->n >= (1u << 28))
        return set_error(ctx, REYN_ERR_CAPACITY, "at most 2^28 - 1 boundary samples per call (got %u)", b->n);
    return check_grid(ctx, &b->grid, b->n, "boundary");
}

int check_sph(reyn_ctx* ctx, const reyn_sph_params* prm)
{
    if (!prm)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "solver parameters are NULL");
    if (prm->kernel < 0 || prm->kernel > 3)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "unknown kernel kind %d", prm->kernel);
    if (!(prm->h_bar > 0.f) || !(prm->h > 0.f) || !(prm->rho_0 > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "h_bar, h and rho_0 must be positive");
    return REYN_OK;
}

namespace {
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}
} // namespace

} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_abi_version(void) { return REYN_ABI_VERSION; }

int reyn_device_count(int* count)
{
    if (!count)
        return REYN_ERR_INVALID_ARGUMENT;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess)
        c = 0;
    *count = c;
    return REYN_OK;
}

int reyn_device_memory(reyn_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes)
{
    if (!ctx || !free_bytes || !total_bytes)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "device memory: NULL argument");
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    REYN_HIP(ctx, hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return REYN_OK;
}

int reyn_ctx_create(int device, reyn_ctx** out)
{
    if (!out)
        return REYN_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return REYN_ERR_HIP;
    if (device < 0 || device >= count)
        return REYN_ERR_INVALID_ARGUMENT;
    reyn_ctx* ctx = new (std::nothrow) reyn_ctx();
    if (!ctx)
        return REYN_ERR_OUT_OF_MEMORY;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess
        || hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess) {
        reyn_ctx_destroy(ctx);
        return REYN_ERR_HIP;
    }
    *out = ctx;
    return REYN_OK;
}

int reyn_ctx_destroy(reyn_ctx* ctx)
{
    if (!ctx)
        return REYN_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    if (ctx->scratch)
        (void)hipFree(ctx->scratch);
    if (ctx->pinned)
        (void)hipHostFree(ctx->pinned);
    for (hipEvent_t e : ctx->prof_pool)
        (void)hipEventDestroy(e);
    if (ctx->ev_start)
        (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop)
        (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return REYN_OK;
}

const char* reyn_last_error(const reyn_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

void* reyn_ctx_stream(reyn_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int reyn_ctx_synchronize(reyn_ctx* ctx)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return REYN_OK;
}

int reyn_malloc(reyn_ctx* ctx, void** ptr, size_t bytes)
{
    if (!ctx || !ptr)
        return REYN_ERR_INVALID_ARGUMENT;
    *ptr = nullptr;
    if (bytes == 0)
        bytes = 4;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    REYN_HIP(ctx, hipMalloc(ptr, bytes));
    return REYN_OK;
}

int reyn_free(reyn_ctx* ctx, void* ptr)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    if (!ptr)
        return REYN_OK;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    REYN_HIP(ctx, hipFree(ptr));
    return REYN_OK;
}

int reyn_memcpy(reyn_ctx* ctx, void* dst, const void* src, size_t bytes, int kind)
{
    if (!ctx || (bytes && (!dst || !src)))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "memcpy: NULL pointer");
    if (bytes == 0)
        return REYN_OK;
    hipMemcpyKind k;
    switch (kind) {
    case 0:
        k = hipMemcpyHostToDevice;
        break;
    case 1:
        k = hipMemcpyDeviceToHost;
        break;
    case 2:
        k = hipMemcpyDeviceToDevice;
        break;
    default:
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "memcpy: bad kind %d", kind);
    }
    REYN_HIP(ctx, hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return REYN_OK;
}

int reyn_fill_u32(reyn_ctx* ctx, uint32_t* ptr, uint32_t value, size_t count)
{
    if (!ctx || (count && !ptr))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "fill: NULL pointer");
    if (count == 0)
        return REYN_OK;
    const size_t nb = (count + kBlock - 1) / kBlock;
    k_fill_u32<<<dim3(static_cast<uint32_t>(nb)), kBlock, 0, ctx->stream>>>(ptr, value, count);
    REYN_LAUNCHED(ctx, "fill_u32");
    return REYN_OK;
}

int reyn_fill_f32(reyn_ctx* ctx, float* ptr, float value, size_t count)
{
    if (!ctx || (count && !ptr))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "fill: NULL pointer");
    if (count == 0)
        return REYN_OK;
    const size_t nb = (count + kBlock - 1) / kBlock;
    k_fill_f32<<<dim3(static_cast<uint32_t>(nb)), kBlock, 0, ctx->stream>>>(ptr, value, count);
    REYN_LAUNCHED(ctx, "fill_f32");
    return REYN_OK;
}

int reyn_profile_enable(reyn_ctx* ctx, int on)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prof_recs.clear();
    ctx->prof_used = 0;
    if (on && ctx->prof_pool.size() < 4096) {
        while (ctx->prof_pool.size() < 4096) {
            hipEvent_t e = nullptr;
            REYN_HIP(ctx, hipEventCreate(&e));
            ctx->prof_pool.push_back(e);
        }
    }
    ctx->prof = on != 0;
    return REYN_OK;
}

int reyn_profile_select(reyn_ctx* ctx, const char* name)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    ctx->prof_only = name ? name : "";
    return REYN_OK;
}

int reyn_profile_summary(reyn_ctx* ctx, char* buf, size_t len, size_t* needed)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // per kernel name, in first-seen order: every launch's duration
    std::vector<const char*> names;
    std::vector<std::vector<float>> times;
    for (const reyn_prof_rec& r : ctx->prof_recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->prof_pool[r.a], ctx->prof_pool[r.b]) != hipSuccess)
            continue;
        size_t k = 0;
        while (k < names.size() && std::strcmp(names[k], r.name) != 0)
            ++k;
        if (k == names.size()) {
            names.push_back(r.name);
            times.emplace_back();
        }
        times[k].push_back(ms);
    }
    std::string out;
    char line[256];
    for (size_t k = 0; k < names.size(); ++k) {
        double total = 0.0, work = 0.0;
        float mx = 0.f;
        for (float ms : times[k]) {
            total += ms;
            mx = std::max(mx, ms);
        }
        uint32_t nwork = 0;
        for (float ms : times[k])
            if (ms >= 0.2f * mx) {
                work += ms;
                ++nwork;
            }
        snprintf(line, sizeof(line), "%s %zu %.6f %u %.6f\n", names[k], times[k].size(), total, nwork, work);
        out += line;
    }
    if (needed)
        *needed = out.size() + 1;
    if (buf && len > 0) {
        const size_t m = std::min(len - 1, out.size());
        std::memcpy(buf, out.data(), m);
        buf[m] = '\0';
    }
    return REYN_OK;
}

int reyn_timer_start(reyn_ctx* ctx)
{
    if (!ctx)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    return REYN_OK;
}

int reyn_timer_stop(reyn_ctx* ctx, float* ms)
{
    if (!ctx || !ms)
        return REYN_ERR_INVALID_ARGUMENT;
    REYN_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    REYN_HIP(ctx, hipEventSynchronize(ctx->ev_stop));
    REYN_HIP(ctx, hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
    return REYN_OK;
}

} // extern "C"
