// reyn C++ host API over the gfx950 C-ABI (include/reyn_amd.h).
//
// Header-only, C++20. It restores the shapes of the reference's host classes
// so code written against reyn's API compiles against this one:
//
//   DeviceBuffer<T>        src/buffer.cuh:15-269
//   Particles              src/particles.cuh:25-75
//   Kernel<D>, C2/B3/W6/COS, IsKernel   src/kernels.cuh:19-174
//   Resort, UniformGrid<R>, UniformGridBuilder   src/datastructure/uniformgrid.cuh
//   Boundary, BoundarySamples, Scene bounds/grid  src/scene/scene.cuh:14-156
//   SESPH<K,R>, PCISPH<K,R>, IISPH<K,R>  src/solvers/*.cuh
//   cfl_time_step, simple_dt_controller  src/timestep/cfl.cuh:13-32
//   Mesh, sample_mesh                    src/scene/loader.h, sample_boundary.cu:391-440
//   LBVH (project, inside)               src/datastructure/lbvh.cuh:614-758, scene.cu:88-140
//   save_to_ply                          src/scene/ply_io.cu:7-56
//   StepTrace: the GUI's dt / iteration plots as data   src/gui.cu:956-957
//   SceneBatch: many scenes per launch (no reference counterpart)
//
// Device errors and misuse throw std::runtime_error (the reference throws on
// misuse and exits on CUDA errors, src/common.h:24-36; a library must not
// exit). All work runs on the stream of one `Context` per host thread.
#pragma once

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdlib>
#include <concepts>
#include <cstring>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../reyn_amd.h"

namespace reyn {

using float3 = reyn_float3;
inline float3 v3(float x, float y, float z) { return float3 { x, y, z }; }
inline float3 v3(float x) { return float3 { x, x, x }; }

// ---- context ----------------------------------------------------------------

class Context {
public:
    explicit Context(int device = 0)
    {
        const int rc = reyn_ctx_create(device, &ctx_);
        if (rc != REYN_OK)
            throw std::runtime_error("reyn: cannot create a context on HIP device " + std::to_string(device));
    }
    ~Context() { reyn_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    reyn_ctx* get() const { return ctx_; }
    void check(int rc, const char* what) const
    {
        if (rc != REYN_OK)
            throw std::runtime_error(std::string("reyn: ") + what + ": " + reyn_last_error(ctx_) + " (code "
                + std::to_string(rc) + ")");
    }
    void synchronize() const { check(reyn_ctx_synchronize(ctx_), "synchronize"); }

    // the context used by objects constructed without an explicit one
    static Context& current()
    {
        thread_local std::unique_ptr<Context> c;
        if (!c)
            c = std::make_unique<Context>(0);
        return *c;
    }

private:
    reyn_ctx* ctx_ = nullptr;
};

// ---- DeviceBuffer<T> (src/buffer.cuh:15-269) ------------------------------------

template <typename T> class DeviceBuffer {
    static_assert(std::is_trivially_copyable_v<T>, "DeviceBuffer holds trivially copyable types");

public:
    explicit DeviceBuffer(size_t n = 1, Context& ctx = Context::current())
        : ctx_(&ctx)
    {
        alloc(n);
    }
    DeviceBuffer(size_t n, const T& value, Context& ctx = Context::current())
        : ctx_(&ctx)
    {
        alloc(n);
        fill(value);
    }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept
        : ctx_(o.ctx_)
        , ptr_(std::exchange(o.ptr_, nullptr))
        , n_(std::exchange(o.n_, 0))
    {
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            ctx_ = o.ctx_;
            ptr_ = std::exchange(o.ptr_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }

    T* ptr() const { return ptr_; }
    size_t size() const { return n_; }
    Context& context() const { return *ctx_; }

    // contents are kept up to min(old, new) elements
    void resize(size_t n)
    {
        if (n == n_)
            return;
        T* old = ptr_;
        const size_t keep = std::min(n, n_);
        ptr_ = nullptr;
        alloc_raw(n);
        if (old && keep)
            ctx_->check(reyn_memcpy(ctx_->get(), ptr_, old, keep * sizeof(T), 2), "resize");
        if (old)
            reyn_free(ctx_->get(), old);
        n_ = n;
    }
    void resize(size_t n, const T& value)
    {
        resize(n);
        fill(value);
    }

    void fill(const T& value)
    {
        if (!n_)
            return;
        if constexpr (std::is_same_v<T, float>)
            ctx_->check(reyn_fill_f32(ctx_->get(), ptr_, value, n_), "fill");
        else if constexpr (sizeof(T) == 4) {
            uint32_t bits;
            std::memcpy(&bits, &value, 4);
            ctx_->check(reyn_fill_u32(ctx_->get(), reinterpret_cast<uint32_t*>(ptr_), bits, n_), "fill");
        } else {
            std::vector<T> h(n_, value);
            upload(h);
        }
    }
    void upload(const std::vector<T>& h)
    {
        if (h.size() != n_)
            resize(h.size());
        if (n_)
            ctx_->check(reyn_memcpy(ctx_->get(), ptr_, h.data(), n_ * sizeof(T), 0), "upload");
    }
    std::vector<T> download() const
    {
        std::vector<T> h(n_);
        if (n_)
            ctx_->check(reyn_memcpy(ctx_->get(), h.data(), ptr_, n_ * sizeof(T), 1), "download");
        return h;
    }

    // gather by `sorted` into tmp, then swap storage (buffer.cuh:184-209)
    void reorder(const DeviceBuffer<uint32_t>& sorted, DeviceBuffer<T>& tmp)
        requires(sizeof(T) == 4)
    {
        if (sorted.size() != n_)
            throw std::runtime_error("reyn: reorder map and buffer differ in size");
        tmp.resize(n_);
        const float* src[1] = { reinterpret_cast<const float*>(ptr_) };
        float* dst[1] = { reinterpret_cast<float*>(tmp.ptr_) };
        ctx_->check(reyn_gather_f32(ctx_->get(), sorted.ptr(), static_cast<uint32_t>(n_), src, dst, 1), "reorder");
        std::swap(ptr_, tmp.ptr_);
    }

    // reductions (buffer.cuh:213-258), formed in FP64 on the device
    double min() const requires std::same_as<T, float> { return reduce(1); }
    double max() const requires std::same_as<T, float> { return reduce(2); }
    double sum() const requires std::same_as<T, float> { return reduce(0); }
    float avg() const requires std::same_as<T, float> { return static_cast<float>(reduce(3)); }

    // swaps storage with another buffer of the same context
    void swap(DeviceBuffer& o) noexcept
    {
        std::swap(ptr_, o.ptr_);
        std::swap(n_, o.n_);
    }

private:
    double reduce(int op) const
    {
        double out = 0.0;
        ctx_->check(reyn_reduce_f32(ctx_->get(), ptr_, static_cast<uint32_t>(n_), op, &out), "reduce");
        return out;
    }
    void alloc_raw(size_t n)
    {
        void* p = nullptr;
        ctx_->check(reyn_malloc(ctx_->get(), &p, std::max<size_t>(n, 1) * sizeof(T)), "malloc");
        ptr_ = static_cast<T*>(p);
    }
    void alloc(size_t n)
    {
        alloc_raw(n);
        n_ = n;
    }
    void release()
    {
        if (ptr_)
            reyn_free(ctx_->get(), ptr_);
        ptr_ = nullptr;
        n_ = 0;
    }

    Context* ctx_;
    T* ptr_ = nullptr;
    size_t n_ = 0;
};

// ---- Particles (src/particles.cuh:25-75) -------------------------------------------

class Particles {
public:
    DeviceBuffer<float> xx, xy, xz, vx, vy, vz, m;
    float rho_0;

    Particles(size_t N, float rho_0_, Context& ctx = Context::current())
        : xx(N, 0.f, ctx)
        , xy(N, 0.f, ctx)
        , xz(N, 0.f, ctx)
        , vx(N, 0.f, ctx)
        , vy(N, 0.f, ctx)
        , vz(N, 0.f, ctx)
        , m(N, 0.f, ctx)
        , rho_0(rho_0_)
    {
    }
    Particles(const Particles&) = delete;
    Particles& operator=(const Particles&) = delete;

    size_t size() const { return xx.size(); }
    // every field to n entries, contents kept up to min(old, new)
    void resize(size_t n)
    {
        for (DeviceBuffer<float>* b : { &xx, &xy, &xz, &vx, &vy, &vz, &m })
            b->resize(n);
    }
    reyn_particles view() const
    {
        return reyn_particles { xx.ptr(), xy.ptr(), xz.ptr(), vx.ptr(), vy.ptr(), vz.ptr(), m.ptr(),
            static_cast<uint32_t>(xx.size()) };
    }
    // reorder of the seven fields by `sorted` (particles.cu:93-113)
    void reorder(const DeviceBuffer<uint32_t>& sorted, DeviceBuffer<float>& tmp)
    {
        for (DeviceBuffer<float>* b : { &vx, &vy, &vz, &m, &xx, &xy, &xz })
            b->reorder(sorted, tmp);
    }
};

// ---- smoothing kernels (src/kernels.cuh:19-174) -------------------------------------

template <typename Derived> class Kernel {
public:
    float h_bar() const { return h_bar_; }
    static constexpr int32_t kind() { return Derived::kKind; }

protected:
    explicit Kernel(float h_bar)
        : h_bar_(h_bar)
    {
    }

private:
    float h_bar_;
};

template <typename T>
concept IsKernel = std::is_base_of_v<Kernel<T>, T>;

class C2 : public Kernel<C2> {
public:
    static constexpr int32_t kKind = REYN_KERNEL_C2;
    explicit C2(float h_bar)
        : Kernel(h_bar)
    {
    }
};
class B3 : public Kernel<B3> {
public:
    static constexpr int32_t kKind = REYN_KERNEL_B3;
    explicit B3(float h_bar)
        : Kernel(h_bar)
    {
    }
};
class W6 : public Kernel<W6> {
public:
    static constexpr int32_t kKind = REYN_KERNEL_W6;
    explicit W6(float h_bar)
        : Kernel(h_bar)
    {
    }
};
class COS : public Kernel<COS> {
public:
    static constexpr int32_t kKind = REYN_KERNEL_COS;
    explicit COS(float h_bar)
        : Kernel(h_bar)
    {
    }
};

#define FOREACH_KERNEL(X)                                                                                           \
    X(C2)                                                                                                           \
    X(B3)                                                                                                           \
    X(W6)                                                                                                           \
    X(COS)

// ---- uniform grid (src/datastructure/uniformgrid.cuh) ------------------------------------

enum class Resort : bool { no = false, yes = true };

// Query-side POD; `grid.sorted` is NULL for Resort::yes.
template <Resort R> struct UniformGrid {
    reyn_grid grid;
};

template <typename T>
concept IsFltDevBufPtr = std::is_same_v<T, DeviceBuffer<float>*>;

class UniformGridBuilder {
public:
    // cell_ratio = cell size / search radius; the reference fixes 0.5. `order`
    // is REYN_CELL_ORDER_XYZ (the reference's keys) or REYN_CELL_ORDER_BRICK.
    UniformGridBuilder(float3 bound_min, float3 bound_max, float search_radius, float cell_ratio = 0.5f,
        Context& ctx = Context::current(), int32_t order = REYN_CELL_ORDER_XYZ)
        : ctx_(&ctx)
    {
        ctx_->check(reyn_grid_builder_create_ordered(ctx.get(), bound_min, bound_max, search_radius, cell_ratio,
                        order, &b_),
            "UniformGridBuilder");
    }
    ~UniformGridBuilder() { reyn_grid_builder_destroy(b_); }
    UniformGridBuilder(const UniformGridBuilder&) = delete;
    UniformGridBuilder& operator=(const UniformGridBuilder&) = delete;

    UniformGrid<Resort::no> construct(float search_radius, const DeviceBuffer<float>& xx,
        const DeviceBuffer<float>& xy, const DeviceBuffer<float>& xz)
    {
        if (deferred_)
            check();
        UniformGrid<Resort::no> g {};
        ctx_->check(reyn_grid_construct(ctx_->get(), b_, search_radius, xx.ptr(), xy.ptr(), xz.ptr(),
                        static_cast<uint32_t>(xx.size()), &g.grid),
            "construct");
        return g;
    }

    // reorders x, y, z and any number of further buffers (uniformgrid.cuh:408-417)
    template <IsFltDevBufPtr... MoreBufs>
    UniformGrid<Resort::yes> construct_and_reorder(float search_radius, DeviceBuffer<float>& tmp,
        DeviceBuffer<float>& xx, DeviceBuffer<float>& xy, DeviceBuffer<float>& xz, MoreBufs... reorder)
    {
        std::vector<DeviceBuffer<float>*> bufs { &xx, &xy, &xz, reorder... };
        return reorder_all(search_radius, tmp, bufs);
    }

    // reorders the particle state plus further buffers (uniformgrid.cuh:475-483)
    template <IsFltDevBufPtr... MoreBufs>
    UniformGrid<Resort::yes> construct_and_reorder(float search_radius, DeviceBuffer<float>& tmp,
        Particles& state, MoreBufs... reorder)
    {
        std::vector<DeviceBuffer<float>*> bufs { &state.xx, &state.xy, &state.xz, &state.vx, &state.vy,
            &state.vz, &state.m, reorder... };
        return reorder_all(search_radius, tmp, bufs);
    }

    reyn_grid geometry() const
    {
        reyn_grid g {};
        reyn_grid_builder_geometry(b_, &g);
        return g;
    }

    // Deferred build checks (no reference counterpart): construct calls return
    // before the build's status is read, and check() throws what they would
    // have thrown; a host loop that waits for the device anyway (a time-step
    // read-back) checks after that wait.
    void defer_check(bool on = true)
    {
        ctx_->check(reyn_grid_builder_defer_check(b_, on ? 1 : 0), "defer_check");
        deferred_ = on;
    }
    void check() { ctx_->check(reyn_grid_builder_check(ctx_->get(), b_), "construct (deferred check)"); }

private:
    UniformGrid<Resort::yes> reorder_all(float r, DeviceBuffer<float>& tmp, std::vector<DeviceBuffer<float>*>& bufs)
    {
        if (deferred_)
            check(); // an earlier build's error, before any buffer is swapped
        const size_t n = bufs[0]->size();
        for (auto* b : bufs)
            if (b->size() != n)
                throw std::runtime_error("Attempted to resort an array of grid construction the dimensions of "
                                         "which don't match the number of particles in the sorting.");
        // one scratch buffer per field so the whole reorder is a single gather;
        // `tmp` keeps the reference's signature and serves as the first one
        while (scratch_.size() + 1 < bufs.size())
            scratch_.emplace_back(n, *ctx_);
        std::vector<DeviceBuffer<float>*> tmps { &tmp };
        for (size_t k = 0; k + 1 < bufs.size(); ++k)
            tmps.push_back(&scratch_[k]);
        std::vector<float*> f(bufs.size()), t(bufs.size());
        for (size_t k = 0; k < bufs.size(); ++k) {
            tmps[k]->resize(n);
            f[k] = bufs[k]->ptr();
            t[k] = tmps[k]->ptr();
        }
        UniformGrid<Resort::yes> g {};
        const int rc = reyn_grid_construct_and_reorder(ctx_->get(), b_, r, f.data(), t.data(),
            static_cast<uint32_t>(bufs.size()), static_cast<uint32_t>(n), &g.grid);
        if (rc == REYN_OK || rc == REYN_ERR_OUT_OF_BOUNDS || rc == REYN_ERR_CAPACITY)
            for (size_t k = 0; k < bufs.size(); ++k)
                bufs[k]->swap(*tmps[k]);
        ctx_->check(rc, "construct_and_reorder");
        return g;
    }

    Context* ctx_;
    reyn_grid_builder* b_ = nullptr;
    bool deferred_ = false;
    std::vector<DeviceBuffer<float>> scratch_;
};

// ---- boundary (src/scene/scene.cuh:14-46; masses sample_boundary.cu:251-297) -------------

struct Boundary {
    reyn_boundary b;
};

class BoundarySamples {
public:
    DeviceBuffer<float> xs, ys, zs, m;
    float3 bound_min, bound_max;

    // samples in any order; they are stored in cell order with Akinci masses
    BoundarySamples(const std::vector<float3>& pts, float h, float rho_0, unsigned mass_iterations = 5,
        Context& ctx = Context::current())
        : xs(pts.size(), ctx)
        , ys(pts.size(), ctx)
        , zs(pts.size(), ctx)
        , m(pts.size(), 1.f, ctx)
    {
        if (pts.empty())
            throw std::runtime_error("reyn: BoundarySamples needs at least one sample");
        std::vector<float> x(pts.size()), y(pts.size()), z(pts.size());
        bound_min = bound_max = pts[0];
        for (size_t i = 0; i < pts.size(); ++i) {
            x[i] = pts[i].x;
            y[i] = pts[i].y;
            z[i] = pts[i].z;
            bound_min = v3(std::min(bound_min.x, x[i]), std::min(bound_min.y, y[i]), std::min(bound_min.z, z[i]));
            bound_max = v3(std::max(bound_max.x, x[i]), std::max(bound_max.y, y[i]), std::max(bound_max.z, z[i]));
        }
        xs.upload(x);
        ys.upload(y);
        zs.upload(z);
        builder_ = std::make_unique<UniformGridBuilder>(bound_min, bound_max, 2.f * h, 0.5f, ctx);
        DeviceBuffer<float> tmp(pts.size(), ctx);
        grid_ = builder_->construct_and_reorder(2.f * h, tmp, xs, ys, zs);
        const reyn_grid& g = grid_.grid;
        near_ = std::make_unique<DeviceBuffer<uint8_t>>(size_t(g.nx) * size_t(g.ny) * size_t(g.nz), ctx);
        ctx.check(reyn_grid_occupancy(ctx.get(), &g, near_->ptr()), "boundary occupancy");
        reyn_boundary v = get().b;
        ctx.check(reyn_boundary_masses(ctx.get(), &v, m.ptr(), h, rho_0, mass_iterations), "boundary masses");
    }

    Boundary get() const
    {
        return Boundary { reyn_boundary { xs.ptr(), ys.ptr(), zs.ptr(), m.ptr(), static_cast<uint32_t>(xs.size()),
            grid_.grid, near_ ? near_->ptr() : nullptr } };
    }

private:
    std::unique_ptr<UniformGridBuilder> builder_;
    UniformGrid<Resort::yes> grid_ {};
    std::unique_ptr<DeviceBuffer<uint8_t>> near_; // cells with samples within reach
};

// Scene::hard_enforce_bounds (src/scene/scene.cu:466-491)
inline void hard_enforce_bounds(float3 bound_min, float3 bound_max, Particles& state)
{
    Context& c = state.xx.context();
    const reyn_particles p = state.view();
    c.check(reyn_hard_enforce_bounds(c.get(), bound_min, bound_max, &p), "hard_enforce_bounds");
}

// ---- time step (src/timestep/cfl.cu) --------------------------------------------------------

inline float cfl_time_step(float lambda, float h, const Particles& state, float3 g)
{
    Context& c = state.xx.context();
    const reyn_particles p = state.view();
    float dt = 0.f;
    c.check(reyn_cfl_time_step(c.get(), lambda, h, &p, g, &dt, nullptr), "cfl_time_step");
    return dt;
}

inline float simple_dt_controller(unsigned iter_last, float dt_last, float lambda, float h, const Particles& state,
    float3 g, unsigned iter_target = 5, float change_factor = 0.01f)
{
    Context& c = state.xx.context();
    const reyn_particles p = state.view();
    float dt = 0.f;
    c.check(reyn_simple_dt_controller(c.get(), iter_last, dt_last, lambda, h, &p, g, iter_target, change_factor,
                &dt),
        "simple_dt_controller");
    return dt;
}

// ---- solvers (src/solvers) --------------------------------------------------------------------

namespace detail {
    template <IsKernel K>
    reyn_sph_params sph(const K& W, float nu, float h, float rho_0, float3 g)
    {
        return reyn_sph_params { K::kKind, W.h_bar(), h, nu, rho_0, g };
    }
    template <Resort R> const reyn_grid* grid_of(const UniformGrid<R>& g) { return &g.grid; }
} // namespace detail

template <IsKernel K, Resort R> class SESPH {
public:
    SESPH(K W, uint32_t N, float nu, float h, float rho_0, DeviceBuffer<float>& rho, DeviceBuffer<float>& ax,
        DeviceBuffer<float>& ay, DeviceBuffer<float>& az, float3 g = v3(0.f, -9.81f, 0.f), float k = 500.f)
        : prm_ { detail::sph(W, nu, h, rho_0, g), k }
        , rho_(rho)
        , ax_(ax)
        , ay_(ay)
        , az_(az)
    {
        for (auto* b : { &rho_, &ax_, &ay_, &az_ })
            b->resize(N);
    }
    void step(Particles& state, const UniformGrid<R> grid, const BoundarySamples& bdy, float dt)
    {
        const reyn_particles p = state.view();
        const Boundary b = bdy.get();
        const reyn_sesph_work w { rho_.ptr(), ax_.ptr(), ay_.ptr(), az_.ptr() };
        Context& c = state.xx.context();
        c.check(reyn_sesph_step(c.get(), &prm_, &p, detail::grid_of(grid), &b.b, &w, dt), "SESPH::step");
    }
    SESPH(const SESPH&) = delete;
    SESPH& operator=(const SESPH&) = delete;

private:
    reyn_sesph_params prm_;
    DeviceBuffer<float>&rho_, &ax_, &ay_, &az_;
};

template <IsKernel K, Resort R> class PCISPH {
public:
    PCISPH(K W, uint32_t N, float nu, float h, float rho_0, DeviceBuffer<float>& rho, DeviceBuffer<float>& ax,
        DeviceBuffer<float>& ay, DeviceBuffer<float>& az, float3 g = v3(0.f, -9.81f, 0.f))
        : rho_(rho)
        , ax_(ax)
        , ay_(ay)
        , az_(az)
    {
        prm_.sph = detail::sph(W, nu, h, rho_0, g);
        if (reyn_pcisph_delta(K::kKind, W.h_bar(), h, &prm_.delta) != REYN_OK)
            throw std::runtime_error("reyn: invalid PCISPH parameters");
        prm_.eta_rho_max = 0.001f;
        prm_.min_iter = 3;
        prm_.max_iter = 0;
        prm_.check_every = 4;
        for (auto* b : { &rho_, &ax_, &ay_, &az_ })
            b->resize(N);
    }
    uint32_t step(Particles& state, const UniformGrid<R> grid, const BoundarySamples& bdy, float dt)
    {
        const reyn_particles p = state.view();
        const Boundary b = bdy.get();
        const reyn_pcisph_work w { rho_.ptr(), ax_.ptr(), ay_.ptr(), az_.ptr() };
        Context& c = state.xx.context();
        c.check(reyn_pcisph_step(c.get(), &prm_, &p, detail::grid_of(grid), &b.b, &w, dt, &info_), "PCISPH::step");
        return info_.iterations;
    }
    reyn_pcisph_params& params() { return prm_; }
    const reyn_solve_info& info() const { return info_; }
    PCISPH(const PCISPH&) = delete;
    PCISPH& operator=(const PCISPH&) = delete;

private:
    reyn_pcisph_params prm_ {};
    reyn_solve_info info_ {};
    DeviceBuffer<float>&rho_, &ax_, &ay_, &az_;
};

template <IsKernel K, Resort R> class IISPH {
public:
    // argument order of the reference's constructor (IISPH.cuh:51-55); the
    // bit-flag buffer is replaced by an on-device counter and is only resized
    IISPH(K W, uint32_t N, float nu, float h, float rho_0, DeviceBuffer<float>& rho, DeviceBuffer<float>& a_ii,
        DeviceBuffer<float>& s_i, DeviceBuffer<float>& ax, DeviceBuffer<float>& ay, DeviceBuffer<float>& az,
        DeviceBuffer<float>& p, DeviceBuffer<uint32_t>& rho_err_threshold, float3 g = v3(0.f, -9.81f, 0.f))
        : rho_(rho)
        , aii_(a_ii)
        , si_(s_i)
        , ax_(ax)
        , ay_(ay)
        , az_(az)
        , p_(p)
    {
        prm_.sph = detail::sph(W, nu, h, rho_0, g);
        prm_.omega = 0.5f;
        prm_.eta_rho_max = 0.005f;
        prm_.min_iter = 3;
        prm_.max_iter = 0;
        prm_.stop_rule = REYN_IISPH_STOP_ANY;
        prm_.warm_start = 1;
        prm_.check_every = 4;
        p_.resize(N, 0.f);
        for (auto* b : { &rho_, &aii_, &si_, &ax_, &ay_, &az_ })
            b->resize(N);
        rho_err_threshold.resize((N + 31) / 32, 0u);
    }
    uint32_t step(Particles& state, const UniformGrid<R> grid, const BoundarySamples& bdy, float dt)
    {
        const reyn_particles p = state.view();
        const Boundary b = bdy.get();
        const reyn_iisph_work w { rho_.ptr(), p_.ptr(), aii_.ptr(), si_.ptr(), ax_.ptr(), ay_.ptr(), az_.ptr() };
        Context& c = state.xx.context();
        c.check(reyn_iisph_step(c.get(), &prm_, &p, detail::grid_of(grid), &b.b, &w, dt, &info_), "IISPH::step");
        return info_.iterations;
    }
    reyn_iisph_params& params() { return prm_; }
    const reyn_solve_info& info() const { return info_; }
    IISPH(const IISPH&) = delete;
    IISPH& operator=(const IISPH&) = delete;

private:
    reyn_iisph_params prm_ {};
    reyn_solve_info info_ {};
    DeviceBuffer<float>&rho_, &aii_, &si_, &ax_, &ay_, &az_, &p_;
};

// ---- meshes: Mesh / load_mesh_from_obj (src/scene/loader.h), sample_mesh ----------
// (sample_boundary.cu:391-440, with the optional relaxation of :100-166)

class Mesh {
public:
    static Mesh load_obj(const std::filesystem::path& path, const std::vector<std::string>& ignore = {},
        const char* only = nullptr)
    {
        std::vector<const char*> ig;
        for (const auto& s : ignore)
            ig.push_back(s.c_str());
        char err[512] = {};
        reyn_mesh* m = nullptr;
        const int rc = reyn_mesh_load_obj(path.string().c_str(), ig.data(), uint32_t(ig.size()), only, &m, err,
            sizeof(err));
        if (rc != REYN_OK)
            throw std::runtime_error(std::string("reyn: load_obj: ") + err);
        return Mesh(m);
    }
    Mesh(Mesh&& o) noexcept : m_(std::exchange(o.m_, nullptr)) { }
    Mesh& operator=(Mesh&& o) noexcept
    {
        std::swap(m_, o.m_);
        return *this;
    }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    ~Mesh() { reyn_mesh_destroy(m_); }
    uint32_t face_count() const
    {
        uint32_t nf = 0;
        reyn_mesh_info(m_, nullptr, &nf, nullptr, nullptr);
        return nf;
    }
    double get_volume() const
    {
        double v = 0.0;
        reyn_mesh_info(m_, nullptr, nullptr, nullptr, &v);
        return v;
    }
    const reyn_mesh* get() const { return m_; }

private:
    explicit Mesh(reyn_mesh* m) : m_(m) { }
    reyn_mesh* m_ = nullptr;
};

// `LBVH` (src/datastructure/lbvh.cuh:614-758) of a mesh on the context's GPU:
// closest-point projection and the inside test of the fluid culling.
class LBVH {
public:
    explicit LBVH(const Mesh& mesh, Context& ctx = Context::current()) : ctx_(&ctx)
    {
        ctx.check(reyn_lbvh_create(ctx.get(), mesh.get(), &b_), "LBVH");
    }
    LBVH(LBVH&& o) noexcept : ctx_(o.ctx_), b_(std::exchange(o.b_, nullptr)) { }
    LBVH(const LBVH&) = delete;
    LBVH& operator=(const LBVH&) = delete;
    LBVH& operator=(LBVH&&) = delete;
    ~LBVH() { reyn_lbvh_destroy(b_); }
    uint32_t height() const
    {
        uint32_t h = 0;
        reyn_lbvh_info(b_, nullptr, &h);
        return h;
    }
    // every point to the closest point of the mesh, in place
    void project(DeviceBuffer<float>& xs, DeviceBuffer<float>& ys, DeviceBuffer<float>& zs) const
    {
        if (ys.size() != xs.size() || zs.size() != xs.size())
            throw std::runtime_error("LBVH::project called with coordinate buffers of differing sizes");
        ctx_->check(reyn_lbvh_project(ctx_->get(), b_, xs.ptr(), ys.ptr(), zs.ptr(), uint32_t(xs.size())),
            "LBVH::project");
    }
    // inside[k] = 1 for points inside the closed mesh
    void inside(const DeviceBuffer<float>& xs, const DeviceBuffer<float>& ys, const DeviceBuffer<float>& zs,
        DeviceBuffer<uint8_t>& out) const
    {
        if (ys.size() != xs.size() || zs.size() != xs.size() || out.size() != xs.size())
            throw std::runtime_error("LBVH::inside called with buffers of differing sizes");
        ctx_->check(
            reyn_lbvh_inside(ctx_->get(), b_, xs.ptr(), ys.ptr(), zs.ptr(), uint32_t(xs.size()), out.ptr()),
            "LBVH::inside");
    }
    const reyn_lbvh* get() const { return b_; }

private:
    Context* ctx_;
    reyn_lbvh* b_ = nullptr;
};

// boundary samples of the mesh surface with Akinci masses; `relaxation_iters`
// rounds of relaxation follow the sampling (the reference defaults to 50 with
// factor 1; here the default is none). `ratio`, when given, receives the
// max/min number density of the last relaxation round.
inline BoundarySamples sample_mesh(const Mesh& mesh, float h, float rho_0, float oversampling = 1.f,
    uint32_t seed = 1614201312u, Context& ctx = Context::current(), uint32_t relaxation_iters = 0,
    float relaxation_factor = 1.f, float* ratio = nullptr)
{
    uint32_t n = 0;
    ctx.check(reyn_mesh_sample_count(mesh.get(), h, oversampling, &n), "sample_mesh count");
    DeviceBuffer<float> xs(n, ctx), ys(n, ctx), zs(n, ctx);
    uint32_t got = 0;
    ctx.check(reyn_mesh_sample(ctx.get(), mesh.get(), h, oversampling, seed, xs.ptr(), ys.ptr(), zs.ptr(), n, &got),
        "sample_mesh");
    if (relaxation_iters)
        ctx.check(reyn_mesh_relax(ctx.get(), mesh.get(), h, oversampling, xs.ptr(), ys.ptr(), zs.ptr(), n,
                      relaxation_iters, relaxation_factor, ratio),
            "sample_mesh relaxation");
    const std::vector<float> x = xs.download(), y = ys.download(), z = zs.download();
    std::vector<float3> pts(n);
    for (uint32_t i = 0; i < n; ++i)
        pts[i] = v3(x[i], y[i], z[i]);
    return BoundarySamples(pts, h, rho_0, 5, ctx);
}

// ---- Scene (src/scene/scene.cuh:48-156, scene.cu:239-491) --------------------------
//
// The reference's two scene kinds: a box of fluid in a boundary sampled from
// an OBJ file (every shape but "fluid"), and the volume of an OBJ shape named
// "fluid" filled with fluid inside the other shapes. Fluid points start on a
// lattice of spacing h (chosen for about N_desired particles), points within
// cull_bdy_radius h of a boundary sample are removed, and the rest are jittered
// by a normal deviate of standard deviation jitter_stddev h. The boundary is
// sampled with relaxation_iters rounds of relaxation (the reference's default
// 50). The random numbers are a counter hash of `seed` in place of curand and
// thrust's engine.
class Scene {
public:
    float ρ₀;
    float h;
    uint32_t N = 0;
    BoundarySamples bdy;
    float3 bound_min, bound_max;

    Scene(const std::filesystem::path& path, uint32_t N_desired, float3 min, float3 max, float rho_0,
        Particles& state, DeviceBuffer<float>& tmp, float bdy_oversampling_factor = 2.f, float cull_bdy_radius = 1.f,
        float jitter_stddev = 0.01f, uint32_t relaxation_iters = 50, uint32_t seed = 1614201312u,
        Context& ctx = Context::current())
        : ρ₀(rho_0)
        , h(std::cbrt(((max.x - min.x) * (max.y - min.y)) * (max.z - min.z) / static_cast<float>(N_desired)))
        , bdy(sample_mesh(Mesh::load_obj(path, { "fluid" }), h, rho_0, bdy_oversampling_factor, seed, ctx,
              relaxation_iters))
        , bound_min(bdy.bound_min)
        , bound_max(bdy.bound_max)
        , grid_builder_(bound_min, bound_max, 2.f * h, 0.5f, ctx)
    {
        const long nx = std::lround(std::floor((max.x - min.x) / h)), ny = std::lround(std::floor((max.y - min.y) / h)),
                   nz = std::lround(std::floor((max.z - min.z) / h));
        const size_t n = size_t(std::labs(nx)) * size_t(std::labs(ny)) * size_t(std::labs(nz));
        if (n == 0)
            throw std::domain_error("Initialization of box failed, zero particles placed");
        state.resize(n);
        ctx.check(reyn_lattice_box(ctx.get(), min, uint32_t(std::labs(nx)), uint32_t(std::labs(ny)),
                      uint32_t(std::labs(nz)), h, state.xx.ptr(), state.xy.ptr(), state.xz.ptr()),
            "Scene lattice");
        finish(state, tmp, n, cull_bdy_radius, jitter_stddev, seed, ctx);
    }

    static Scene from_obj(const std::filesystem::path& path, uint32_t N_desired, float rho_0, Particles& state,
        DeviceBuffer<float>& tmp, float bdy_oversampling_factor = 2.f, float cull_bdy_radius = 1.f,
        float jitter_stddev = 0.01f, uint32_t relaxation_iters = 50, uint32_t seed = 1614201312u,
        Context& ctx = Context::current())
    {
        const Mesh fluid = Mesh::load_obj(path, {}, "fluid");
        const float h = static_cast<float>(std::cbrt(fluid.get_volume() / static_cast<double>(N_desired)));
        return Scene(path, fluid, h, rho_0, state, tmp, bdy_oversampling_factor, cull_bdy_radius, jitter_stddev,
            relaxation_iters, seed, ctx);
    }

    // Scene::hard_enforce_bounds (scene.cu:466-491)
    void hard_enforce_bounds(Particles& state) const { reyn::hard_enforce_bounds(bound_min, bound_max, state); }

    // Scene::get_grid (scene.cuh:136-156): the grid of radius search_radius h,
    // the state and the further buffers reordered into cell order
    template <IsFltDevBufPtr... MoreBufs>
    UniformGrid<Resort::yes> get_grid(Particles& state, DeviceBuffer<float>& tmp, float search_radius = 2.f,
        MoreBufs... reorder)
    {
        return grid_builder_.construct_and_reorder(search_radius * h, tmp, state, reorder...);
    }

private:
    UniformGridBuilder grid_builder_;

    // the fluid volume of `fluid`, the boundary from the other shapes of path
    Scene(const std::filesystem::path& path, const Mesh& fluid, float h_, float rho_0, Particles& state,
        DeviceBuffer<float>& tmp, float bdy_oversampling_factor, float cull_bdy_radius, float jitter_stddev,
        uint32_t relaxation_iters, uint32_t seed, Context& ctx)
        : ρ₀(rho_0)
        , h(h_)
        , bdy(sample_mesh(Mesh::load_obj(path, { "fluid" }), h_, rho_0, bdy_oversampling_factor, seed, ctx,
              relaxation_iters))
        , bound_min(bdy.bound_min)
        , bound_max(bdy.bound_max)
        , grid_builder_(grow_min(bdy.bound_min, fluid, h_), grow_max(bdy.bound_max, fluid, h_), 2.f * h_, 0.5f, ctx)
    {
        // the scene bounds also hold the fluid's box, with half a spacing to spare
        bound_min = grow_min(bound_min, fluid, h);
        bound_max = grow_max(bound_max, fluid, h);
        const auto [lo, hi] = mesh_box(fluid);
        const long nx = std::lround(std::floor((hi.x - lo.x) / h)), ny = std::lround(std::floor((hi.y - lo.y) / h)),
                   nz = std::lround(std::floor((hi.z - lo.z) / h));
        const size_t n = size_t(std::labs(nx)) * size_t(std::labs(ny)) * size_t(std::labs(nz));
        if (n == 0)
            throw std::domain_error("Initialization of box failed, zero particles placed");
        // lattice candidates in the fluid's box, kept where inside the fluid
        state.resize(n);
        ctx.check(reyn_lattice_box(ctx.get(), lo, uint32_t(std::labs(nx)), uint32_t(std::labs(ny)),
                      uint32_t(std::labs(nz)), h, state.xx.ptr(), state.xy.ptr(), state.xz.ptr()),
            "Scene lattice");
        const LBVH bvh(fluid, ctx);
        DeviceBuffer<uint8_t> keep(n, ctx);
        bvh.inside(state.xx, state.xy, state.xz, keep);
        uint32_t kept = 0;
        float* xyz[3] = { state.xx.ptr(), state.xy.ptr(), state.xz.ptr() };
        ctx.check(reyn_compact_f32(ctx.get(), keep.ptr(), uint32_t(n), xyz, 3, &kept), "Scene fluid volume");
        state.resize(kept);
        finish(state, tmp, kept, cull_bdy_radius, jitter_stddev, seed, ctx);
    }

    static std::pair<float3, float3> mesh_box(const Mesh& m)
    {
        uint32_t nv = 0;
        reyn_mesh_info(m.get(), &nv, nullptr, nullptr, nullptr);
        std::vector<double> x(nv), y(nv), z(nv);
        std::vector<uint32_t> f(3 * size_t(m.face_count()));
        reyn_mesh_data(m.get(), x.data(), y.data(), z.data(), f.data());
        double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
        for (uint32_t v = 0; v < nv; ++v) {
            const double p[3] = { x[v], y[v], z[v] };
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], p[a]);
                hi[a] = std::max(hi[a], p[a]);
            }
        }
        return { v3(float(lo[0]), float(lo[1]), float(lo[2])), v3(float(hi[0]), float(hi[1]), float(hi[2])) };
    }
    static float3 grow_min(float3 b, const Mesh& fluid, float h)
    {
        const float3 lo = mesh_box(fluid).first;
        return v3(std::min(b.x, lo.x - 0.5f * h), std::min(b.y, lo.y - 0.5f * h), std::min(b.z, lo.z - 0.5f * h));
    }
    static float3 grow_max(float3 b, const Mesh& fluid, float h)
    {
        const float3 hi = mesh_box(fluid).second;
        return v3(std::max(b.x, hi.x + 0.5f * h), std::max(b.y, hi.y + 0.5f * h), std::max(b.z, hi.z + 0.5f * h));
    }

    // rest mass and zero velocity, the cull near the boundary, the jitter
    void finish(Particles& state, DeviceBuffer<float>& tmp, size_t n, float cull_bdy_radius, float jitter_stddev,
        uint32_t seed, Context& ctx)
    {
        state.m.fill(ρ₀ * h * h * h);
        for (DeviceBuffer<float>* b : { &state.vx, &state.vy, &state.vz })
            b->fill(0.f);
        DeviceBuffer<uint8_t> keep(n, ctx);
        const reyn_boundary b = bdy.get().b;
        ctx.check(reyn_boundary_cull(ctx.get(), &b, cull_bdy_radius * h, state.xx.ptr(), state.xy.ptr(),
                      state.xz.ptr(), uint32_t(n), keep.ptr()),
            "Scene cull");
        float* xyz[3] = { state.xx.ptr(), state.xy.ptr(), state.xz.ptr() };
        uint32_t kept = 0;
        ctx.check(reyn_compact_f32(ctx.get(), keep.ptr(), uint32_t(n), xyz, 3, &kept), "Scene cull");
        state.resize(kept);
        N = kept;
        ctx.check(reyn_jitter_normal(ctx.get(), state.xx.ptr(), state.xy.ptr(), state.xz.ptr(), kept, jitter_stddev * h,
                      seed ^ 0x5bd1e995u),
            "Scene jitter");
        tmp.resize(kept);
        ctx.synchronize();
    }
};

// ---- save_to_ply (src/scene/ply_io.cu:7-56) ----------------------------------------
//
// Binary PLY of the positions in the host's byte order: one vertex element of
// float32 x, y, z, written as x_i, y_i, z_i per particle.
inline void save_to_ply(const std::filesystem::path& path, const DeviceBuffer<float>& xs,
    const DeviceBuffer<float>& ys, const DeviceBuffer<float>& zs)
{
    const size_t n = xs.size();
    if (ys.size() != n || zs.size() != n)
        throw std::runtime_error("save_to_ply called on buffers with differing sizes");
    static_assert(std::endian::native == std::endian::little || std::endian::native == std::endian::big,
        "save_to_ply needs a little- or big-endian host");
    const char* order = std::endian::native == std::endian::little ? "little" : "big";
    const std::vector<float> x = xs.download(), y = ys.download(), z = zs.download();
    std::vector<float> xyz(3 * n);
    for (size_t i = 0; i < n; ++i) {
        xyz[3 * i] = x[i];
        xyz[3 * i + 1] = y[i];
        xyz[3 * i + 2] = z[i];
    }
    std::ofstream out(path, std::ios::binary);
    if (!out)
        throw std::runtime_error("save_to_ply: cannot open " + path.string());
    out << "ply\nformat binary_" << order << "_endian 1.0\n"
        << "comment particle positions written by reyn_amd\n"
        << "element vertex " << n << "\n"
        << "property float32 x\nproperty float32 y\nproperty float32 z\nend_header\n";
    out.write(reinterpret_cast<const char*>(xyz.data()), std::streamsize(xyz.size() * sizeof(float)));
    if (!out)
        throw std::runtime_error("save_to_ply: write failed for " + path.string());
}

// ---- step telemetry (the GUI's dt and iteration plots, src/gui.cu:956-957) ---------

struct StepTrace {
    std::vector<float> dt_ms; // time step of each step in milliseconds, as plotted
    std::vector<unsigned> iters; // pressure iterations of each step
    void record(float dt, unsigned iterations)
    {
        dt_ms.push_back(dt * 1000.f);
        iters.push_back(iterations);
    }
    void write_csv(const std::filesystem::path& path) const
    {
        std::ofstream out(path);
        if (!out)
            throw std::runtime_error("StepTrace: cannot open " + path.string());
        out << "step,dt_ms,iterations\n";
        out.precision(9); // float values round-trip
        for (size_t k = 0; k < dt_ms.size(); ++k)
            out << k + 1 << ',' << dt_ms[k] << ',' << iters[k] << '\n';
    }
};

// ---- scene batches (no reference counterpart) -------------------------------------
//
// Many scenes advanced together, each step being the reference's main-loop
// step (main.cu:59-78) per scene; scenes join and leave between steps and are
// addressed by id. A scene's result is the same alone and in any batch.

struct BatchSceneState {
    std::vector<float> x, y, z, vx, vy, vz, rho, p;
};

class SceneBatch {
public:
    SceneBatch(const reyn_batch_params& prm, const std::vector<reyn_batch_scene>& scenes = {},
        Context& ctx = Context::current())
        : ctx_(&ctx)
    {
        ctx.check(reyn_batch_create(ctx.get(), &prm, scenes.data(), uint32_t(scenes.size()), &b_),
            "SceneBatch create");
        for (size_t k = 0; k < scenes.size(); ++k)
            n_.push_back({ uint32_t(k), scenes[k].n });
    }
    ~SceneBatch() { reyn_batch_destroy(b_); }
    SceneBatch(const SceneBatch&) = delete;
    SceneBatch& operator=(const SceneBatch&) = delete;

    std::vector<uint32_t> add(const std::vector<reyn_batch_scene>& scenes)
    {
        std::vector<uint32_t> ids(scenes.size());
        ctx_->check(reyn_batch_add(b_, scenes.data(), uint32_t(scenes.size()), ids.data()), "SceneBatch add");
        for (size_t k = 0; k < scenes.size(); ++k)
            n_.push_back({ ids[k], scenes[k].n });
        return ids;
    }
    // scenes generated on the device from their descriptions
    std::vector<uint32_t> add_dam_breaks(const std::vector<reyn_dam_break>& scenes)
    {
        std::vector<uint32_t> ids(scenes.size());
        ctx_->check(reyn_batch_add_dam_breaks(b_, scenes.data(), uint32_t(scenes.size()), ids.data()),
            "SceneBatch add_dam_breaks");
        for (size_t k = 0; k < scenes.size(); ++k) {
            uint32_t n = 0;
            ctx_->check(reyn_batch_scene_size(b_, ids[k], &n), "SceneBatch scene size");
            n_.push_back({ ids[k], n });
        }
        return ids;
    }
    void remove(const std::vector<uint32_t>& ids)
    {
        ctx_->check(reyn_batch_remove(b_, ids.data(), uint32_t(ids.size())), "SceneBatch remove");
        std::erase_if(n_, [&](const auto& e) { return std::find(ids.begin(), ids.end(), e.first) != ids.end(); });
    }
    void step(uint32_t steps = 1) { ctx_->check(reyn_batch_step(b_, steps), "SceneBatch step"); }
    reyn_batch_scene_info info(uint32_t id) const
    {
        reyn_batch_scene_info out {};
        ctx_->check(reyn_batch_info(b_, id, &out), "SceneBatch info");
        return out;
    }
    BatchSceneState download(uint32_t id) const
    {
        uint32_t n = 0;
        for (const auto& e : n_)
            if (e.first == id)
                n = e.second;
        BatchSceneState s;
        for (auto* v : { &s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.rho, &s.p })
            v->resize(n);
        ctx_->check(reyn_batch_download(b_, id, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(),
                        s.vz.data(), s.rho.data(), s.p.data()),
            "SceneBatch download");
        return s;
    }
    reyn_batch* get() const { return b_; }

private:
    Context* ctx_;
    reyn_batch* b_ = nullptr;
    std::vector<std::pair<uint32_t, uint32_t>> n_; // id, particles
};

} // namespace reyn
