// reyn oracle — TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// A CPU restatement of reyn's per-timestep SPH path (JulianKarrer/reyn),
// written from the reference's source as read, templated on the arithmetic
// type: `double` is the FP64 oracle the GPU results are checked against,
// `float` is the OpenMP "port" timed as the CPU baseline. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library, and only as the checker or as the baseline; the product path never
// calls it.
//
// Scheme choices shared with the GPU library so results are comparable:
//   * the neighbour test is evaluated in FP32 exactly as the reference does
//     (x_ij = x_i - x_j, d2 = fmaf(x,x, fmaf(y,y, z*z)), d2 <= r_c_2), so the
//     neighbour sets of oracle and GPU are identical; the map bodies then run
//     in T;
//   * cell coordinates are clamped to the grid, and the in-cell order of the
//     counting sort is ascending point index (the reference's is racy).
//
// Every function names the reference lines it follows. What pins each one
// (closed forms, brute force, invariants) is listed in tests/test_oracle.py;
// no function here is left without such a check.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include <thrust/execution_policy.h>
#include <thrust/functional.h>
#include <thrust/gather.h>
#include <thrust/iterator/counting_iterator.h>
#include <thrust/scan.h>
#include <thrust/system/omp/execution_policy.h>
#include <thrust/transform_reduce.h>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

// ---- grid geometry and keys (src/datastructure/uniformgrid.cu:65-96,
//      uniformgrid.cuh:17-64, src/utils/vector.cuh:137-149) -----------------

struct OGrid {
    float bmin[3];
    float cell;
    float r2;
    int n[3];
    int range;
    const uint32_t* prefix;
    const uint32_t* sorted; // nullptr: arrays are in cell order
    int order; // 0: the reference's XYZ keys; 1: 2 x 2 x 2 bricks
};

// The key of cell (x, y, z): z nx ny + y nx + x (uniformgrid.cuh:33-40), or in
// brick order 8 brick + 4 (z & 1) + 2 (y & 1) + (x & 1) with the bricks in
// XYZ order over (nx/2, ny/2, nz/2).
inline uint32_t okey(int order, const int n[3], int x, int y, int z)
{
    if (order == 1) {
        const uint32_t nbx = uint32_t(n[0] / 2), nby = uint32_t(n[1] / 2);
        const uint32_t brick = (uint32_t(z >> 1) * nby + uint32_t(y >> 1)) * nbx + uint32_t(x >> 1);
        return brick * 8u + uint32_t(((z & 1) << 2) | ((y & 1) << 1) | (x & 1));
    }
    return uint32_t(z) * uint32_t(n[0]) * uint32_t(n[1]) + uint32_t(y) * uint32_t(n[0]) + uint32_t(x);
}

inline int cell_axis(float x, float lo, float cell, int n)
{
    float f = std::floor((x - lo) / cell);
    f = std::fmax(f, 0.f);
    f = std::fmin(f, static_cast<float>(n - 1));
    return static_cast<int>(f);
}

// The GPU library's walk budget and the most points per cell a grid build
// accepts, (2^15 candidates) / (2 range + 1)^3: a build over the limit fails
// there with REYN_ERR_CAPACITY, and here with a negative return, so no walk
// over a built grid has more candidates than the budget. The walk below is
// never clipped.
constexpr uint32_t kWalkBudget = 1u << 15;
inline uint32_t walk_cell_cap(int range)
{
    return kWalkBudget / uint32_t((2 * range + 1) * (2 * range + 1) * (2 * range + 1));
}

// Strip walk of `ff_nbrs` (uniformgrid.cuh:113-178), clipped to the grid.
// f(j) is called for every neighbour j, in walk order.
template <class T, class F>
inline void walk(const OGrid& g, const T* xs, const T* ys, const T* zs, T xi, T yi, T zi, F&& f)
{
    const float fx = static_cast<float>(xi), fy = static_cast<float>(yi), fz = static_cast<float>(zi);
    const int cx = cell_axis(fx, g.bmin[0], g.cell, g.n[0]);
    const int cy = cell_axis(fy, g.bmin[1], g.cell, g.n[1]);
    const int cz = cell_axis(fz, g.bmin[2], g.cell, g.n[2]);
    const int R = g.range;
    const int x0 = std::max(cx - R, 0), x1 = std::min(cx + R, g.n[0] - 1);
    const int y0 = std::max(cy - R, 0), y1 = std::min(cy + R, g.n[1] - 1);
    const int z0 = std::max(cz - R, 0), z1 = std::min(cz + R, g.n[2] - 1);
    // a strip of cells x0..x1 at (y, z): in XYZ order their keys are
    // consecutive, one range; in brick order each cell is its own range,
    // visited in ascending x (the same candidates in the same order)
    const bool brick = g.order == 1;
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int xa = x0; xa <= x1; xa = brick ? xa + 1 : x1 + 1) {
                const int xb = brick ? xa : x1;
                const uint32_t s = g.prefix[okey(g.order, g.n, xa, y, z)];
                const uint32_t e = g.prefix[okey(g.order, g.n, xb, y, z) + 1];
                for (uint32_t q = s; q < e; ++q) {
                    const uint32_t j = g.sorted ? g.sorted[q] : q;
                    const float dx = fx - static_cast<float>(xs[j]);
                    const float dy = fy - static_cast<float>(ys[j]);
                    const float dz = fz - static_cast<float>(zs[j]);
                    const float d2 = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
                    if (d2 <= g.r2)
                        f(j);
                }
            }
}

// ---- smoothing kernels (src/kernels.cuh:19-174) ---------------------------

constexpr double kPi = 3.14159265358979323846264338327950288;

template <class T> struct Kern {
    int kind;
    T hinv, ws, dws;
};

// alpha per family: exact in double for the oracle, rounded to float for the
// port as the reference's constructors do (kernels.cuh:32,81,104,124,151)
template <class T> Kern<T> make_kern(int kind, float h_bar)
{
    double alpha = 0.0;
    switch (kind) {
    case 0:
        alpha = 21.0 / (2.0 * kPi);
        break;
    case 1:
        alpha = 16.0 / kPi;
        break;
    case 2:
        alpha = 1365.0 / (64.0 * kPi);
        break;
    case 3:
        alpha = kPi / (4.0 * kPi * kPi - 30.0);
        break;
    }
    Kern<T> k;
    k.kind = kind;
    const T a = static_cast<T>(alpha);
    const T h = static_cast<T>(h_bar);
    k.hinv = T(1) / h;
    k.ws = a / (h * h * h);
    k.dws = a / (h * h * h * h);
    return k;
}

template <class T> inline T sat(T a) { return std::fmin(std::fmax(a, T(0)), T(1)); }

template <class T> inline T shape_w(int kind, T q)
{
    const T t1 = std::fmax(T(0), T(1) - q);
    switch (kind) {
    case 0: { // Wendland C2: (1-q)^4 (4q+1)
        const T t2 = t1 * t1;
        return t2 * t2 * (T(4) * q + T(1));
    }
    case 1: { // cubic B-spline: (1-q)^3 - 4 (1/2-q)^3
        const T t2 = sat(T(0.5) - q);
        return t1 * t1 * t1 - T(4) * t2 * t2 * t2;
    }
    case 2: { // Wendland C6: (1-q)^8 (32q^3 + 25q^2 + 8q + 1)
        const T t2 = t1 * t1, t4 = t2 * t2, q2 = q * q;
        return t4 * t4 * (T(32) * q2 * q + T(25) * q2 + T(8) * q + T(1));
    }
    default: { // double cosine: 4 cos(pi s) + cos(2 pi s) + 3, s = min(1,q)
        const T pis = static_cast<T>(kPi) * std::fmin(T(1), q);
        return T(4) * std::cos(pis) + std::cos(T(2) * pis) + T(3);
    }
    }
}

template <class T> inline T shape_dw(int kind, T q)
{
    const T t1 = std::fmax(T(0), T(1) - q);
    switch (kind) {
    case 0:
        return t1 * t1 * t1 * (T(-20) * q);
    case 1: {
        const T t2 = sat(T(0.5) - q);
        return T(-3) * t1 * t1 + T(12) * t2 * t2;
    }
    case 2: {
        const T t2 = t1 * t1;
        const T t7 = t2 * t2 * t2 * t1;
        return t7 * ((T(16) * q * q + T(7) * q + T(1)) * (T(-22) * q));
    }
    default: {
        const T pis = static_cast<T>(kPi) * std::fmin(T(1), q);
        const T two_pi = static_cast<T>(2.0 * kPi);
        return T(-2) * two_pi * std::sin(pis) - std::sin(T(2) * pis) * two_pi;
    }
    }
}

template <class T> inline T kW(const Kern<T>& k, T d2)
{
    const T len = std::sqrt(d2);
    return k.ws * shape_w(k.kind, sat(len * k.hinv));
}

template <class T> inline void kdW(const Kern<T>& k, T x, T y, T z, T d2, T out[3])
{
    const T len = std::sqrt(d2);
    T scale = k.dws * shape_dw(k.kind, len * k.hinv) / len;
    if (!std::isfinite(scale))
        scale = T(0);
    out[0] = scale * x;
    out[1] = scale * y;
    out[2] = scale * z;
}

template <class T> inline T dot3(const T a[3], const T b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// ---- views ------------------------------------------------------------------

template <class T> struct Fluid {
    T *x, *y, *z, *vx, *vy, *vz, *m;
    uint32_t n;
};
template <class T> struct Bdy {
    const T *x, *y, *z, *m;
    uint32_t n;
    OGrid grid;
};

struct SphPrm {
    int kernel;
    float h_bar, h, nu, rho0;
    float g[3];
};

// Calls f(j, xij[3], d2) over fluid neighbours in T.
template <class T, class F>
inline void nbrs(const OGrid& g, const Fluid<T>& s, uint32_t i, F&& f)
{
    const T xi = s.x[i], yi = s.y[i], zi = s.z[i];
    walk(g, s.x, s.y, s.z, xi, yi, zi, [&](uint32_t j) {
        const T d[3] = { xi - s.x[j], yi - s.y[j], zi - s.z[j] };
        f(j, d, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    });
}
template <class T, class F>
inline void bnbrs(const Bdy<T>* b, T xi, T yi, T zi, F&& f)
{
    if (!b || b->n == 0)
        return;
    walk(b->grid, b->x, b->y, b->z, xi, yi, zi, [&](uint32_t k) {
        const T d[3] = { xi - b->x[k], yi - b->y[k], zi - b->z[k] };
        f(k, d, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    });
}

// ---- phases (src/solvers/solver.cuh:18-105, solver.cu:3-32) ----------------

template <class T> void density(const OGrid& g, const Fluid<T>& s, const Bdy<T>* b, const Kern<T>& k, T* rho)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i) {
        T acc = 0, accb = 0;
        nbrs(g, s, uint32_t(i), [&](uint32_t j, const T*, T d2) { acc += s.m[j] * kW(k, d2); });
        bnbrs(b, s.x[i], s.y[i], s.z[i], [&](uint32_t j, const T*, T d2) { accb += b->m[j] * kW(k, d2); });
        rho[i] = acc + accb;
    }
}

template <class T>
void nonpressure(const OGrid& g, const Fluid<T>& s, const Kern<T>& k, const SphPrm& p, const T* rho, T* ax, T* ay, T* az)
{
    const T c = T(10) * T(p.nu);
    const T eps = T(0.01) * T(p.h) * T(p.h);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i) {
        T a[3] = { T(p.g[0]), T(p.g[1]), T(p.g[2]) };
        const T vi[3] = { s.vx[i], s.vy[i], s.vz[i] };
        nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
            const T vij[3] = { vi[0] - s.vx[j], vi[1] - s.vy[j], vi[2] - s.vz[j] };
            const T sc = c * s.m[j] / rho[j] * dot3(vij, d) / (d2 + eps);
            T dw[3];
            kdW(k, d[0], d[1], d[2], d2, dw);
            a[0] += sc * dw[0];
            a[1] += sc * dw[1];
            a[2] += sc * dw[2];
        });
        ax[i] = a[0];
        ay[i] = a[1];
        az[i] = a[2];
    }
}

template <class T> void integrate(const Fluid<T>& s, const T* ax, const T* ay, const T* az, T dt, bool positions)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i) {
        s.vx[i] += dt * ax[i];
        s.vy[i] += dt * ay[i];
        s.vz[i] += dt * az[i];
        if (positions) {
            s.x[i] += dt * s.vx[i];
            s.y[i] += dt * s.vy[i];
            s.z[i] += dt * s.vz[i];
        }
    }
}

// sum_k m_k grad W_ik over the boundary
template <class T> inline void bgrad(const Bdy<T>* b, const Kern<T>& k, T xi, T yi, T zi, T out[3])
{
    out[0] = out[1] = out[2] = 0;
    bnbrs(b, xi, yi, zi, [&](uint32_t j, const T* d, T d2) {
        T dw[3];
        kdW(k, d[0], d[1], d[2], d2, dw);
        out[0] += b->m[j] * dw[0];
        out[1] += b->m[j] * dw[1];
        out[2] += b->m[j] * dw[2];
    });
}

// ---- SESPH (src/solvers/SESPH.cu:6-95) -------------------------------------

template <class T>
void sesph_step(const OGrid& g, const Fluid<T>& s, const Bdy<T>* b, const SphPrm& p, T kstiff, T dt, T* rho, T* ax, T* ay, T* az)
{
    const Kern<T> k = make_kern<T>(p.kernel, p.h_bar);
    density(g, s, b, k, rho);
    const T rho0 = p.rho0;
    const T c = T(10) * T(p.nu);
    const T eps = T(0.01) * T(p.h) * T(p.h);
    auto eos = [&](T r) { return std::fmax(T(0), kstiff * (r / rho0 - T(1))); };
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i) {
        const T ri = rho[i];
        const T pri = eos(ri) / (ri * ri);
        const T vi[3] = { s.vx[i], s.vy[i], s.vz[i] };
        T a[3] = { T(p.g[0]), T(p.g[1]), T(p.g[2]) };
        nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
            const T rj = rho[j], mj = s.m[j];
            const T vij[3] = { vi[0] - s.vx[j], vi[1] - s.vy[j], vi[2] - s.vz[j] };
            T dw[3];
            kdW(k, d[0], d[1], d[2], d2, dw);
            const T visc = c * mj / rj * dot3(vij, d) / (d2 + eps);
            const T prs = mj * (pri + eos(rj) / (rj * rj));
            for (int q = 0; q < 3; ++q)
                a[q] += (visc - prs) * dw[q];
        });
        T gb[3];
        bgrad(b, k, s.x[i], s.y[i], s.z[i], gb);
        ax[i] = a[0] - pri * gb[0];
        ay[i] = a[1] - pri * gb[1];
        az[i] = a[2] - pri * gb[2];
    }
    integrate(s, ax, ay, az, dt, true);
}

// ---- PCISPH (src/solvers/PCISPH.cu:15-111, PCISPH.cuh:49-77) --------------

template <class T>
uint32_t pcisph_step(const OGrid& g, const Fluid<T>& s, const Bdy<T>* b, const SphPrm& p, double delta, float eta,
    uint32_t min_iter, uint32_t max_iter, uint32_t fixed_iters, T dt, T* rho, T* ax, T* ay, T* az, double* avg_out)
{
    const Kern<T> k = make_kern<T>(p.kernel, p.h_bar);
    const T kk = static_cast<T>((delta / double(dt)) / double(dt));
    const T rho0 = p.rho0;
    density(g, s, b, k, rho);
    nonpressure(g, s, k, p, rho, ax, ay, az);
    std::vector<T> gb(3 * size_t(s.n));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i)
        bgrad(b, k, s.x[i], s.y[i], s.z[i], &gb[3 * size_t(i)]);
    const double thr = double((eta + 1.f) * p.rho0);
    uint32_t l = 0;
    double avg = 0;
    for (;;) {
        integrate(s, ax, ay, az, dt, false);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < long(s.n); ++i) {
            const T vi[3] = { s.vx[i], s.vy[i], s.vz[i] };
            T acc = 0;
            nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
                const T vij[3] = { vi[0] - s.vx[j], vi[1] - s.vy[j], vi[2] - s.vz[j] };
                T dw[3];
                kdW(k, d[0], d[1], d[2], d2, dw);
                acc += s.m[j] * (kW(k, d2) + dt * dot3(dw, vij));
            });
            bnbrs(b, s.x[i], s.y[i], s.z[i], [&](uint32_t j, const T* d, T d2) {
                T dw[3];
                kdW(k, d[0], d[1], d[2], d2, dw);
                acc += b->m[j] * (kW(k, d2) + dt * dot3(dw, vi));
            });
            rho[i] = acc;
        }
        double sum = 0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
        for (long i = 0; i < long(s.n); ++i)
            sum += double(rho[i]);
        avg = sum / double(s.n);
        const T r0sqinv = T(1) / (rho0 * rho0);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < long(s.n); ++i) {
            const T ri = rho[i];
            const T pi = std::fmax(T(0), kk * (ri - rho0));
            const T pri = pi / (ri * ri);
            const T pr0 = pi * r0sqinv;
            T a[3] = { 0, 0, 0 };
            nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
                const T rj = rho[j];
                const T pj = std::fmax(T(0), kk * (rj - rho0));
                T dw[3];
                kdW(k, d[0], d[1], d[2], d2, dw);
                const T c = s.m[j] * (pri + pj / (rj * rj));
                for (int q = 0; q < 3; ++q)
                    a[q] += c * dw[q];
            });
            const T* G = &gb[3 * size_t(i)];
            const T cb = T(0.5) * (pri + pr0);
            ax[i] = -a[0] - cb * G[0];
            ay[i] = -a[1] - cb * G[1];
            az[i] = -a[2] - cb * G[2];
        }
        ++l;
        if (fixed_iters) {
            if (l >= fixed_iters)
                break;
        } else {
            const bool keep = (l < min_iter) || (avg > thr);
            if (!keep || (max_iter && l >= max_iter))
                break;
        }
    }
    integrate(s, ax, ay, az, dt, true);
    if (avg_out)
        *avg_out = avg;
    return l;
}

// ---- IISPH (src/solvers/IISPH.cu:10-262) -----------------------------------

template <class T> inline T jacobi_update(T omega, T aii, T Ap, T s)
{
    return (double(aii) < 1e-6 && double(aii) > -1e-6) ? T(0) : std::fmax(T(0), omega * (s - Ap) / aii);
}

struct IisphPrm {
    float omega, eta;
    uint32_t min_iter, max_iter;
    int stop_rule; // 0 any particle above eta rho0, 1 mean error <= eta
    int warm_start;
};

template <class T>
uint32_t iisph_step(const OGrid& g, const Fluid<T>& s, const Bdy<T>* b, const SphPrm& p, const IisphPrm& ip,
    uint32_t fixed_iters, T dt, T* rho, T* pr, T* aii, T* si, T* ax, T* ay, T* az, double* mean_err_out,
    uint32_t* viol_out)
{
    const Kern<T> k = make_kern<T>(p.kernel, p.h_bar);
    const T rho0 = p.rho0;
    const T omega = ip.omega;
    density(g, s, b, k, rho);
    nonpressure(g, s, k, p, rho, ax, ay, az);
    integrate(s, ax, ay, az, dt, false);
    std::vector<T> gb(3 * size_t(s.n));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(s.n); ++i) {
        T* G = &gb[3 * size_t(i)];
        bgrad(b, k, s.x[i], s.y[i], s.z[i], G);
        const T vi[3] = { s.vx[i], s.vy[i], s.vz[i] };
        T gf[3] = { 0, 0, 0 };
        T sq = 0, div = 0;
        nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
            T dw[3];
            kdW(k, d[0], d[1], d[2], d2, dw);
            const T mj = s.m[j];
            const T mdw[3] = { mj * dw[0], mj * dw[1], mj * dw[2] };
            const T vij[3] = { vi[0] - s.vx[j], vi[1] - s.vy[j], vi[2] - s.vz[j] };
            div += dot3(mdw, vij);
            sq += dot3(mdw, dw);
            for (int q = 0; q < 3; ++q)
                gf[q] += mdw[q];
        });
        const T Gs[3] = { gf[0] + G[0], gf[1] + G[1], gf[2] + G[2] };
        const T a = -dt * (dot3(Gs, Gs) + s.m[i] * sq) / (rho0 * rho0);
        const T sv = (rho0 - rho[i]) / dt - (div + dot3(vi, G));
        aii[i] = a;
        si[i] = sv;
        pr[i] = ip.warm_start ? pr[i] * T(0.5) : jacobi_update(omega, a, T(0), sv);
    }
    const T r0sqinv = T(1) / (rho0 * rho0);
    auto pacc = [&]() {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < long(s.n); ++i) {
            const T pi = pr[i];
            T a[3] = { 0, 0, 0 };
            nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
                T dw[3];
                kdW(k, d[0], d[1], d[2], d2, dw);
                const T c = s.m[j] * (pi + pr[j]);
                for (int q = 0; q < 3; ++q)
                    a[q] += c * dw[q];
            });
            const T* G = &gb[3 * size_t(i)];
            ax[i] = -r0sqinv * (a[0] + pi * G[0]);
            ay[i] = -r0sqinv * (a[1] + pi * G[1]);
            az[i] = -r0sqinv * (a[2] + pi * G[2]);
        }
    };
    const T thr = T(ip.eta * p.rho0);
    std::vector<T> pnew(s.n);
    uint32_t l = 0;
    double mean = 0;
    uint32_t viol = 0;
    for (;;) {
        pacc();
        double esum = 0;
        long vcount = 0;
#pragma omp parallel for reduction(+ : esum, vcount) schedule(static)
        for (long i = 0; i < long(s.n); ++i) {
            const T ai[3] = { ax[i], ay[i], az[i] };
            T lap = 0;
            nbrs(g, s, uint32_t(i), [&](uint32_t j, const T* d, T d2) {
                T dw[3];
                kdW(k, d[0], d[1], d[2], d2, dw);
                const T aij[3] = { ai[0] - ax[j], ai[1] - ay[j], ai[2] - az[j] };
                lap += s.m[j] * dot3(aij, dw);
            });
            const T* G = &gb[3 * size_t(i)];
            lap += dot3(ai, G);
            const T Ap = dt * lap;
            pnew[i] = pr[i] + jacobi_update(omega, aii[i], Ap, si[i]);
            const T err = (Ap - si[i]) * dt;
            vcount += err > thr ? 1 : 0;
            esum += err > T(0) ? double(err) : 0.0;
        }
        std::copy(pnew.begin(), pnew.end(), pr);
        ++l;
        mean = s.n ? esum / double(s.n) / double(p.rho0) : 0.0;
        viol = uint32_t(vcount);
        if (fixed_iters) {
            if (l >= fixed_iters)
                break;
        } else {
            const bool ok = ip.stop_rule == 1 ? (mean <= double(ip.eta)) : (viol == 0);
            if ((l >= ip.min_iter && ok) || (ip.max_iter && l >= ip.max_iter) || l >= 100000)
                break;
        }
    }
    pacc();
    integrate(s, ax, ay, az, dt, true);
    if (mean_err_out)
        *mean_err_out = mean;
    if (viol_out)
        *viol_out = viol;
    return l;
}

// ---- boundary masses (src/scene/sample_boundary.cu:251-297) ----------------

template <class T> void boundary_masses(const Bdy<T>& b, T* m, float h, float rho0, uint32_t iters)
{
    const Kern<T> k = make_kern<T>(1, 2.f * h);
    std::vector<T> old(b.n, T(1));
    for (uint32_t it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < long(b.n); ++i) {
            T sum = 0;
            const T xi = b.x[i], yi = b.y[i], zi = b.z[i];
            walk(b.grid, b.x, b.y, b.z, xi, yi, zi, [&](uint32_t j) {
                const T d[3] = { xi - b.x[j], yi - b.y[j], zi - b.z[j] };
                sum += old[j] * kW(k, d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            });
            m[i] = T(rho0) * old[i] / sum;
        }
        std::copy(m, m + b.n, old.begin());
    }
}

} // namespace

// ============================ extern "C" surface ============================

struct orc_grid {
    float bmin[3];
    float cell;
    float r2;
    int32_t n[3];
    int32_t range;
    const uint32_t* prefix;
    const uint32_t* sorted;
    int32_t order;
};

static OGrid to_og(const orc_grid* g)
{
    OGrid o;
    for (int a = 0; a < 3; ++a) {
        o.bmin[a] = g->bmin[a];
        o.n[a] = g->n[a];
    }
    o.cell = g->cell;
    o.r2 = g->r2;
    o.range = g->range;
    o.prefix = g->prefix;
    o.sorted = g->sorted;
    o.order = g->order;
    return o;
}

ORC_API int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

ORC_API void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// UniformGridBuilder geometry (uniformgrid.cu:65-96) with cell = ratio * r;
// in brick order the cell counts are rounded up to even.
ORC_API int orc_grid_geometry_ordered(const float* bmin, const float* bmax, float r, float ratio, int order,
    orc_grid* out)
{
    const float cell = r * ratio;
    const float pad = static_cast<float>(r * 2.0);
    for (int a = 0; a < 3; ++a) {
        const float lo = bmin[a] - pad;
        const float hi = bmax[a] + pad;
        const float e = std::ceil((hi - lo) / cell);
        if (!(e > 0.f))
            return -1;
        out->bmin[a] = lo;
        out->n[a] = static_cast<int32_t>(e);
        if (order == 1)
            out->n[a] += out->n[a] & 1;
    }
    out->cell = cell;
    out->r2 = r * r;
    out->range = static_cast<int32_t>(std::ceil(1.0 / double(ratio) - 1e-6));
    out->prefix = nullptr;
    out->sorted = nullptr;
    out->order = order;
    return 0;
}

ORC_API int orc_grid_geometry(const float* bmin, const float* bmax, float r, float ratio, orc_grid* out)
{
    return orc_grid_geometry_ordered(bmin, bmax, r, ratio, 0, out);
}

// Counting sort (uniformgrid.cu:20-63, 98-143) with ascending index inside a
// cell. Returns the number of points outside the query-safe cells, or -1 (with
// the tables complete) when a cell holds more than walk_cell_cap(range) points.
ORC_API long orc_grid_build(const orc_grid* geom, const float* x, const float* y, const float* z, uint32_t n,
    uint32_t* counts, uint32_t* prefix, uint32_t* sorted, uint32_t* keys)
{
    const long nx = geom->n[0], ny = geom->n[1], nz = geom->n[2];
    const size_t C = size_t(nx) * ny * nz;
    std::memset(counts, 0, C * sizeof(uint32_t));
    long outside = 0;
    const int R = geom->range;
    for (uint32_t i = 0; i < n; ++i) {
        const int cx = cell_axis(x[i], geom->bmin[0], geom->cell, int(nx));
        const int cy = cell_axis(y[i], geom->bmin[1], geom->cell, int(ny));
        const int cz = cell_axis(z[i], geom->bmin[2], geom->cell, int(nz));
        keys[i] = okey(geom->order, geom->n, cx, cy, cz);
        counts[keys[i]]++;
        if (cx < R || cy < R || cz < R || cx > nx - 1 - R || cy > ny - 1 - R || cz > nz - 1 - R)
            ++outside;
    }
    uint32_t run = 0;
    for (size_t c = 0; c < C; ++c) {
        prefix[c] = run;
        run += counts[c];
    }
    prefix[C] = run;
    std::vector<uint32_t> fill(prefix, prefix + C);
    for (uint32_t i = 0; i < n; ++i)
        sorted[fill[keys[i]]++] = i;
    const uint32_t cap = walk_cell_cap(R);
    for (size_t c = 0; c < C; ++c)
        if (counts[c] > cap)
            return -1;
    return outside;
}

ORC_API uint32_t orc_grid_cell_capacity(int range) { return walk_cell_cap(range); }

// Reference grid test statistics (uniformgrid.cu:185-208) through the walk.
ORC_API void orc_query_stats(const orc_grid* geom, const float* x, const float* y, const float* z, uint32_t n,
    uint32_t* count, float* len2, float* vx, float* vy, float* vz)
{
    const OGrid g = to_og(geom);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
        uint32_t c = 0;
        float l2 = 0, a = 0, b = 0, d = 0;
        const float xi = x[i], yi = y[i], zi = z[i];
        walk(g, x, y, z, xi, yi, zi, [&](uint32_t j) {
            const float dx = xi - x[j], dy = yi - y[j], dz = zi - z[j];
            c += (j == uint32_t(i)) ? 0u : 1u;
            l2 += std::fma(dx, dx, std::fma(dy, dy, dz * dz));
            a += dx;
            b += dy;
            d += dz;
        });
        count[i] = c;
        len2[i] = l2;
        vx[i] = a;
        vy[i] = b;
        vz[i] = d;
    }
}

// Brute-force O(N^2) counterpart (uniformgrid.cu:209-229).
ORC_API void orc_brute_stats(float r2, const float* x, const float* y, const float* z, uint32_t n, uint32_t* count,
    float* len2, float* vx, float* vy, float* vz)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < long(n); ++i) {
        uint32_t c = 0;
        float l2 = 0, a = 0, b = 0, d = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const float dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            const float q = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
            if (q <= r2) {
                c += (j == uint32_t(i)) ? 0u : 1u;
                l2 += q;
                a += dx;
                b += dy;
                d += dz;
            }
        }
        count[i] = c;
        len2[i] = l2;
        vx[i] = a;
        vy[i] = b;
        vz[i] = d;
    }
}

// ---- the reference tests' own inputs ---------------------------------------
// Point sets generated with the same standard-library calls as the
// reference's tests, so the tests here run on the very inputs the reference's
// tests use.

// src/datastructure/uniformgrid.cu:257-275: side^3 points uniform in
// [0, box]^3 from std::mt19937(161) and uniform_real_distribution<float>.
ORC_API void orc_reference_grid_points(uint32_t side, float h, float* x, float* y, float* z)
{
    const uint32_t n = side * side * side;
    const float box_size = h * static_cast<float>(side);
    std::mt19937 rng(161);
    std::uniform_real_distribution<float> uniform_dist(0.f, 1.f);
    for (uint32_t i = 0; i < n; ++i) {
        x[i] = box_size * uniform_dist(rng);
        y[i] = box_size * uniform_dist(rng);
        z[i] = box_size * uniform_dist(rng);
    }
}

// src/kernels.cu:31-52: 9 edge cases then std::srand(16142069) samples in
// [0, 1.5 h_bar]^3 for N = 25^3.
ORC_API void orc_reference_kernel_points(float h_bar, uint32_t n, float* x, float* y, float* z)
{
    const float EPS = std::numeric_limits<float>::epsilon();
    const double edge[9] = { 0.0, h_bar, double(h_bar - EPS), double(h_bar + EPS), double(EPS),
        double(static_cast<float>(h_bar + 1e-20)), double(static_cast<float>(h_bar - 1e-20)), 1e-20,
        double(std::numeric_limits<float>::max()) };
    for (uint32_t i = 0; i < 9 && i < n; ++i)
        x[i] = y[i] = z[i] = static_cast<float>(edge[i]);
    std::srand(16142069);
    for (uint32_t i = 9; i < n; ++i) {
        const float a = static_cast<float>(h_bar * 1.5 * static_cast<float>(std::rand()) / static_cast<float>(RAND_MAX));
        const float b = static_cast<float>(h_bar * 1.5 * static_cast<float>(std::rand()) / static_cast<float>(RAND_MAX));
        const float c = static_cast<float>(h_bar * 1.5 * static_cast<float>(std::rand()) / static_cast<float>(RAND_MAX));
        x[i] = a;
        y[i] = b;
        z[i] = c;
    }
}

// Brute force for the query points idx[0..nq) against all n points.
ORC_API void orc_brute_stats_subset(float r2, const float* x, const float* y, const float* z, uint32_t n,
    const uint32_t* idx, uint32_t nq, uint32_t* count, float* len2, float* vx, float* vy, float* vz)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long q = 0; q < long(nq); ++q) {
        const uint32_t i = idx[q];
        uint32_t c = 0;
        float l2 = 0, a = 0, b = 0, d = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const float dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            const float s2 = std::fma(dx, dx, std::fma(dy, dy, dz * dz));
            if (s2 <= r2) {
                c += (j == i) ? 0u : 1u;
                l2 += s2;
                a += dx;
                b += dy;
                d += dz;
            }
        }
        count[q] = c;
        len2[q] = l2;
        vx[q] = a;
        vy[q] = b;
        vz[q] = d;
    }
}

// Neighbour lists through the walk: for each i, its neighbours in walk order
// (offsets has n+1 entries; call once with list == NULL for the sizes).
ORC_API void orc_neighbour_lists(const orc_grid* geom, const float* x, const float* y, const float* z, uint32_t n,
    uint32_t* offsets, uint32_t* list)
{
    const OGrid g = to_og(geom);
    if (!list) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < long(n); ++i) {
            uint32_t c = 0;
            walk(g, x, y, z, x[i], y[i], z[i], [&](uint32_t) { ++c; });
            offsets[i + 1] = c;
        }
        offsets[0] = 0;
        for (uint32_t i = 0; i < n; ++i)
            offsets[i + 1] += offsets[i];
        return;
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
        uint32_t o = offsets[i];
        walk(g, x, y, z, x[i], y[i], z[i], [&](uint32_t j) { list[o++] = j; });
    }
}

ORC_API void orc_gather_f32(const uint32_t* sorted, uint32_t n, const float* src, float* dst)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i)
        dst[i] = src[sorted[i]];
}

#define ORC_KERNEL_EVAL(T, SUF)                                                                                    \
    ORC_API void orc_kernel_eval_##SUF(int kind, float h_bar, const T* dx, const T* dy, const T* dz, uint32_t n,  \
        T* w, T* gx, T* gy, T* gz)                                                                                 \
    {                                                                                                              \
        const Kern<T> k = make_kern<T>(kind, h_bar);                                                               \
        for (uint32_t i = 0; i < n; ++i) {                                                                         \
            const T d2 = dx[i] * dx[i] + dy[i] * dy[i] + dz[i] * dz[i];                                            \
            w[i] = kW(k, d2);                                                                                      \
            T g[3];                                                                                                \
            kdW(k, dx[i], dy[i], dz[i], d2, g);                                                                    \
            gx[i] = g[0];                                                                                          \
            gy[i] = g[1];                                                                                          \
            gz[i] = g[2];                                                                                          \
        }                                                                                                          \
    }
ORC_KERNEL_EVAL(double, f64)
ORC_KERNEL_EVAL(float, f32)

// shape functions alone, for the known-answer tests of the published forms
ORC_API double orc_shape_w(int kind, double q) { return shape_w<double>(kind, q); }
ORC_API double orc_shape_dw(int kind, double q) { return shape_dw<double>(kind, q); }

struct orc_sph {
    int32_t kernel;
    float h_bar, h, nu, rho0;
    float g[3];
};
struct orc_iisph {
    float omega, eta;
    uint32_t min_iter, max_iter;
    int32_t stop_rule, warm_start;
};

static SphPrm to_prm(const orc_sph* p)
{
    SphPrm s;
    s.kernel = p->kernel;
    s.h_bar = p->h_bar;
    s.h = p->h;
    s.nu = p->nu;
    s.rho0 = p->rho0;
    for (int a = 0; a < 3; ++a)
        s.g[a] = p->g[a];
    return s;
}

#define ORC_SOLVERS(T, SUF)                                                                                         \
    ORC_API void orc_density_##SUF(const orc_grid* fg, T* x, T* y, T* z, T* m, uint32_t n, const orc_grid* bg,      \
        const T* bx, const T* by, const T* bz, const T* bm, uint32_t nb, int kernel, float h_bar, T* rho)             \
    {                                                                                                               \
        Fluid<T> s { x, y, z, nullptr, nullptr, nullptr, m, n };                                                    \
        Bdy<T> b { bx, by, bz, bm, nb, bg ? to_og(bg) : OGrid {} };                                                 \
        density(to_og(fg), s, nb ? &b : nullptr, make_kern<T>(kernel, h_bar), rho);                                 \
    }                                                                                                               \
    ORC_API void orc_nonpressure_##SUF(const orc_grid* fg, T* x, T* y, T* z, T* vx, T* vy, T* vz, T* m, uint32_t n, \
        const orc_sph* p, const T* rho, T* ax, T* ay, T* az)                                                         \
    {                                                                                                               \
        Fluid<T> s { x, y, z, vx, vy, vz, m, n };                                                                   \
        const SphPrm q = to_prm(p);                                                                                 \
        nonpressure(to_og(fg), s, make_kern<T>(q.kernel, q.h_bar), q, rho, ax, ay, az);                             \
    }                                                                                                               \
    ORC_API void orc_sesph_step_##SUF(const orc_grid* fg, T* x, T* y, T* z, T* vx, T* vy, T* vz, T* m, uint32_t n,  \
        const orc_grid* bg, const T* bx, const T* by, const T* bz, const T* bm, uint32_t nb, const orc_sph* p,       \
        float k, float dt, T* rho, T* ax, T* ay, T* az)                                                              \
    {                                                                                                               \
        Fluid<T> s { x, y, z, vx, vy, vz, m, n };                                                                   \
        Bdy<T> b { bx, by, bz, bm, nb, bg ? to_og(bg) : OGrid {} };                                                 \
        sesph_step(to_og(fg), s, nb ? &b : nullptr, to_prm(p), T(k), T(dt), rho, ax, ay, az);                       \
    }                                                                                                               \
    ORC_API uint32_t orc_pcisph_step_##SUF(const orc_grid* fg, T* x, T* y, T* z, T* vx, T* vy, T* vz, T* m,         \
        uint32_t n, const orc_grid* bg, const T* bx, const T* by, const T* bz, const T* bm, uint32_t nb,             \
        const orc_sph* p, double delta, float eta, uint32_t min_iter, uint32_t max_iter, uint32_t fixed_iters,     \
        float dt, T* rho, T* ax, T* ay, T* az, double* avg)                                                          \
    {                                                                                                               \
        Fluid<T> s { x, y, z, vx, vy, vz, m, n };                                                                   \
        Bdy<T> b { bx, by, bz, bm, nb, bg ? to_og(bg) : OGrid {} };                                                 \
        return pcisph_step(to_og(fg), s, nb ? &b : nullptr, to_prm(p), delta, eta, min_iter, max_iter,             \
            fixed_iters, T(dt), rho, ax, ay, az, avg);                                                              \
    }                                                                                                               \
    ORC_API uint32_t orc_iisph_step_##SUF(const orc_grid* fg, T* x, T* y, T* z, T* vx, T* vy, T* vz, T* m,          \
        uint32_t n, const orc_grid* bg, const T* bx, const T* by, const T* bz, const T* bm, uint32_t nb,             \
        const orc_sph* p, const orc_iisph* ip, uint32_t fixed_iters, float dt, T* rho, T* pr, T* aii, T* si, T* ax, \
        T* ay, T* az, double* mean_err, uint32_t* viol)                                                              \
    {                                                                                                               \
        Fluid<T> s { x, y, z, vx, vy, vz, m, n };                                                                   \
        Bdy<T> b { bx, by, bz, bm, nb, bg ? to_og(bg) : OGrid {} };                                                 \
        IisphPrm q { ip->omega, ip->eta, ip->min_iter, ip->max_iter, ip->stop_rule, ip->warm_start };              \
        return iisph_step(to_og(fg), s, nb ? &b : nullptr, to_prm(p), q, fixed_iters, T(dt), rho, pr, aii, si, ax,  \
            ay, az, mean_err, viol);                                                                                \
    }                                                                                                               \
    ORC_API void orc_boundary_masses_##SUF(const orc_grid* bg, const T* bx, const T* by, const T* bz, uint32_t nb,  \
        T* m, float h, float rho0, uint32_t iters)                                                                   \
    {                                                                                                               \
        Bdy<T> b { bx, by, bz, nullptr, nb, to_og(bg) };                                                            \
        boundary_masses(b, m, h, rho0, iters);                                                                      \
    }
ORC_SOLVERS(double, f64)
ORC_SOLVERS(float, f32)

// CFL time step (src/timestep/cfl.cu:5-37) with max |v| in FP64.
ORC_API float orc_cfl_time_step(float lambda, float h, const float* vx, const float* vy, const float* vz, uint32_t n,
    const float* g, double* vmax_out)
{
    double vmax = 0.0;
#pragma omp parallel for reduction(max : vmax) schedule(static)
    for (long i = 0; i < long(n); ++i) {
        const double a = vx[i], b = vy[i], c = vz[i];
        double v = std::sqrt(a * a + b * b + c * c);
        if (v != v)
            v = INFINITY;
        vmax = std::max(vmax, v);
    }
    if (vmax_out)
        *vmax_out = vmax;
    const float dt_vel = static_cast<float>(double(lambda * h) / vmax);
    const double gn = std::sqrt(double(g[0]) * g[0] + double(g[1]) * g[1] + double(g[2]) * g[2]);
    const float dt_g = std::sqrt(static_cast<float>(2.0 * double(lambda) * double(h) / gn));
    return std::fmin(dt_vel, dt_g);
}

// Scene::hard_enforce_bounds (src/scene/scene.cu:466-483).
ORC_API void orc_hard_enforce_bounds(const float* lo, const float* hi, float* x, float* y, float* z, float* vx,
    float* vy, float* vz, uint32_t n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
        if (x[i] <= lo[0] || y[i] <= lo[1] || z[i] <= lo[2] || hi[0] <= x[i] || hi[1] <= y[i] || hi[2] <= z[i]) {
            const float c = static_cast<float>(-0.1);
            vx[i] *= c;
            vy[i] *= c;
            vz[i] *= c;
        }
        x[i] = std::fmax(std::fmin(x[i], hi[0]), lo[0]);
        y[i] = std::fmax(std::fmin(y[i], hi[1]), lo[1]);
        z[i] = std::fmax(std::fmin(z[i], hi[2]), lo[2]);
    }
}

// PCISPH template-particle stiffness (src/solvers/PCISPH.cuh:49-77), in T.
ORC_API double orc_pcisph_delta_f64(int kind, float h_bar, float h)
{
    const Kern<double> k = make_kern<double>(kind, h_bar);
    double sx = 0, sy = 0, sz = 0, sum_sq = 0;
    for (int x = -2; x <= 2; ++x)
        for (int y = -2; y <= 2; ++y)
            for (int z = -2; z <= 2; ++z) {
                const double d[3] = { -double(x) * h, -double(y) * h, -double(z) * h };
                double g[3];
                kdW(k, d[0], d[1], d[2], d[0] * d[0] + d[1] * d[1] + d[2] * d[2], g);
                sx += g[0];
                sy += g[1];
                sz += g[2];
                sum_sq += g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
            }
    const double V0 = double(h) * h * h;
    return 1.0 / (2.0 * V0 * V0 * (sx * sx + sy * sy + sz * sz + sum_sq));
}

// ============================ the CPU baseline ==============================
// The float port of the reference main loop (src/main.cu:60-78) as the CPU
// baseline runs it: the grid is a parallel counting sort (atomic per-cell
// counts as uniformgrid.cu:20-34, rocThrust's OpenMP exclusive scan as
// uniformgrid.cu:128-133, a scatter through atomic slots as :40-63, then each
// cell's indices put in ascending order, so the result is the serial build's
// bit for bit); the reorder is rocThrust's OpenMP gather (buffer.cuh:184-209);
// every neighbour loop is an OpenMP loop over particles. Whole steps are timed
// with std::chrono::steady_clock.

namespace {

struct PortGrid {
    std::vector<uint32_t> counts, prefix, sorted, keys, fill;
    std::vector<float> tmp;
};

void grid_build_par(const orc_grid* geom, const float* x, const float* y, const float* z, uint32_t n, PortGrid& w)
{
    const long nx = geom->n[0], ny = geom->n[1], nz = geom->n[2];
    const size_t C = size_t(nx) * ny * nz;
    w.counts.assign(C + 1, 0u);
    w.prefix.resize(C + 1);
    w.sorted.resize(n);
    w.keys.resize(n);
    uint32_t* counts = w.counts.data();
    uint32_t* keys = w.keys.data();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
        const int cx = cell_axis(x[i], geom->bmin[0], geom->cell, int(nx));
        const int cy = cell_axis(y[i], geom->bmin[1], geom->cell, int(ny));
        const int cz = cell_axis(z[i], geom->bmin[2], geom->cell, int(nz));
        const uint32_t k = okey(geom->order, geom->n, cx, cy, cz);
        keys[i] = k;
#pragma omp atomic
        counts[k] += 1u;
    }
    thrust::exclusive_scan(thrust::omp::par, counts, counts + C + 1, w.prefix.data());
    w.fill.assign(w.prefix.begin(), w.prefix.end() - 1);
    uint32_t* fill = w.fill.data();
    uint32_t* sorted = w.sorted.data();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
        uint32_t slot;
#pragma omp atomic capture
        slot = fill[keys[i]]++;
        sorted[slot] = uint32_t(i);
    }
    const uint32_t* prefix = w.prefix.data();
#pragma omp parallel for schedule(dynamic, 4096)
    for (long c = 0; c < long(C); ++c)
        if (prefix[c + 1] - prefix[c] > 1)
            std::sort(sorted + prefix[c], sorted + prefix[c + 1]);
}

// fields[k] <- fields[k][sorted] through w.tmp
void reorder_par(PortGrid& w, float* const* fields, uint32_t nfields, uint32_t n)
{
    w.tmp.resize(n);
    for (uint32_t k = 0; k < nfields; ++k) {
        thrust::gather(thrust::omp::par, w.sorted.begin(), w.sorted.end(), fields[k], w.tmp.begin());
        std::copy(w.tmp.begin(), w.tmp.end(), fields[k]);
    }
}

// cfl_time_step / simple_dt_controller (src/timestep/cfl.cu:5-52), max |v| in FP64
float port_cfl(float lambda, float h, const float* vx, const float* vy, const float* vz, uint32_t n, const float* g)
{
    const double vmax = thrust::transform_reduce(
        thrust::omp::par, thrust::counting_iterator<uint32_t>(0), thrust::counting_iterator<uint32_t>(n),
        [=](uint32_t i) {
            const double a = vx[i], b = vy[i], c = vz[i];
            const double v = std::sqrt(a * a + b * b + c * c);
            return v != v ? double(INFINITY) : v;
        },
        0.0, thrust::maximum<double>());
    const float dt_vel = static_cast<float>(double(lambda * h) / vmax);
    const double gn = std::sqrt(double(g[0]) * g[0] + double(g[1]) * g[1] + double(g[2]) * g[2]);
    const float dt_g = std::sqrt(static_cast<float>(2.0 * double(lambda) * double(h) / gn));
    return std::fmin(dt_vel, dt_g);
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

ORC_API void orc_grid_build_par(const orc_grid* geom, const float* x, const float* y, const float* z, uint32_t n,
    uint32_t* prefix, uint32_t* sorted)
{
    PortGrid w;
    grid_build_par(geom, x, y, z, n, w);
    std::copy(w.prefix.begin(), w.prefix.end(), prefix);
    std::copy(w.sorted.begin(), w.sorted.end(), sorted);
}

// `steps` steps of the IISPH main loop (main.cu:60-78) in float: the time-step
// controller (lambda, iter_target, change_factor) from *dt and *iters, the
// grid with the reorder of x, v, m and the warm-start pressure, the IISPH step,
// the bounds. The state arrays are updated in place; step_seconds[k] and
// step_iters[k] receive each step's wall time and iteration count.
ORC_API void orc_port_iisph_run(const orc_grid* geom, float* x, float* y, float* z, float* vx, float* vy, float* vz,
    float* m, float* p, uint32_t n, const orc_grid* bg, const float* bx, const float* by, const float* bz,
    const float* bm, uint32_t nb, const orc_sph* prm, const orc_iisph* ip, const float* lo, const float* hi,
    float lambda, uint32_t iter_target, float change_factor, float* dt, uint32_t* iters, uint32_t steps,
    double* step_seconds, uint32_t* step_iters)
{
    PortGrid w;
    std::vector<float> rho(n), aii(n), si(n), ax(n), ay(n), az(n);
    const SphPrm q = to_prm(prm);
    const IisphPrm iq { ip->omega, ip->eta, ip->min_iter, ip->max_iter, ip->stop_rule, ip->warm_start };
    Bdy<float> b { bx, by, bz, bm, nb, bg ? to_og(bg) : OGrid {} };
    float* fields[8] = { x, y, z, vx, vy, vz, m, p };
    for (uint32_t k = 0; k < steps; ++k) {
        const auto t0 = std::chrono::steady_clock::now();
        const float factor = *iters < iter_target ? 1.f + change_factor : 1.f - change_factor;
        *dt = std::fmin(port_cfl(lambda, q.h, vx, vy, vz, n, q.g), *dt * factor);
        grid_build_par(geom, x, y, z, n, w);
        reorder_par(w, fields, 8, n);
        orc_grid g = *geom;
        g.prefix = w.prefix.data();
        g.sorted = nullptr;
        Fluid<float> s { x, y, z, vx, vy, vz, m, n };
        *iters = iisph_step(to_og(&g), s, nb ? &b : nullptr, q, iq, 0, *dt, rho.data(), p, aii.data(), si.data(),
            ax.data(), ay.data(), az.data(), nullptr, nullptr);
        orc_hard_enforce_bounds(lo, hi, x, y, z, vx, vy, vz, n);
        step_seconds[k] = seconds_since(t0);
        step_iters[k] = *iters;
    }
}

// `steps` steps of the SESPH main loop in float with dt = cfl_time_step(lambda)
// every step: grid with the reorder of x, v, m, the SESPH step (stiffness k),
// the bounds. *time receives the simulated time advanced.
ORC_API void orc_port_sesph_run(const orc_grid* geom, float* x, float* y, float* z, float* vx, float* vy, float* vz,
    float* m, uint32_t n, const orc_grid* bg, const float* bx, const float* by, const float* bz, const float* bm,
    uint32_t nb, const orc_sph* prm, float k, const float* lo, const float* hi, float lambda, uint32_t steps,
    double* step_seconds, double* time)
{
    PortGrid w;
    std::vector<float> rho(n), ax(n), ay(n), az(n);
    const SphPrm q = to_prm(prm);
    Bdy<float> b { bx, by, bz, bm, nb, bg ? to_og(bg) : OGrid {} };
    float* fields[7] = { x, y, z, vx, vy, vz, m };
    double t = 0.0;
    for (uint32_t s = 0; s < steps; ++s) {
        const auto t0 = std::chrono::steady_clock::now();
        const float dt = port_cfl(lambda, q.h, vx, vy, vz, n, q.g);
        grid_build_par(geom, x, y, z, n, w);
        reorder_par(w, fields, 7, n);
        orc_grid g = *geom;
        g.prefix = w.prefix.data();
        g.sorted = nullptr;
        Fluid<float> st { x, y, z, vx, vy, vz, m, n };
        sesph_step(to_og(&g), st, nb ? &b : nullptr, q, k, dt, rho.data(), ax.data(), ay.data(), az.data());
        orc_hard_enforce_bounds(lo, hi, x, y, z, vx, vy, vz, n);
        step_seconds[s] = seconds_since(t0);
        t += dt;
    }
    *time = t;
}
