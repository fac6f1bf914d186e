// Triangle meshes: Wavefront OBJ loading on the host, stratified area sampling
// of the surface and a generalized-winding-number inside test on the GPU.
//
// The reference loads OBJ files with tinyobjloader (src/scene/loader.cpp:13-107:
// polygons triangulated by the loader, vertex indices checked, shapes skipped
// by name),
// samples the surface by inverting the CDF of the face areas with one stratum
// per sample (sample_boundary.cu:179-249, 299-389; count ceil(area /
// (1.5 sqrt(3) h_bdy^2))) using curand, optionally relaxes the samples, and
// culls fluid with inside predicates over an LBVH. Here the jitter comes from
// a counter hash of (seed, sample, axis), the CDF is formed on the host in
// double, and the inside test sums solid angles over all faces (no BVH).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "mesh.hpp"
#include "sph.hpp"

// No fused multiply-adds anywhere in this file (mesh.hpp turns contraction
// off): the samples and the area CDF are formed operation by operation, as a
// restatement in numpy forms them.

namespace reyn {
namespace {

__host__ __device__ inline uint32_t mmix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// U[0, 1) with 24 bits from (seed, sample, axis); the scene generator's hash
__device__ inline double mhash(uint32_t seed, uint32_t idx, uint32_t axis)
{
    const uint32_t h = mmix32(mmix32(seed * 0x9e3779b9u + 0x7f4a7c15u) ^ (idx * 3u + axis));
    return double(h >> 8) * (1.0 / 16777216.0);
}

// Sample i of n: a stratified area coordinate, the face whose inclusive CDF
// entry is the first >= it, and a uniform point of that face (s, t folded into
// the triangle), formed in double and rounded to float.
__global__ __launch_bounds__(kBlock) void k_mesh_sample(MeshDev m, const double* __restrict__ cdf, uint32_t n,
    double total, double bin, uint32_t seed, float* __restrict__ xs, float* __restrict__ ys, float* __restrict__ zs)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const double a = (double(i) / double(n)) * total + bin * mhash(seed, i, 0);
    int lo = 0, hi = int(m.nf) - 1, tri = int(m.nf) - 1;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        if (cdf[mid] >= a) {
            tri = mid;
            hi = mid - 1;
        } else {
            lo = mid + 1;
        }
    }
    const uint32_t i1 = m.f[3 * tri], i2 = m.f[3 * tri + 1], i3 = m.f[3 * tri + 2];
    const double u = mhash(seed, i, 1), v = mhash(seed, i, 2);
    const bool in = u + v <= 1.0;
    const double s = in ? u : 1.0 - u, t = in ? v : 1.0 - v;
    const double ax = m.x[i1], ay = m.y[i1], az = m.z[i1];
    const double e1x = m.x[i2] - ax, e1y = m.y[i2] - ay, e1z = m.z[i2] - az;
    const double e2x = m.x[i3] - ax, e2y = m.y[i3] - ay, e2z = m.z[i3] - az;
    xs[i] = static_cast<float>((ax + s * e1x) + t * e2x);
    ys[i] = static_cast<float>((ay + s * e1y) + t * e2y);
    zs[i] = static_cast<float>((az + s * e1z) + t * e2z);
}

// Generalized winding number of point k: the solid angles of all faces seen
// from it (Van Oosterom and Strackee), over 4 pi; inside when above 1/2.
__global__ __launch_bounds__(kBlock) void k_mesh_inside(MeshDev m, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n, uint8_t* __restrict__ inside)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n)
        return;
    const double px = xs[k], py = ys[k], pz = zs[k];
    double w = 0.0;
    for (uint32_t f = 0; f < m.nf; ++f)
        w += solid_angle2(m, f, px, py, pz);
    inside[k] = w / (4.0 * M_PI) > 0.5 ? 1 : 0;
}

// ---- relaxation of the samples -------------------------------------------------

// Number density of sample i: sum_j W(x_ij) over the samples (B3 of support
// 2 h_bdy; the grid's radius covers it).
__global__ __launch_bounds__(kBlock) void k_relax_density(GridDev g, KParams kp, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n, float* __restrict__ nd)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    float acc = 0.f;
    walk<false>(g, ld3(xs, ys, zs, i), xs, ys, zs, [&](uint32_t, f3, float d2) { acc += kW<KB3>(kp, d2); });
    nd[i] = acc;
}

// The shift of `_shift_neg_rho_grad` (sample_boundary.cu:46-77): down the
// symmetric number-density gradient, -k h_bdy^2 sum_j (n_i + n_j) / n_0^2
// grad W_ij, written to separate arrays (the reference shifts in place while
// neighbours read).
__global__ __launch_bounds__(kBlock) void k_relax_shift(GridDev g, KParams kp, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n, const float* __restrict__ nd, float k,
    float h_bdy, float n0_sq_inv, float* __restrict__ ox, float* __restrict__ oy, float* __restrict__ oz)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const f3 xi = ld3(xs, ys, zs, i);
    const float ti = n0_sq_inv * nd[i];
    f3 acc { 0.f, 0.f, 0.f };
    walk<false>(g, xi, xs, ys, zs, [&](uint32_t j, f3 xij, float d2) {
        acc += (ti + n0_sq_inv * nd[j]) * kdW<KB3>(kp, xij, d2);
    });
    st3(ox, oy, oz, i, xi + (-k * h_bdy * h_bdy) * acc);
}

// Moves every point to the closest point of the mesh (first face on ties).
__global__ __launch_bounds__(kBlock) void k_project(MeshDev m, const float* __restrict__ ix,
    const float* __restrict__ iy, const float* __restrict__ iz, uint32_t n, float* __restrict__ xs,
    float* __restrict__ ys, float* __restrict__ zs)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n)
        return;
    const D3 p { ix[k], iy[k], iz[k] };
    D3 best = p;
    double bd = INFINITY;
    for (uint32_t f = 0; f < m.nf; ++f) {
        double dd;
        int feature;
        const D3 q = closest_on_face(m, f, p, &dd, &feature);
        if (dd < bd) {
            bd = dd;
            best = q;
        }
    }
    xs[k] = static_cast<float>(best.x);
    ys[k] = static_cast<float>(best.y);
    zs[k] = static_cast<float>(best.z);
}

// Rest number density of a hexagonal lattice of spacing h_bdy in a plane,
// points q a + r b for q, r in [-2, 2] (`_hex_plane_rest_density`,
// sample_boundary.cu:80-98).
__global__ void k_hex_rest(KParams kp, float h_bdy, float* out)
{
    float rho = 0.f;
    for (int q = -2; q <= 2; ++q)
        for (int r = -2; r <= 2; ++r) {
            const f3 x { h_bdy * (float(q) + 0.5f * float(r)), h_bdy * (0.8660254037844386f * float(r)), 0.f };
            rho += kW<KB3>(kp, dot(x, x));
        }
    *out = rho;
}

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
    return s;
}

// One OBJ face corner: "v", "v/t", "v//n" or "v/t/n"; negative indices count
// back from the last vertex read. Returns the zero-based vertex index or -1.
long corner_index(const std::string& tok, size_t nverts)
{
    const std::string head = tok.substr(0, tok.find('/'));
    if (head.empty())
        return -1;
    char* end = nullptr;
    const long v = std::strtol(head.c_str(), &end, 10);
    if (!end || *end != '\0' || v == 0)
        return -1;
    return v > 0 ? v - 1 : long(nverts) + v;
}

int load_obj(const char* path, const char* const* ignore, uint32_t n_ignore, const char* only, reyn_mesh* m,
    std::string* err)
{
    std::ifstream in(path);
    if (!in) {
        *err = std::string("cannot open ") + path;
        return REYN_ERR_INVALID_ARGUMENT;
    }
    std::vector<std::string> skip;
    for (uint32_t k = 0; k < n_ignore; ++k)
        if (ignore[k])
            skip.push_back(lower(ignore[k]));
    const std::string want = only ? lower(only) : std::string();
    bool use = true;
    auto pick = [&](const std::string& name) {
        const std::string l = lower(name);
        use = true;
        for (const std::string& s : skip)
            if (l.find(s) != std::string::npos)
                use = false;
        if (only && l.find(want) == std::string::npos)
            use = false;
    };
    pick("");
    std::string line;
    size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#')
            continue;
        if (key == "v") {
            double a = 0, b = 0, c = 0;
            if (!(ls >> a >> b >> c)) {
                *err = "malformed vertex on line " + std::to_string(lineno);
                return REYN_ERR_INVALID_ARGUMENT;
            }
            m->x.push_back(a);
            m->y.push_back(b);
            m->z.push_back(c);
        } else if (key == "o" || key == "g") {
            std::string name;
            std::getline(ls, name);
            pick(name);
        } else if (key == "f") {
            std::vector<std::string> toks;
            std::string t;
            while (ls >> t)
                toks.push_back(t);
            if (toks.size() < 3) {
                *err = "face with fewer than 3 corners on line " + std::to_string(lineno);
                return REYN_ERR_INVALID_ARGUMENT;
            }
            std::vector<long> idx(toks.size());
            for (size_t k = 0; k < toks.size(); ++k) {
                idx[k] = corner_index(toks[k], m->x.size());
                if (idx[k] < 0 || idx[k] >= long(m->x.size())) {
                    *err = "invalid vertex index on line " + std::to_string(lineno);
                    return REYN_ERR_INVALID_ARGUMENT;
                }
            }
            // polygons as fans from their first corner, as the loader the
            // reference pins (tinyobjloader v1.0.6, CMakeLists.txt:145-151)
            // triangulates them by default
            if (use)
                for (size_t k = 2; k < idx.size(); ++k)
                    for (long v : { idx[0], idx[k - 1], idx[k] })
                        m->f.push_back(static_cast<uint32_t>(v));
        }
    }
    return REYN_OK;
}

} // namespace

int upload_mesh(reyn_ctx* ctx, const reyn_mesh* m, DevMesh* out)
{
    const size_t nv = m->x.size(), nf = m->f.size() / 3;
    const size_t bytes = sizeof(double) * 3 * nv + sizeof(uint32_t) * 3 * nf;
    REYN_HIP(ctx, hipMalloc(&out->mem, bytes ? bytes : 8));
    char* p = static_cast<char*>(out->mem);
    double* x = reinterpret_cast<double*>(p);
    uint32_t* f = reinterpret_cast<uint32_t*>(p + sizeof(double) * 3 * nv);
    REYN_HIP(ctx, hipMemcpyAsync(x, m->x.data(), sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream));
    REYN_HIP(ctx, hipMemcpyAsync(x + nv, m->y.data(), sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream));
    REYN_HIP(ctx, hipMemcpyAsync(x + 2 * nv, m->z.data(), sizeof(double) * nv, hipMemcpyHostToDevice, ctx->stream));
    REYN_HIP(ctx, hipMemcpyAsync(f, m->f.data(), sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, ctx->stream));
    out->d = MeshDev { x, x + nv, x + 2 * nv, f, static_cast<uint32_t>(nf) };
    return REYN_OK;
}

namespace {

void face_areas(const reyn_mesh* m, std::vector<double>* cdf)
{
    const size_t nf = m->f.size() / 3;
    cdf->resize(nf);
    double run = 0.0;
    for (size_t k = 0; k < nf; ++k) {
        const uint32_t a = m->f[3 * k], b = m->f[3 * k + 1], c = m->f[3 * k + 2];
        const double e1x = m->x[b] - m->x[a], e1y = m->y[b] - m->y[a], e1z = m->z[b] - m->z[a];
        const double e2x = m->x[c] - m->x[a], e2y = m->y[c] - m->y[a], e2z = m->z[c] - m->z[a];
        const double cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
        run += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        (*cdf)[k] = run;
    }
}

uint32_t sample_count(double total, float h, float oversampling)
{
    const double h_bdy = double(h) / double(oversampling);
    const float cell = 1.5f * std::sqrt(3.f) * static_cast<float>(h_bdy) * static_cast<float>(h_bdy);
    return static_cast<uint32_t>(std::ceil(total / double(cell)));
}

} // namespace
} // namespace reyn

using namespace reyn;

extern "C" {

int reyn_mesh_load_obj(const char* path, const char* const* ignore, uint32_t n_ignore, const char* only,
    reyn_mesh** out, char* err, size_t err_len)
{
    if (!path || !out || (n_ignore && !ignore))
        return REYN_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    reyn_mesh* m = new (std::nothrow) reyn_mesh();
    if (!m)
        return REYN_ERR_OUT_OF_MEMORY;
    std::string e;
    const int rc = load_obj(path, ignore, n_ignore, only, m, &e);
    if (rc != REYN_OK) {
        if (err && err_len)
            std::snprintf(err, err_len, "%s", e.c_str());
        delete m;
        return rc;
    }
    *out = m;
    return REYN_OK;
}

int reyn_mesh_create(const double* x, const double* y, const double* z, uint32_t n_vertices, const uint32_t* faces,
    uint32_t n_faces, reyn_mesh** out)
{
    if (!out || (n_vertices && (!x || !y || !z)) || (n_faces && !faces))
        return REYN_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    for (size_t k = 0; k < size_t(n_faces) * 3; ++k)
        if (faces[k] >= n_vertices)
            return REYN_ERR_OUT_OF_BOUNDS;
    reyn_mesh* m = new (std::nothrow) reyn_mesh();
    if (!m)
        return REYN_ERR_OUT_OF_MEMORY;
    m->x.assign(x, x + n_vertices);
    m->y.assign(y, y + n_vertices);
    m->z.assign(z, z + n_vertices);
    m->f.assign(faces, faces + size_t(n_faces) * 3);
    *out = m;
    return REYN_OK;
}

int reyn_mesh_destroy(reyn_mesh* m)
{
    delete m;
    return REYN_OK;
}

int reyn_mesh_info(const reyn_mesh* m, uint32_t* n_vertices, uint32_t* n_faces, double* area, double* volume)
{
    if (!m)
        return REYN_ERR_INVALID_ARGUMENT;
    const size_t nf = m->f.size() / 3;
    if (n_vertices)
        *n_vertices = static_cast<uint32_t>(m->x.size());
    if (n_faces)
        *n_faces = static_cast<uint32_t>(nf);
    if (area) {
        std::vector<double> cdf;
        face_areas(m, &cdf);
        *area = cdf.empty() ? 0.0 : cdf.back();
    }
    if (volume) { // signed volume: sum of a . (b x c) / 6 over the faces
        double v = 0.0;
        for (size_t k = 0; k < nf; ++k) {
            const uint32_t a = m->f[3 * k], b = m->f[3 * k + 1], c = m->f[3 * k + 2];
            const double cx = m->y[b] * m->z[c] - m->z[b] * m->y[c];
            const double cy = m->z[b] * m->x[c] - m->x[b] * m->z[c];
            const double cz = m->x[b] * m->y[c] - m->y[b] * m->x[c];
            v += m->x[a] * cx + m->y[a] * cy + m->z[a] * cz;
        }
        *volume = v / 6.0;
    }
    return REYN_OK;
}

int reyn_mesh_data(const reyn_mesh* m, double* x, double* y, double* z, uint32_t* faces)
{
    if (!m)
        return REYN_ERR_INVALID_ARGUMENT;
    if (x)
        std::copy(m->x.begin(), m->x.end(), x);
    if (y)
        std::copy(m->y.begin(), m->y.end(), y);
    if (z)
        std::copy(m->z.begin(), m->z.end(), z);
    if (faces)
        std::copy(m->f.begin(), m->f.end(), faces);
    return REYN_OK;
}

int reyn_mesh_sample_count(const reyn_mesh* m, float h, float oversampling, uint32_t* count)
{
    if (!m || !count || !(h > 0.f) || !(oversampling > 0.f))
        return REYN_ERR_INVALID_ARGUMENT;
    std::vector<double> cdf;
    face_areas(m, &cdf);
    *count = cdf.empty() ? 0u : sample_count(cdf.back(), h, oversampling);
    return REYN_OK;
}

int reyn_mesh_sample(reyn_ctx* ctx, const reyn_mesh* m, float h, float oversampling, uint32_t seed, float* xs,
    float* ys, float* zs, uint32_t capacity, uint32_t* count)
{
    if (!ctx || !m || !count || !(h > 0.f) || !(oversampling > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "mesh sample: bad argument");
    std::vector<double> cdf;
    face_areas(m, &cdf);
    if (cdf.empty() || !(cdf.back() > 0.0))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "mesh sample: the mesh has no area");
    const uint32_t n = sample_count(cdf.back(), h, oversampling);
    *count = n;
    if (n > capacity || !xs || !ys || !zs)
        return set_error(ctx, REYN_ERR_CAPACITY, "mesh sample: %u samples exceed the capacity %u", n, capacity);
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    DevMesh dm;
    double* dcdf = nullptr;
    int rc = upload_mesh(ctx, m, &dm);
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&dcdf), sizeof(double) * cdf.size()), "hipMalloc cdf");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMemcpyAsync(dcdf, cdf.data(), sizeof(double) * cdf.size(), hipMemcpyHostToDevice,
                                ctx->stream), "copy cdf");
    if (rc == REYN_OK) {
        k_mesh_sample<<<blocks_for(n), kBlock, 0, ctx->stream>>>(dm.d, dcdf, n, cdf.back(),
            cdf.back() / double(n), seed, xs, ys, zs);
        rc = check_hip(ctx, hipGetLastError(), "k_mesh_sample");
    }
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(ctx->stream);
    if (dcdf)
        (void)hipFree(dcdf);
    if (dm.mem)
        (void)hipFree(dm.mem);
    return rc;
}

int reyn_mesh_relax(reyn_ctx* ctx, const reyn_mesh* m, float h, float oversampling, float* xs, float* ys,
    float* zs, uint32_t n, uint32_t iterations, float factor, float* ratio)
{
    if (!ctx || !m || (n && (!xs || !ys || !zs)) || !(h > 0.f) || !(oversampling > 0.f))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "mesh relax: bad argument");
    if (n == 0 || iterations == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    // the grid reaches 2 h_grid >= the kernel support 2 h_bdy (the reference
    // takes h_grid = h_bdy * oversampling = h, too short for oversampling < 1)
    const float h_bdy = static_cast<float>(double(h) / double(oversampling)), h_grid = fmaxf(h, h_bdy);
    const KParams kp = make_kparams(REYN_KERNEL_B3, 2.f * h_bdy);
    // bounds of the samples, padded by 5 h_grid as in the reference
    reyn_float3 lo {}, hi {};
    float* lv[3] = { &lo.x, &lo.y, &lo.z };
    float* hv[3] = { &hi.x, &hi.y, &hi.z };
    const float* arr[3] = { xs, ys, zs };
    for (int a = 0; a < 3; ++a) {
        double mn = 0.0, mx = 0.0;
        REYN_TRY(reyn_reduce_f32(ctx, arr[a], n, 1, &mn));
        REYN_TRY(reyn_reduce_f32(ctx, arr[a], n, 2, &mx));
        *lv[a] = static_cast<float>(mn) - 5.f * h_grid;
        *hv[a] = static_cast<float>(mx) + 5.f * h_grid;
    }
    reyn_lbvh* bvh = nullptr;
    REYN_TRY(reyn_lbvh_create(ctx, m, &bvh));
    reyn_grid_builder* gb = nullptr;
    int rc = reyn_grid_builder_create(ctx, lo, hi, 2.f * h_grid, 0.5f, &gb);
    float* work = nullptr; // tmp x y z, shifted x y z, number density, rest density
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&work), sizeof(float) * (7 * size_t(n) + 1)),
            "hipMalloc relax");
    float* tmp[3] = { work, work + n, work + 2 * size_t(n) };
    float* sh[3] = { work + 3 * size_t(n), work + 4 * size_t(n), work + 5 * size_t(n) };
    float* nd = work + 6 * size_t(n);
    float* rest = work + 7 * size_t(n);
    float n0 = 0.f;
    if (rc == REYN_OK) {
        k_hex_rest<<<1, 1, 0, ctx->stream>>>(kp, h_bdy, rest);
        rc = check_hip(ctx, hipGetLastError(), "k_hex_rest");
    }
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMemcpyAsync(&n0, rest, sizeof(float), hipMemcpyDeviceToHost, ctx->stream), "copy");
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    const float n0_sq_inv = 1.f / (n0 * n0);
    // the reorder swaps `cur` with `tmp` and the projection `cur` with `sh`;
    // the last positions are copied back into the caller's arrays when they
    // ended up in the work buffers
    float* cur[3] = { xs, ys, zs };
    for (uint32_t it = 0; it < iterations && rc == REYN_OK; ++it) {
        reyn_grid g {};
        rc = reyn_grid_construct_and_reorder(ctx, gb, 2.f * h_grid, cur, tmp, 3, n, &g);
        if (rc != REYN_OK)
            break;
        const GridDev gd = to_dev(g);
        k_relax_density<<<blocks_for(n), kBlock, 0, ctx->stream>>>(gd, kp, cur[0], cur[1], cur[2], n, nd);
        k_relax_shift<<<blocks_for(n), kBlock, 0, ctx->stream>>>(gd, kp, cur[0], cur[1], cur[2], n, nd, factor,
            h_bdy, n0_sq_inv, sh[0], sh[1], sh[2]);
        rc = check_hip(ctx, hipGetLastError(), "relax kernels");
        if (rc == REYN_OK)
            rc = reyn_lbvh_project(ctx, bvh, sh[0], sh[1], sh[2], n);
        for (int a = 0; a < 3; ++a)
            std::swap(cur[a], sh[a]);
    }
    float* const mine[3] = { xs, ys, zs };
    for (int a = 0; a < 3 && rc == REYN_OK; ++a)
        if (cur[a] != mine[a])
            rc = check_hip(ctx,
                hipMemcpyAsync(mine[a], cur[a], sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream), "copy");
    if (rc == REYN_OK && ratio) { // spread of the number density before the last shift
        double mn = 0.0, mx = 0.0;
        rc = reyn_reduce_f32(ctx, nd, n, 1, &mn);
        if (rc == REYN_OK)
            rc = reyn_reduce_f32(ctx, nd, n, 2, &mx);
        *ratio = static_cast<float>(mx / mn);
    }
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    (void)hipStreamSynchronize(ctx->stream);
    if (work)
        (void)hipFree(work);
    reyn_lbvh_destroy(bvh);
    reyn_grid_builder_destroy(gb);
    return rc;
}

int reyn_mesh_project(reyn_ctx* ctx, const reyn_mesh* m, float* xs, float* ys, float* zs, uint32_t n)
{
    if (!ctx || !m || (n && (!xs || !ys || !zs)))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "mesh project: NULL argument");
    if (n == 0)
        return REYN_OK;
    if (m->f.empty())
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "mesh project: the mesh has no faces");
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    DevMesh dm;
    float* in = nullptr;
    int rc = upload_mesh(ctx, m, &dm);
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipMalloc(reinterpret_cast<void**>(&in), sizeof(float) * 3 * size_t(n)), "hipMalloc");
    float* const io[3] = { xs, ys, zs };
    for (int a = 0; a < 3 && rc == REYN_OK; ++a)
        rc = check_hip(ctx, hipMemcpyAsync(in + a * size_t(n), io[a], sizeof(float) * n, hipMemcpyDeviceToDevice,
                                ctx->stream), "copy");
    if (rc == REYN_OK) {
        k_project<<<blocks_for(n), kBlock, 0, ctx->stream>>>(dm.d, in, in + n, in + 2 * size_t(n), n, xs, ys, zs);
        rc = check_hip(ctx, hipGetLastError(), "k_project");
    }
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(ctx->stream);
    if (in)
        (void)hipFree(in);
    if (dm.mem)
        (void)hipFree(dm.mem);
    return rc;
}

int reyn_mesh_inside(reyn_ctx* ctx, const reyn_mesh* m, const float* xs, const float* ys, const float* zs,
    uint32_t n, uint8_t* inside)
{
    if (!ctx || !m || (n && (!xs || !ys || !zs || !inside)))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "mesh inside: NULL argument");
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    DevMesh dm;
    int rc = upload_mesh(ctx, m, &dm);
    if (rc == REYN_OK) {
        k_mesh_inside<<<blocks_for(n), kBlock, 0, ctx->stream>>>(dm.d, xs, ys, zs, n, inside);
        rc = check_hip(ctx, hipGetLastError(), "k_mesh_inside");
    }
    if (rc == REYN_OK)
        rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(ctx->stream);
    if (dm.mem)
        (void)hipFree(dm.mem);
    return rc;
}

} // extern "C"
