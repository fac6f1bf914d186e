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

struct reyn_mesh {
    std::vector<double> x, y, z;
    std::vector<uint32_t> f; // 3 per face
};

// No fused multiply-adds anywhere in this file: the samples and the area CDF
// are formed operation by operation, as a restatement in numpy forms them.
#pragma clang fp contract(off)

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

struct MeshDev {
    const double *x, *y, *z;
    const uint32_t* f;
    uint32_t nf;
};

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
    for (uint32_t f = 0; f < m.nf; ++f) {
        const uint32_t i1 = m.f[3 * f], i2 = m.f[3 * f + 1], i3 = m.f[3 * f + 2];
        const double ax = m.x[i1] - px, ay = m.y[i1] - py, az = m.z[i1] - pz;
        const double bx = m.x[i2] - px, by = m.y[i2] - py, bz = m.z[i2] - pz;
        const double cx = m.x[i3] - px, cy = m.y[i3] - py, cz = m.z[i3] - pz;
        const double la = sqrt(ax * ax + ay * ay + az * az);
        const double lb = sqrt(bx * bx + by * by + bz * bz);
        const double lc = sqrt(cx * cx + cy * cy + cz * cz);
        const double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
        const double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb
            + (bx * cx + by * cy + bz * cz) * la;
        w += 2.0 * atan2(det, den);
    }
    inside[k] = w / (4.0 * M_PI) > 0.5 ? 1 : 0;
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

struct DevMesh {
    void* mem = nullptr;
    MeshDev d {};
};

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
