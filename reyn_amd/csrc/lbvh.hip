// Linear BVH over the faces of a triangle mesh, built on the GPU: Morton codes
// of the face centroids made unique by the face index, a radix sort, the
// binary radix tree of the sorted codes with every internal node built in
// parallel, and the boxes bottom-up. Queries find the closest point of the
// mesh by a depth-first descent that visits the nearer child first and prunes
// boxes farther than the best distance so far.
//
// The reference builds the same kind of tree (`LBVH`, src/datastructure/
// lbvh.cuh:614-758, after Karras, "Maximizing Parallelism in the Construction
// of BVHs, Octrees, and k-d Trees", 2012) for `LBVH::project` and the
// `InsidePredicate` (src/scene/scene.cu:88-140). Differences: the boxes are
// exact double min/max of the vertices; ties between faces at the same
// distance go to the lowest face index, so a query returns exactly what a
// brute-force scan over the faces in order returns; and the inside test reads
// the angle-weighted pseudo-normal of the closest feature (face, edge or
// vertex; Baerentzen and Aanaes 2005) instead of the closest face's normal,
// which gives the wrong side where the closest point lies on an edge or vertex.
#include <algorithm>
#include <cmath>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "mesh.hpp"

namespace reyn {
namespace {

constexpr int kStack = 64; // >= the tree height: 64-bit keys give at most 64 levels
constexpr uint32_t kInternal = 1u; // child encoding: index << 1 | 1 for an internal node

struct BvhDev {
    MeshDev m;
    const uint32_t *left, *right, *leaf_face;
    const double *ibox, *lbox; // 6 per node: lo x y z, hi x y z
    const double *vn, *en; // pseudo-normals: 3 per vertex, 9 per face (edges AB, AC, BC)
    uint32_t nf;
    double slack; // distance margin for pruning (rounding of the face distances)
};

// The 10 bits of x spread to every third bit of 30.
__device__ inline uint32_t spread10(uint32_t x)
{
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < 10; ++b)
        r |= ((x >> b) & 1u) << (3 * b);
    return r;
}

__device__ inline uint32_t quantize10(double u)
{
    const double s = fmin(fmax(u * 1024.0, 0.0), 1023.0);
    return static_cast<uint32_t>(s);
}

// key = Morton code of the centroid in the mesh's box (x highest) << 32 | face
__global__ __launch_bounds__(kBlock) void k_bvh_keys(MeshDev m, D3 lo, D3 inv, uint64_t* __restrict__ keys)
{
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= m.nf)
        return;
    const D3 a = vertex(m, m.f[3 * f]), b = vertex(m, m.f[3 * f + 1]), c = vertex(m, m.f[3 * f + 2]);
    const D3 cen = dscale(1.0 / 3.0, dadd(dadd(a, b), c));
    const uint32_t qx = quantize10((cen.x - lo.x) * inv.x);
    const uint32_t qy = quantize10((cen.y - lo.y) * inv.y);
    const uint32_t qz = quantize10((cen.z - lo.z) * inv.z);
    const uint32_t code = (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
    keys[f] = (uint64_t(code) << 32) | f;
}

__device__ inline int common_prefix(const uint64_t* __restrict__ k, int n, int a, int b)
{
    if (b < 0 || b >= n)
        return -1;
    return __clzll(static_cast<long long>(k[a] ^ k[b]));
}

// Internal node i of the radix tree over n unique sorted keys: the direction
// of its range, the range's other end by exponential and binary search, and
// the split where the highest differing bit changes.
__global__ __launch_bounds__(kBlock) void k_bvh_tree(const uint64_t* __restrict__ keys, int n, uint32_t* __restrict__ left,
    uint32_t* __restrict__ right, uint32_t* __restrict__ iparent, uint32_t* __restrict__ lparent)
{
    const int i = static_cast<int>(blockIdx.x * kBlock + threadIdx.x);
    if (i >= n - 1)
        return;
    const int d = common_prefix(keys, n, i, i + 1) >= common_prefix(keys, n, i, i - 1) ? 1 : -1;
    const int dmin = common_prefix(keys, n, i, i - d);
    int lmax = 2;
    while (common_prefix(keys, n, i, i + lmax * d) > dmin)
        lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (common_prefix(keys, n, i, i + (l + t) * d) > dmin)
            l += t;
    const int j = i + l * d;
    const int dnode = common_prefix(keys, n, i, j);
    int s = 0;
    for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (common_prefix(keys, n, i, i + (s + t) * d) > dnode)
            s += t;
        if (t == 1)
            break;
    }
    const int split = i + s * d + (d < 0 ? -1 : 0);
    const int first = i < j ? i : j, last = i < j ? j : i;
    const bool lleaf = first == split, rleaf = last == split + 1;
    left[i] = (uint32_t(split) << 1) | (lleaf ? 0u : kInternal);
    right[i] = (uint32_t(split + 1) << 1) | (rleaf ? 0u : kInternal);
    (lleaf ? lparent : iparent)[split] = uint32_t(i);
    (rleaf ? lparent : iparent)[split + 1] = uint32_t(i);
}

__device__ inline void face_box(const MeshDev& m, uint32_t f, double* box)
{
    const D3 a = vertex(m, m.f[3 * f]), b = vertex(m, m.f[3 * f + 1]), c = vertex(m, m.f[3 * f + 2]);
    box[0] = fmin(fmin(a.x, b.x), c.x);
    box[1] = fmin(fmin(a.y, b.y), c.y);
    box[2] = fmin(fmin(a.z, b.z), c.z);
    box[3] = fmax(fmax(a.x, b.x), c.x);
    box[4] = fmax(fmax(a.y, b.y), c.y);
    box[5] = fmax(fmax(a.z, b.z), c.z);
}

// Leaf k's face and box, then upwards: the second of two siblings to arrive
// at a parent (device-scope counter) forms the parent's box from both.
__global__ __launch_bounds__(kBlock) void k_bvh_boxes(MeshDev m, const uint64_t* __restrict__ keys, uint32_t n,
    const uint32_t* __restrict__ left, const uint32_t* __restrict__ right, const uint32_t* __restrict__ iparent,
    const uint32_t* __restrict__ lparent, uint32_t* __restrict__ arrived, uint32_t* __restrict__ leaf_face,
    double* lbox, double* ibox, uint32_t* height)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n)
        return;
    const uint32_t f = static_cast<uint32_t>(keys[k] & 0xffffffffu);
    leaf_face[k] = f;
    double box[6];
    face_box(m, f, box);
    for (int a = 0; a < 6; ++a)
        lbox[6 * size_t(k) + a] = box[a];
    if (n == 1)
        return;
    // the leaf's depth (levels from the root down to it, the leaf included);
    // the parents are complete since the tree launch
    uint32_t depth = 2;
    for (uint32_t up = lparent[k]; up != 0 && depth <= n; up = iparent[up])
        ++depth;
    atomicMax(height, depth);
    uint32_t node = lparent[k];
    for (;;) {
        __threadfence();
        if (atomicAdd(arrived + node, 1u) == 0u)
            return; // the sibling finishes this node
        __threadfence();
        double u[6];
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = c == 0 ? left[node] : right[node];
            const double* src = ((ch & kInternal) ? ibox : lbox) + 6 * size_t(ch >> 1);
            for (int a = 0; a < 6; ++a) {
                const double v = __hip_atomic_load(src + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                u[a] = c == 0 ? v : (a < 3 ? fmin(u[a], v) : fmax(u[a], v));
            }
        }
        for (int a = 0; a < 6; ++a)
            ibox[6 * size_t(node) + a] = u[a];
        if (node == 0)
            return;
        node = iparent[node];
    }
}

// Squared distance from p to the box (0 inside).
__device__ inline double box_d2(const double* b, D3 p)
{
    const double dx = fmax(fmax(b[0] - p.x, 0.0), p.x - b[3]);
    const double dy = fmax(fmax(b[1] - p.y, 0.0), p.y - b[4]);
    const double dz = fmax(fmax(b[2] - p.z, 0.0), p.z - b[5]);
    return dx * dx + dy * dy + dz * dz;
}

struct Hit {
    D3 q;
    double d2;
    uint32_t face;
    int feature;
};

__device__ inline void test_leaf(const BvhDev& b, uint32_t leaf, D3 p, Hit* h, double* lim)
{
    const uint32_t f = b.leaf_face[leaf];
    double d2;
    int feature;
    const D3 q = closest_on_face(b.m, f, p, &d2, &feature);
    if (d2 < h->d2 || (d2 == h->d2 && f < h->face)) {
        *h = Hit { q, d2, f, feature };
        const double r = sqrt(d2) + b.slack;
        *lim = r * r;
    }
}

// Closest point of the mesh to p; among faces at the same distance the
// lowest face index.
__device__ Hit bvh_closest(const BvhDev& b, D3 p)
{
    Hit h { p, INFINITY, 0xffffffffu, kFace };
    double lim = INFINITY; // boxes beyond this squared distance cannot hold a closer face
    if (b.nf == 1) {
        test_leaf(b, 0, p, &h, &lim);
        return h;
    }
    uint32_t stack[kStack];
    double sd[kStack];
    int sp = 0;
    uint32_t node = 0; // the root
    for (;;) {
        const uint32_t ch[2] = { b.left[node], b.right[node] };
        double cd[2];
        for (int c = 0; c < 2; ++c)
            cd[c] = box_d2(((ch[c] & kInternal) ? b.ibox : b.lbox) + 6 * size_t(ch[c] >> 1), p);
        // leaves are tested at once; internal children are descended nearer first
        uint32_t next[2];
        double nd[2];
        int nn = 0;
        for (int c = 0; c < 2; ++c) {
            if (cd[c] > lim)
                continue;
            if (ch[c] & kInternal) {
                next[nn] = ch[c] >> 1;
                nd[nn] = cd[c];
                ++nn;
            } else {
                test_leaf(b, ch[c] >> 1, p, &h, &lim);
            }
        }
        if (nn == 2 && nd[1] < nd[0]) {
            std::swap(next[0], next[1]);
            std::swap(nd[0], nd[1]);
        }
        if (nn == 2) {
            stack[sp] = next[1];
            sd[sp] = nd[1];
            ++sp;
        }
        if (nn >= 1 && nd[0] <= lim) {
            node = next[0];
            continue;
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            if (sd[sp] <= lim) {
                node = stack[sp];
                found = true;
                break;
            }
        }
        if (!found)
            break;
    }
    return h;
}

__global__ __launch_bounds__(kBlock) void k_bvh_project(BvhDev b, float* __restrict__ xs, float* __restrict__ ys,
    float* __restrict__ zs, uint32_t n)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n)
        return;
    const Hit h = bvh_closest(b, D3 { xs[k], ys[k], zs[k] });
    xs[k] = static_cast<float>(h.q.x);
    ys[k] = static_cast<float>(h.q.y);
    zs[k] = static_cast<float>(h.q.z);
}

// Inside when p lies behind the pseudo-normal of the closest feature; points
// on the surface count as outside.
__global__ __launch_bounds__(kBlock) void k_bvh_inside(BvhDev b, const float* __restrict__ xs,
    const float* __restrict__ ys, const float* __restrict__ zs, uint32_t n, uint8_t* __restrict__ inside)
{
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n)
        return;
    const D3 p { xs[k], ys[k], zs[k] };
    const Hit h = bvh_closest(b, p);
    const uint32_t* fv = b.m.f + 3 * size_t(h.face);
    D3 nrm;
    switch (h.feature) {
    case kVertA:
    case kVertB:
    case kVertC: {
        const uint32_t v = fv[h.feature - kVertA];
        nrm = D3 { b.vn[3 * size_t(v)], b.vn[3 * size_t(v) + 1], b.vn[3 * size_t(v) + 2] };
        break;
    }
    case kEdgeAB:
    case kEdgeAC:
    case kEdgeBC: {
        const double* e = b.en + 9 * size_t(h.face) + 3 * (h.feature - kEdgeAB);
        nrm = D3 { e[0], e[1], e[2] };
        break;
    }
    default:
        nrm = dcross(dsub(vertex(b.m, fv[1]), vertex(b.m, fv[0])), dsub(vertex(b.m, fv[2]), vertex(b.m, fv[0])));
    }
    inside[k] = (h.d2 > 0.0 && ddot(dsub(p, h.q), nrm) < 0.0) ? 1 : 0;
}

// Angle-weighted vertex normals and per-face edge normals (the sum of the
// unit normals of the faces sharing the edge), in double on the host.
void pseudo_normals(const reyn_mesh* m, std::vector<double>* vn, std::vector<double>* en)
{
    const size_t nv = m->x.size(), nf = m->f.size() / 3;
    vn->assign(3 * nv, 0.0);
    en->assign(9 * nf, 0.0);
    std::map<std::pair<uint32_t, uint32_t>, D3> edges;
    std::vector<D3> unit(nf);
    auto vtx = [&](uint32_t v) { return D3 { m->x[v], m->y[v], m->z[v] }; };
    for (size_t f = 0; f < nf; ++f) {
        const uint32_t id[3] = { m->f[3 * f], m->f[3 * f + 1], m->f[3 * f + 2] };
        const D3 n = dcross(dsub(vtx(id[1]), vtx(id[0])), dsub(vtx(id[2]), vtx(id[0])));
        const double len = std::sqrt(ddot(n, n));
        unit[f] = len > 0.0 ? dscale(1.0 / len, n) : D3 { 0.0, 0.0, 0.0 };
        for (int c = 0; c < 3; ++c) {
            const D3 u = dsub(vtx(id[(c + 1) % 3]), vtx(id[c])), w = dsub(vtx(id[(c + 2) % 3]), vtx(id[c]));
            const double lu = std::sqrt(ddot(u, u)), lw = std::sqrt(ddot(w, w));
            if (lu == 0.0 || lw == 0.0)
                continue;
            const double ang = std::acos(std::clamp(ddot(u, w) / (lu * lw), -1.0, 1.0));
            for (int a = 0; a < 3; ++a)
                (*vn)[3 * size_t(id[c]) + a] += ang * (a == 0 ? unit[f].x : a == 1 ? unit[f].y : unit[f].z);
        }
        for (int e = 0; e < 3; ++e) {
            const uint32_t u = id[e], v = id[(e + 1) % 3];
            D3& s = edges[{ std::min(u, v), std::max(u, v) }];
            s = dadd(s, unit[f]);
        }
    }
    // the face's edges in the order of the closest-point features: AB, AC, BC
    const int ends[3][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 } };
    for (size_t f = 0; f < nf; ++f)
        for (int e = 0; e < 3; ++e) {
            const uint32_t u = m->f[3 * f + ends[e][0]], v = m->f[3 * f + ends[e][1]];
            const D3 s = edges[{ std::min(u, v), std::max(u, v) }];
            (*en)[9 * f + 3 * e] = s.x;
            (*en)[9 * f + 3 * e + 1] = s.y;
            (*en)[9 * f + 3 * e + 2] = s.z;
        }
}

} // namespace
} // namespace reyn

using namespace reyn;

struct reyn_lbvh {
    int device = 0;
    uint32_t nf = 0, height = 0;
    DevMesh mesh;
    void* mem = nullptr;
    BvhDev d {};
};

namespace {

void free_lbvh(reyn_lbvh* b)
{
    if (!b)
        return;
    if (b->mem)
        (void)hipFree(b->mem);
    if (b->mesh.mem)
        (void)hipFree(b->mesh.mem);
    delete b;
}

int check_bvh_ctx(reyn_ctx* ctx, const reyn_lbvh* b)
{
    if (!ctx || !b)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lbvh: NULL argument");
    if (ctx->device != b->device)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lbvh: built on another device than the context's");
    return REYN_OK;
}

int build_lbvh(reyn_ctx* ctx, const reyn_mesh* m, reyn_lbvh* b)
{
    const uint32_t nf = static_cast<uint32_t>(m->f.size() / 3), nv = static_cast<uint32_t>(m->x.size());
    b->device = ctx->device;
    b->nf = nf;
    REYN_TRY(upload_mesh(ctx, m, &b->mesh));
    // bounds of the vertices, the scale of the pruning margin
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    const std::vector<double>* vs[3] = { &m->x, &m->y, &m->z };
    for (int a = 0; a < 3; ++a)
        for (double v : *vs[a]) {
            lo[a] = std::min(lo[a], v);
            hi[a] = std::max(hi[a], v);
        }
    double diag = 0.0;
    D3 inv {};
    double* iv[3] = { &inv.x, &inv.y, &inv.z };
    for (int a = 0; a < 3; ++a) {
        const double ext = hi[a] - lo[a];
        *iv[a] = ext > 0.0 ? 1.0 / ext : 0.0;
        diag += ext * ext;
    }
    std::vector<double> vn, en;
    pseudo_normals(m, &vn, &en);
    // one allocation: nodes, boxes, normals, then the build's scratch
    const size_t ni = nf > 1 ? nf - 1 : 0;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) / 256 * 256;
        return at;
    };
    const size_t o_left = take(4 * ni), o_right = take(4 * ni), o_leaf = take(4 * size_t(nf));
    const size_t o_ibox = take(48 * ni), o_lbox = take(48 * size_t(nf));
    const size_t o_vn = take(8 * vn.size()), o_en = take(8 * en.size());
    const size_t keep = off;
    const size_t o_keys = take(8 * size_t(nf)), o_sorted = take(8 * size_t(nf));
    const size_t o_ipar = take(4 * ni), o_lpar = take(4 * size_t(nf)), o_arr = take(4 * ni), o_h = take(4);
    size_t sort_bytes = 0;
    REYN_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr),
                      static_cast<uint64_t*>(nullptr), int(nf), 0, 64, ctx->stream));
    const size_t o_sort = take(sort_bytes);
    char* base = nullptr;
    REYN_HIP(ctx, hipMalloc(reinterpret_cast<void**>(&base), off));
    b->mem = base;
    auto u32 = [&](size_t o) { return reinterpret_cast<uint32_t*>(base + o); };
    auto f64 = [&](size_t o) { return reinterpret_cast<double*>(base + o); };
    auto u64 = [&](size_t o) { return reinterpret_cast<uint64_t*>(base + o); };
    REYN_HIP(ctx, hipMemcpyAsync(f64(o_vn), vn.data(), 8 * vn.size(), hipMemcpyHostToDevice, ctx->stream));
    REYN_HIP(ctx, hipMemcpyAsync(f64(o_en), en.data(), 8 * en.size(), hipMemcpyHostToDevice, ctx->stream));
    REYN_HIP(ctx, hipMemsetAsync(base + keep, 0, off - keep, ctx->stream));
    const MeshDev md = b->mesh.d;
    k_bvh_keys<<<blocks_for(nf), kBlock, 0, ctx->stream>>>(md, D3 { lo[0], lo[1], lo[2] }, inv, u64(o_keys));
    REYN_HIP(ctx, hipGetLastError());
    REYN_HIP(ctx, hipcub::DeviceRadixSort::SortKeys(base + o_sort, sort_bytes, u64(o_keys), u64(o_sorted), int(nf),
                      0, 64, ctx->stream));
    if (nf > 1) {
        k_bvh_tree<<<blocks_for(uint32_t(ni)), kBlock, 0, ctx->stream>>>(u64(o_sorted), int(nf), u32(o_left),
            u32(o_right), u32(o_ipar), u32(o_lpar));
        REYN_HIP(ctx, hipGetLastError());
    }
    k_bvh_boxes<<<blocks_for(nf), kBlock, 0, ctx->stream>>>(md, u64(o_sorted), nf, u32(o_left), u32(o_right),
        u32(o_ipar), u32(o_lpar), u32(o_arr), u32(o_leaf), f64(o_lbox), f64(o_ibox), u32(o_h));
    REYN_HIP(ctx, hipGetLastError());
    uint32_t height = 0;
    REYN_HIP(ctx, hipMemcpyAsync(&height, u32(o_h), 4, hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->height = nf > 1 ? height : 1;
    if (b->height > uint32_t(kStack))
        return set_error(ctx, REYN_ERR_CAPACITY, "lbvh: tree deeper than the traversal stack");
    b->d = BvhDev { md, u32(o_left), u32(o_right), u32(o_leaf), f64(o_ibox), f64(o_lbox), f64(o_vn), f64(o_en), nf,
        1e-12 * std::sqrt(diag) };
    return REYN_OK;
}

} // namespace

extern "C" {

int reyn_lbvh_create(reyn_ctx* ctx, const reyn_mesh* m, reyn_lbvh** out)
{
    if (!ctx || !m || !out)
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lbvh create: NULL argument");
    *out = nullptr;
    if (m->f.empty())
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lbvh create: the mesh has no faces");
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    reyn_lbvh* b = new (std::nothrow) reyn_lbvh();
    if (!b)
        return set_error(ctx, REYN_ERR_OUT_OF_MEMORY, "lbvh create: host allocation");
    const int rc = build_lbvh(ctx, m, b);
    if (rc != REYN_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        free_lbvh(b);
        return rc;
    }
    *out = b;
    return REYN_OK;
}

int reyn_lbvh_destroy(reyn_lbvh* b)
{
    if (b) {
        (void)hipSetDevice(b->device);
        (void)hipDeviceSynchronize();
    }
    free_lbvh(b);
    return REYN_OK;
}

int reyn_lbvh_info(const reyn_lbvh* b, uint32_t* n_faces, uint32_t* height)
{
    if (!b)
        return REYN_ERR_INVALID_ARGUMENT;
    if (n_faces)
        *n_faces = b->nf;
    if (height)
        *height = b->height;
    return REYN_OK;
}

int reyn_lbvh_nodes(reyn_ctx* ctx, const reyn_lbvh* b, uint32_t* left, uint32_t* right, uint32_t* leaf_face,
    double* internal_box, double* leaf_box)
{
    REYN_TRY(check_bvh_ctx(ctx, b));
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    const size_t ni = b->nf - 1, nf = b->nf;
    struct Item {
        void* dst;
        const void* src;
        size_t bytes;
    } items[5] = { { left, b->d.left, 4 * ni }, { right, b->d.right, 4 * ni }, { leaf_face, b->d.leaf_face, 4 * nf },
        { internal_box, b->d.ibox, 48 * ni }, { leaf_box, b->d.lbox, 48 * nf } };
    for (const Item& it : items)
        if (it.dst && it.bytes)
            REYN_HIP(ctx, hipMemcpyAsync(it.dst, it.src, it.bytes, hipMemcpyDeviceToHost, ctx->stream));
    REYN_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return REYN_OK;
}

int reyn_lbvh_project(reyn_ctx* ctx, const reyn_lbvh* b, float* xs, float* ys, float* zs, uint32_t n)
{
    REYN_TRY(check_bvh_ctx(ctx, b));
    if (n && (!xs || !ys || !zs))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lbvh project: NULL array");
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    k_bvh_project<<<blocks_for(n), kBlock, 0, ctx->stream>>>(b->d, xs, ys, zs, n);
    REYN_HIP(ctx, hipGetLastError());
    return REYN_OK;
}

int reyn_lbvh_inside(reyn_ctx* ctx, const reyn_lbvh* b, const float* xs, const float* ys, const float* zs,
    uint32_t n, uint8_t* inside)
{
    REYN_TRY(check_bvh_ctx(ctx, b));
    if (n && (!xs || !ys || !zs || !inside))
        return set_error(ctx, REYN_ERR_INVALID_ARGUMENT, "lbvh inside: NULL array");
    if (n == 0)
        return REYN_OK;
    REYN_HIP(ctx, hipSetDevice(ctx->device));
    k_bvh_inside<<<blocks_for(n), kBlock, 0, ctx->stream>>>(b->d, xs, ys, zs, n, inside);
    REYN_HIP(ctx, hipGetLastError());
    return REYN_OK;
}

} // extern "C"
