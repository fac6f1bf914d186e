// Device-side triangle-mesh pieces shared by the sampling, relaxation and
// inside-test code (mesh.hip) and the LBVH (lbvh.hip).
//
// Everything here is formed operation by operation without fused
// multiply-adds (the pragma below stays in force for the including file), so
// the brute-force and the LBVH closest-point searches, compiled in different
// files, produce the same bits for the same face.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

#pragma clang fp contract(off)

struct reyn_mesh {
    std::vector<double> x, y, z;
    std::vector<uint32_t> f; // 3 per face
};

namespace reyn {

struct MeshDev {
    const double *x, *y, *z;
    const uint32_t* f;
    uint32_t nf;
};

// One device allocation holding x, y, z (double) and the faces.
struct DevMesh {
    void* mem = nullptr;
    MeshDev d {};
};
int upload_mesh(reyn_ctx* ctx, const reyn_mesh* m, DevMesh* out);

struct D3 {
    double x, y, z;
};
__host__ __device__ inline D3 dsub(D3 a, D3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__host__ __device__ inline D3 dadd(D3 a, D3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__host__ __device__ inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline D3 dscale(double s, D3 a) { return { s * a.x, s * a.y, s * a.z }; }
__host__ __device__ inline D3 dmad(D3 a, double s, D3 b) { return { a.x + s * b.x, a.y + s * b.y, a.z + s * b.z }; }
__host__ __device__ inline D3 dcross(D3 a, D3 b)
{
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}

__device__ inline D3 vertex(const MeshDev& m, uint32_t v) { return D3 { m.x[v], m.y[v], m.z[v] }; }

// Feature of a triangle that holds the closest point: its interior, one of
// its vertices or one of its edges.
enum TriFeature : int { kFace = 0, kVertA, kVertB, kVertC, kEdgeAB, kEdgeAC, kEdgeBC };

// Closest point of triangle (a, b, c) to p, by the Voronoi regions of its
// vertices, edges and face; `feature` receives the region.
__device__ inline D3 closest_on_triangle(D3 p, D3 a, D3 b, D3 c, int* feature)
{
    const D3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
    const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        *feature = kVertA;
        return a;
    }
    const D3 bp = dsub(p, b);
    const double d3 = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        *feature = kVertB;
        return b;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        *feature = kEdgeAB;
        return dmad(a, d1 / (d1 - d3), ab);
    }
    const D3 cp = dsub(p, c);
    const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        *feature = kVertC;
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        *feature = kEdgeAC;
        return dmad(a, d2 / (d2 - d6), ac);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        *feature = kEdgeBC;
        return dmad(b, (d4 - d3) / ((d4 - d3) + (d5 - d6)), dsub(c, b));
    }
    *feature = kFace;
    const double den = 1.0 / (va + vb + vc);
    return dmad(dmad(a, vb * den, ab), vc * den, ac);
}

// Closest point of face f; `d2` its squared distance to p.
__device__ inline D3 closest_on_face(const MeshDev& m, uint32_t f, D3 p, double* d2, int* feature)
{
    const D3 q = closest_on_triangle(p, vertex(m, m.f[3 * f]), vertex(m, m.f[3 * f + 1]), vertex(m, m.f[3 * f + 2]),
        feature);
    const D3 d = dsub(q, p);
    *d2 = ddot(d, d);
    return q;
}

// Twice the signed solid angle of face f seen from p (Van Oosterom and
// Strackee); the generalized winding number is the sum over all faces / 4 pi.
__device__ inline double solid_angle2(const MeshDev& m, uint32_t f, double px, double py, double pz)
{
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
    return 2.0 * atan2(det, den);
}

} // namespace reyn
