// Exact signed distance to a triangle mesh through a BVH, the sampler of the BVH
// contact variant (simulation/redmax/Body/BodyBVHObj.cpp:285-356 over
// Body/BVH/AABBTree.h:74-82, i.e. libigl's AABB tree and
// signed_distance_pseudonormal, libigl not vendored in the reference):
//   * the closest point over all triangles (FP64, Ericson's closest point on a
//     triangle by Voronoi regions), the BVH only prunes;
//   * the sign from the angle-weighted pseudonormal (Baerentzen & Aanaes 2005) of
//     the feature the closest point lies on: the face normal in the interior, the
//     uniform edge normal on an edge, the angle-weighted vertex normal at a vertex,
//     decided by barycentric coordinates <= 1e-12 as libigl's pseudonormal_test;
//   * the pseudonormal itself as the distance gradient (BodyBVHObj.cpp:295-301),
//     the Hessian zero (:520-525).
// Equidistant triangles are resolved to the smallest face index, so the result
// does not depend on the traversal order.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include "contact_math.h"

namespace atac {

// Node n of the BVH: lo/hi of its box (FP32, rounded outward from the FP64
// vertices), and in the w lanes: internal node (w0 >= 0): children w0 and w0 + 1;
// leaf (w0 < 0): triangles tri[-w0 - 1 .. -w0 - 1 + w1) of the leaf order.
struct MeshDev {
    const float4* nodes;    // 2 per node: (lo, w0) (hi, w1), w as int bits
    const int32_t* tri;     // leaf order -> face index f, stored as ~f for a degenerate face
    const double* V;        // nv x 3
    const int32_t* F;       // nf x 3
    const double* FN;       // nf x 3 face normals (zero for a degenerate face)
    const int32_t* FE;      // nf x 3: edge (opposite corner c) of face f
    const double* EN;       // ne x 3 edge normals
    const double* VN;       // nv x 3 vertex normals
    double bmin[3], bmax[3];  // the vertices' bounding box (BodyMeshObj.cpp:179-182)
    int32_t n_nodes, nf;
};

constexpr int kBvhStack = 64;        // traversal stack (the build keeps the depth below it)
constexpr double kPseudoEps = 1e-12; // barycentric threshold of the feature test

struct D3 {
    double x, y, z;
};
ATAC_HD D3 d3(double x, double y, double z) { return D3{x, y, z}; }
ATAC_HD D3 sub(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
ATAC_HD D3 add(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
ATAC_HD D3 scl(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
ATAC_HD double dt(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ATAC_HD D3 row3(const double* p, int64_t i) { return d3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

// Closest point of segment (u, v) to p (u itself for a zero-length segment).
ATAC_HD D3 closest_on_segment(D3 p, D3 u, D3 v) {
    const D3 uv = sub(v, u);
    const double l2 = dt(uv, uv);
    const double t = l2 > 0.0 ? fmin(fmax(dt(sub(p, u), uv) / l2, 0.0), 1.0) : 0.0;
    return add(u, scl(uv, t));
}

// Closest point of triangle (a, b, c) to p and its barycentric coordinates
// (Ericson, Real-Time Collision Detection, 5.1.5): the Voronoi region of p is
// found from the same dot products; vertex regions give exact (1,0,0)-type
// coordinates and edge regions an exact zero. A degenerate triangle (zero area,
// flagged when the mesh is prepared) is taken as its three edges, the first
// nearest in the order ab, bc, ca: the region tests would compare rounding
// noise there. Its coordinates are returned as (1/3, 1/3, 1/3), which selects
// its face normal, zero.
ATAC_HD D3 closest_on_triangle(D3 p, D3 a, D3 b, D3 c, double bary[3], bool degenerate) {
    if (degenerate) {
        bary[0] = bary[1] = bary[2] = 1.0 / 3.0;
        D3 best = closest_on_segment(p, a, b);
        D3 q = sub(p, best);
        double bd = dt(q, q);
        const D3 c1 = closest_on_segment(p, b, c);
        q = sub(p, c1);
        const double d1 = dt(q, q);
        if (d1 < bd) { bd = d1; best = c1; }
        const D3 c2 = closest_on_segment(p, c, a);
        q = sub(p, c2);
        if (dt(q, q) < bd) best = c2;
        return best;
    }
    const D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dt(ab, ap), d2 = dt(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { bary[0] = 1.0; bary[1] = 0.0; bary[2] = 0.0; return a; }
    const D3 bp = sub(p, b);
    const double d3v = dt(ab, bp), d4 = dt(ac, bp);
    if (d3v >= 0.0 && d4 <= d3v) { bary[0] = 0.0; bary[1] = 1.0; bary[2] = 0.0; return b; }
    const double vc = d1 * d4 - d3v * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3v <= 0.0) {
        const double v = d1 / (d1 - d3v);
        bary[0] = 1.0 - v; bary[1] = v; bary[2] = 0.0;
        return add(a, scl(ab, v));
    }
    const D3 cp = sub(p, c);
    const double d5 = dt(ab, cp), d6 = dt(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) { bary[0] = 0.0; bary[1] = 0.0; bary[2] = 1.0; return c; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        bary[0] = 1.0 - w; bary[1] = 0.0; bary[2] = w;
        return add(a, scl(ac, w));
    }
    const double va = d3v * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3v) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3v) / ((d4 - d3v) + (d5 - d6));
        bary[0] = 0.0; bary[1] = 1.0 - w; bary[2] = w;
        return add(b, scl(sub(c, b), w));
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    bary[0] = 1.0 - v - w; bary[1] = v; bary[2] = w;
    return add(a, add(scl(ab, v), scl(ac, w)));
}

// Squared distance from p to the box (FP64 from the FP32 bounds).
ATAC_HD double box_dist2(D3 p, float4 lo, float4 hi) {
    const double dx = fmax(fmax((double)lo.x - p.x, 0.0), p.x - (double)hi.x);
    const double dy = fmax(fmax((double)lo.y - p.y, 0.0), p.y - (double)hi.y);
    const double dz = fmax(fmax((double)lo.z - p.z, 0.0), p.z - (double)hi.z);
    return dx * dx + dy * dy + dz * dz;
}

ATAC_HD int node_w(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_int(v);
#else
    int r;
    std::memcpy(&r, &v, sizeof r);
    return r;
#endif
}

// Pseudonormal of face f for a closest point with barycentric coordinates b
// (libigl pseudonormal_test: two coordinates <= eps -> the remaining vertex's
// normal; one -> the normal of the edge opposite that corner; none -> the face's).
ATAC_HD D3 pseudonormal(const MeshDev& M, int f, const double b[3]) {
    const int nz = (b[0] <= kPseudoEps) + (b[1] <= kPseudoEps) + (b[2] <= kPseudoEps);
    if (nz >= 2) {
        for (int c = 0; c < 3; ++c)
            if (b[c] > kPseudoEps) return row3(M.VN, M.F[3 * f + c]);
    } else if (nz == 1) {
        for (int c = 0; c < 3; ++c)
            if (b[c] <= kPseudoEps) return row3(M.EN, M.FE[3 * f + c]);
    }
    return row3(M.FN, f);
}

// Signed distance of p (body frame) to the mesh, its pseudonormal and face:
// libigl's signed_distance_pseudonormal (sign s = +1 when (p - c) . n >= 0).
ATAC_HD double mesh_signed_distance(const MeshDev& M, D3 p, D3& normal, int& face) {
    int stack[kBvhStack];
    int sp = 0;
    double best = INFINITY;
    int bf = -1;
    D3 bc = p;
    double bb[3] = {1.0, 0.0, 0.0};
    int node = 0;
    if (M.n_nodes > 0) {
        for (;;) {
            const float4 lo = M.nodes[2 * node], hi = M.nodes[2 * node + 1];
            const int w0 = node_w(lo.w);
            if (w0 < 0) {   // leaf
                const int first = -w0 - 1, cnt = node_w(hi.w);
                for (int k = 0; k < cnt; ++k) {
                    const int tf = M.tri[first + k];
                    const int f = tf < 0 ? ~tf : tf;
                    double b[3];
                    const D3 c = closest_on_triangle(p, row3(M.V, M.F[3 * f]), row3(M.V, M.F[3 * f + 1]),
                                                     row3(M.V, M.F[3 * f + 2]), b, tf < 0);
                    const D3 pc = sub(p, c);
                    const double d2 = dt(pc, pc);
                    if (d2 < best || (d2 == best && f < bf)) {
                        best = d2; bf = f; bc = c;
                        bb[0] = b[0]; bb[1] = b[1]; bb[2] = b[2];
                    }
                }
            } else {        // nearer child first; a child is skipped if its box is farther than the best
                const float4 l0 = M.nodes[2 * w0], h0 = M.nodes[2 * w0 + 1];
                const float4 l1 = M.nodes[2 * (w0 + 1)], h1 = M.nodes[2 * (w0 + 1) + 1];
                const double e0 = box_dist2(p, l0, h0), e1 = box_dist2(p, l1, h1);
                const int near = e0 <= e1 ? w0 : w0 + 1, far = e0 <= e1 ? w0 + 1 : w0;
                const double en = fmin(e0, e1), ef = fmax(e0, e1);
                if (ef <= best && sp < kBvhStack) stack[sp++] = far;
                if (en <= best) { node = near; continue; }
            }
            // next node from the stack whose box can still hold the closest point
            bool found = false;
            while (sp > 0) {
                const int n2 = stack[--sp];
                if (box_dist2(p, M.nodes[2 * n2], M.nodes[2 * n2 + 1]) <= best) { node = n2; found = true; break; }
            }
            if (!found) break;
        }
    }
    face = bf;
    if (bf < 0) { normal = d3(0.0, 0.0, 0.0); return INFINITY; }
    normal = pseudonormal(M, bf, bb);
    const double s = dt(sub(p, bc), normal) >= 0.0 ? 1.0 : -1.0;
    return s * sqrt(best);
}

// Inside the vertices' bounding box (BodyMeshObj::filter_single, BodyMeshObj.cpp:278-286).
ATAC_HD bool mesh_box_filter(const MeshDev& M, D3 x) {
    return !(x.x < M.bmin[0] || x.y < M.bmin[1] || x.z < M.bmin[2] || x.x > M.bmax[0] || x.y > M.bmax[1] ||
             x.z > M.bmax[2]);
}

}  // namespace atac
