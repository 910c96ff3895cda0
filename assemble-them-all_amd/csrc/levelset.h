// Mesh -> signed distance grid, device side: the per-node arithmetic of the
// level-set construction (sdfgen::make_level_set3, makelevelset3.cpp:6-161 for
// the distance and crossing primitives). The float / double mix is the
// reference's, and the including file turns off FMA contraction, so a node's
// value is the same bits the sequential construction produces.
#pragma once
#include <cstdint>

#include "contact_math.h"

namespace atac {
namespace ls {

struct F3 {
    float x, y, z;
};
ATAC_HD F3 f3(float a, float b, float c) { return F3{a, b, c}; }
ATAC_HD F3 sub(const F3& a, const F3& b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
ATAC_HD F3 add(const F3& a, const F3& b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
ATAC_HD F3 scale(float s, const F3& a) { return f3(s * a.x, s * a.y, s * a.z); }
ATAC_HD float mag2(const F3& a) {
    float l = a.x * a.x;
    l += a.y * a.y;
    l += a.z * a.z;
    return l;
}
ATAC_HD float dotf(const F3& a, const F3& b) {
    float d = a.x * b.x;
    d += a.y * b.y;
    d += a.z * b.z;
    return d;
}
ATAC_HD float distf(const F3& a, const F3& b) {
    const float e0 = a.x - b.x, e1 = a.y - b.y, e2 = a.z - b.z;
    float d = e0 * e0;
    d += e1 * e1;
    d += e2 * e2;
    return sqrtf(d);   // correctly rounded in HIP device code, as on the host
}
template <class T> ATAC_HD T lmin(T a, T b) { return (b < a) ? b : a; }
template <class T> ATAC_HD T lmax(T a, T b) { return (a < b) ? b : a; }
ATAC_HD int lclamp(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// closest point of segment x1-x2 to x0: the parameter goes through double
ATAC_HD F3 seg_closest(const F3& x0, const F3& x1, const F3& x2) {
    const F3 d = sub(x2, x1);
    const double m2 = mag2(d);
    float s = (float)((double)dotf(sub(x2, x0), d) / m2);
    s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
    return add(scale(s, x1), scale(1.0f - s, x2));
}

// distance from x0 to triangle x1-x2-x3: barycentric projection on its plane,
// clamped to the two edges the projection's signs leave
ATAC_HD float tri_dist(const F3& x0, const F3& x1, const F3& x2, const F3& x3) {
    const F3 x13 = sub(x1, x3), x23 = sub(x2, x3), x03 = sub(x0, x3);
    const float m13 = mag2(x13), m23 = mag2(x23), d = dotf(x13, x23);
    const float invdet = 1.0f / lmax(m13 * m23 - d * d, 1e-30f);
    const float a = dotf(x13, x03), b = dotf(x23, x03);
    const float w23 = invdet * (m23 * a - d * b);
    const float w31 = invdet * (m13 * b - d * a);
    const float w12 = 1.0f - w23 - w31;
    if (w23 >= 0.0f && w31 >= 0.0f && w12 >= 0.0f)
        return distf(x0, add(add(scale(w23, x1), scale(w31, x2)), scale(w12, x3)));
    // w23 > 0: edges 1-2 and 1-3; else w31 > 0: 1-2 and 2-3; else 1-3 and 2-3
    const F3 a1 = (w23 > 0.0f || w31 > 0.0f) ? x2 : x3;
    const F3 b0 = w23 > 0.0f ? x1 : x2;
    return lmin(distf(x0, seg_closest(x0, x1, a1)), distf(x0, seg_closest(x0, b0, x3)));
}

// sign of the doubled area of (0,0)-(x1,y1)-(x2,y2), ties broken by the
// coordinates (0 only for a degenerate triangle)
ATAC_HD int orientation(double x1, double y1, double x2, double y2, double& area2) {
    area2 = y1 * x2 - x1 * y2;
    if (area2 > 0) return 1;
    if (area2 < 0) return -1;
    if (y2 > y1) return 1;
    if (y2 < y1) return -1;
    if (x1 > x2) return 1;
    if (x1 < x2) return -1;
    return 0;
}

// (x0, y0) in triangle (x1,y1)-(x2,y2)-(x3,y3): barycentric coordinates a, b, c
ATAC_HD bool in_triangle_2d(double x0, double y0, double x1, double y1, double x2, double y2, double x3, double y3,
                            double& a, double& b, double& c) {
    x1 -= x0; x2 -= x0; x3 -= x0;
    y1 -= y0; y2 -= y0; y3 -= y0;
    const int sa = orientation(x2, y2, x3, y3, a);
    if (sa == 0) return false;
    if (orientation(x3, y3, x1, y1, b) != sa) return false;
    if (orientation(x1, y1, x2, y2, c) != sa) return false;
    const double s = a + b + c;
    a /= s;
    b /= s;
    c /= s;
    return true;
}

}  // namespace ls
}  // namespace atac
