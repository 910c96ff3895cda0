// TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference's mesh -> signed
// distance grid construction, used by tests/ as the checker of the GPU builder.
//
// Follows sdfgen::make_level_set3 (simulation/externals/sdfgen/makelevelset3.cpp:
// point_segment_closest_point :6-18, point_triangle_distance :27-49, check_neighbour
// :87-100, sweep :102-125, orientation :129-140, point_in_triangle_2d :144-161,
// make_level_set3 :163-230) with the float / double mix of the original (Vec3f
// arithmetic in float, segment parameter through double, grid coordinates and the
// 2-D crossing test in double), and the grid placement of
// BodySDFObj::precompute_SDF (simulation/redmax/Body/BodySDFObj.cpp:38-77) and
// SDFMesh::precompute_SDF (simulation/redmax/Mesh.cpp:54-95). Built with
// -ffp-contract=off so every float operation rounds as written.
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

struct F3 {
    float v[3];
};
inline F3 f3(float a, float b, float c) { F3 r; r.v[0] = a; r.v[1] = b; r.v[2] = c; return r; }
inline F3 operator-(const F3& a, const F3& b) { return f3(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
inline F3 operator+(const F3& a, const F3& b) { return f3(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
inline F3 operator*(float s, const F3& a) { return f3(s * a.v[0], s * a.v[1], s * a.v[2]); }
inline float mag2(const F3& a) { float l = a.v[0] * a.v[0]; l += a.v[1] * a.v[1]; l += a.v[2] * a.v[2]; return l; }
inline float dot(const F3& a, const F3& b) { float d = a.v[0] * b.v[0]; d += a.v[1] * b.v[1]; d += a.v[2] * b.v[2]; return d; }
inline float dist(const F3& a, const F3& b) {
    const float e0 = a.v[0] - b.v[0], e1 = a.v[1] - b.v[1], e2 = a.v[2] - b.v[2];
    float d = e0 * e0; d += e1 * e1; d += e2 * e2;
    return std::sqrt(d);
}
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }   // std::min
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }   // std::max
template <class T> inline T smin3(T a, T b, T c) { return smin(a, smin(b, c)); }
template <class T> inline T smax3(T a, T b, T c) { return smax(a, smax(b, c)); }
inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

F3 seg_closest(const F3& x0, const F3& x1, const F3& x2) {
    const F3 dx = x2 - x1;
    const double m2 = mag2(dx);
    float s12 = (float)(dot(x2 - x0, dx) / m2);
    if (s12 < 0) s12 = 0;
    else if (s12 > 1) s12 = 1;
    return s12 * x1 + (1 - s12) * x2;
}
float seg_dist(const F3& x0, const F3& x1, const F3& x2) { return dist(x0, seg_closest(x0, x1, x2)); }

float tri_dist(const F3& x0, const F3& x1, const F3& x2, const F3& x3) {
    const F3 x13 = x1 - x3, x23 = x2 - x3, x03 = x0 - x3;
    const float m13 = mag2(x13), m23 = mag2(x23), d = dot(x13, x23);
    const float invdet = 1.f / smax(m13 * m23 - d * d, 1e-30f);
    const float a = dot(x13, x03), b = dot(x23, x03);
    const float w23 = invdet * (m23 * a - d * b);
    const float w31 = invdet * (m13 * b - d * a);
    const float w12 = 1 - w23 - w31;
    if (w23 >= 0 && w31 >= 0 && w12 >= 0) return dist(x0, w23 * x1 + w31 * x2 + w12 * x3);
    if (w23 > 0) return smin(seg_dist(x0, x1, x2), seg_dist(x0, x1, x3));
    if (w31 > 0) return smin(seg_dist(x0, x1, x2), seg_dist(x0, x2, x3));
    return smin(seg_dist(x0, x1, x3), seg_dist(x0, x2, x3));
}

int orientation(double x1, double y1, double x2, double y2, double& twice_signed_area) {
    twice_signed_area = y1 * x2 - x1 * y2;
    if (twice_signed_area > 0) return 1;
    if (twice_signed_area < 0) return -1;
    if (y2 > y1) return 1;
    if (y2 < y1) return -1;
    if (x1 > x2) return 1;
    if (x1 < x2) return -1;
    return 0;
}

bool point_in_triangle_2d(double x0, double y0, double x1, double y1, double x2, double y2, double x3, double y3,
                          double& a, double& b, double& c) {
    x1 -= x0; x2 -= x0; x3 -= x0;
    y1 -= y0; y2 -= y0; y3 -= y0;
    const int signa = orientation(x2, y2, x3, y3, a);
    if (signa == 0) return false;
    const int signb = orientation(x3, y3, x1, y1, b);
    if (signb != signa) return false;
    const int signc = orientation(x1, y1, x2, y2, c);
    if (signc != signa) return false;
    const double sum = a + b + c;
    a /= sum; b /= sum; c /= sum;
    return true;
}

struct LS {
    const F3* x;
    const uint32_t* tri;
    int64_t nf;
    F3 origin;
    float dx, dy, dz;
    int ni, nj, nk;
    std::vector<float> phi;
    std::vector<int> closest;
    int64_t id(int i, int j, int k) const { return i + (int64_t)ni * (j + (int64_t)nj * k); }
    F3 node(int i, int j, int k) const { return f3(i * dx + origin.v[0], j * dy + origin.v[1], k * dz + origin.v[2]); }
    void check(const F3& gx, int i0, int j0, int k0, int i1, int j1, int k1) {
        const int c = closest[id(i1, j1, k1)];
        if (c < 0) return;
        const uint32_t* t = tri + 3 * (int64_t)c;
        const float d = tri_dist(gx, x[t[0]], x[t[1]], x[t[2]]);
        if (d < phi[id(i0, j0, k0)]) {
            phi[id(i0, j0, k0)] = d;
            closest[id(i0, j0, k0)] = c;
        }
    }
    void sweep(int di, int dj, int dk) {
        const int i0 = di > 0 ? 1 : ni - 2, i1 = di > 0 ? ni : -1;
        const int j0 = dj > 0 ? 1 : nj - 2, j1 = dj > 0 ? nj : -1;
        const int k0 = dk > 0 ? 1 : nk - 2, k1 = dk > 0 ? nk : -1;
        for (int k = k0; k != k1; k += dk)
            for (int j = j0; j != j1; j += dj)
                for (int i = i0; i != i1; i += di) {
                    const F3 gx = node(i, j, k);
                    check(gx, i, j, k, i - di, j, k);
                    check(gx, i, j, k, i, j - dj, k);
                    check(gx, i, j, k, i - di, j - dj, k);
                    check(gx, i, j, k, i, j, k - dk);
                    check(gx, i, j, k, i - di, j, k - dk);
                    check(gx, i, j, k, i, j - dj, k - dk);
                    check(gx, i, j, k, i - di, j - dj, k - dk);
                }
    }
};

}  // namespace

extern "C" {

// make_level_set3 (makelevelset3.cpp:163-230). x: 3*nv floats, tri: 3*nf indices,
// phi: ni*nj*nk floats (x fastest), closest (optional): the closest triangle per node.
int oracle_make_level_set3(const float* x, int64_t nv, const uint32_t* tri, int64_t nf, const float* origin, float dx,
                           float dy, float dz, int ni, int nj, int nk, int exact_band, float* phi, int32_t* closest) {
    if (ni < 1 || nj < 1 || nk < 1) return -1;
    for (int64_t e = 0; e < 3 * nf; ++e)
        if (tri[e] >= (uint64_t)nv) return -1;
    LS L;
    std::vector<F3> xs((size_t)nv);
    for (int64_t v = 0; v < nv; ++v) xs[v] = f3(x[3 * v], x[3 * v + 1], x[3 * v + 2]);
    L.x = xs.data(); L.tri = tri; L.nf = nf;
    L.origin = f3(origin[0], origin[1], origin[2]);
    L.dx = dx; L.dy = dy; L.dz = dz; L.ni = ni; L.nj = nj; L.nk = nk;
    const int64_t n = (int64_t)ni * nj * nk;
    L.phi.assign((size_t)n, (ni + nj + nk) * (dx + dy + dz));
    L.closest.assign((size_t)n, -1);
    std::vector<int> count((size_t)n, 0);
    for (int64_t t = 0; t < nf; ++t) {
        const uint32_t p = tri[3 * t], q = tri[3 * t + 1], r = tri[3 * t + 2];
        const F3 &xp = xs[p], &xq = xs[q], &xr = xs[r];
        const double fip = ((double)xp.v[0] - L.origin.v[0]) / dx, fjp = ((double)xp.v[1] - L.origin.v[1]) / dy,
                     fkp = ((double)xp.v[2] - L.origin.v[2]) / dz;
        const double fiq = ((double)xq.v[0] - L.origin.v[0]) / dx, fjq = ((double)xq.v[1] - L.origin.v[1]) / dy,
                     fkq = ((double)xq.v[2] - L.origin.v[2]) / dz;
        const double fir = ((double)xr.v[0] - L.origin.v[0]) / dx, fjr = ((double)xr.v[1] - L.origin.v[1]) / dy,
                     fkr = ((double)xr.v[2] - L.origin.v[2]) / dz;
        int i0 = clampi(int(smin3(fip, fiq, fir)) - exact_band, 0, ni - 1);
        int i1 = clampi(int(smax3(fip, fiq, fir)) + exact_band + 1, 0, ni - 1);
        int j0 = clampi(int(smin3(fjp, fjq, fjr)) - exact_band, 0, nj - 1);
        int j1 = clampi(int(smax3(fjp, fjq, fjr)) + exact_band + 1, 0, nj - 1);
        int k0 = clampi(int(smin3(fkp, fkq, fkr)) - exact_band, 0, nk - 1);
        int k1 = clampi(int(smax3(fkp, fkq, fkr)) + exact_band + 1, 0, nk - 1);
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i) {
                    const float d = tri_dist(L.node(i, j, k), xp, xq, xr);
                    if (d < L.phi[L.id(i, j, k)]) {
                        L.phi[L.id(i, j, k)] = d;
                        L.closest[L.id(i, j, k)] = (int)t;
                    }
                }
        j0 = clampi((int)std::ceil(smin3(fjp, fjq, fjr)), 0, nj - 1);
        j1 = clampi((int)std::floor(smax3(fjp, fjq, fjr)), 0, nj - 1);
        k0 = clampi((int)std::ceil(smin3(fkp, fkq, fkr)), 0, nk - 1);
        k1 = clampi((int)std::floor(smax3(fkp, fkq, fkr)), 0, nk - 1);
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j) {
                double a, b, c;
                if (point_in_triangle_2d(j, k, fjp, fkp, fjq, fkq, fjr, fkr, a, b, c)) {
                    const double fi = a * fip + b * fiq + c * fir;
                    const int ii = int(std::ceil(fi));
                    if (ii < 0) ++count[L.id(0, j, k)];
                    else if (ii < ni) ++count[L.id(ii, j, k)];
                }
            }
    }
    for (int pass = 0; pass < 2; ++pass) {
        L.sweep(+1, +1, +1); L.sweep(-1, -1, -1); L.sweep(+1, +1, -1); L.sweep(-1, -1, +1);
        L.sweep(+1, -1, +1); L.sweep(-1, +1, -1); L.sweep(+1, -1, -1); L.sweep(-1, +1, +1);
    }
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j) {
            int total = 0;
            for (int i = 0; i < ni; ++i) {
                total += count[L.id(i, j, k)];
                if (total % 2 == 1) L.phi[L.id(i, j, k)] = -L.phi[L.id(i, j, k)];
            }
        }
    for (int64_t e = 0; e < n; ++e) {
        phi[e] = L.phi[e];
        if (closest) closest[e] = L.closest[e];
    }
    return 0;
}

// The grid placement of precompute_SDF (BodySDFObj.cpp:38-77 with min_size = res;
// Mesh.cpp:54-95 with min_size = 20): V is 3*nv doubles (columns of the Eigen
// matrix). Outputs the float origin (min_box) and max_box, the spacing (double) and
// the node counts.
int oracle_sdf_mesh_grid(const double* V, int64_t nv, double dx, int min_size, float* min_box, float* max_box,
                         double* d, int* n) {
    if (nv < 1 || min_size < 1) return -1;
    float mn[3], mx[3];
    double bmin[3], bmax[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = std::numeric_limits<float>::max();
        mx[a] = -std::numeric_limits<float>::max();
        bmin[a] = V[a];
        bmax[a] = V[a];
    }
    for (int64_t i = 0; i < nv; ++i)
        for (int a = 0; a < 3; ++a) {
            const float p = (float)V[3 * i + a];
            if (p < mn[a]) mn[a] = p;
            if (p > mx[a]) mx[a] = p;
            bmin[a] = std::fmin(bmin[a], V[3 * i + a]);
            bmax[a] = std::fmax(bmax[a], V[3 * i + a]);
        }
    for (int a = 0; a < 3; ++a) d[a] = smin(dx, (bmax[a] - bmin[a]) / min_size);
    const float unit[3] = {(float)d[0], (float)d[1], (float)d[2]};
    for (int a = 0; a < 3; ++a) {
        mn[a] -= (float)1 * unit[a];
        mx[a] += (float)1 * unit[a];
    }
    for (int a = 0; a < 3; ++a) {
        const float half = (float)(0.5 * d[a]);
        const float ext = mx[a] - mn[a] + half;
        n[a] = (int)(unsigned int)(ext / (float)d[a]) + 1;
        min_box[a] = mn[a];
        max_box[a] = mx[a];
    }
    return 0;
}

}  // extern "C"
