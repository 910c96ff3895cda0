// FP64 CPU oracle for the point-vs-SDF contact path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
// product library. It restates, operation for operation, the reference's FP64
// routines (paths relative to the reference repository root):
//
//   query_SDF / query_dSDF / query_ddSDF   simulation/redmax/Body/BodySDFObj.cpp:123-205
//   distance                               simulation/redmax/Body/BodySDFObj.cpp:207-213
//   collision (plain)                      simulation/redmax/Body/BodySDFObj.cpp:239-274
//   collision (with derivatives)           simulation/redmax/Body/BodySDFObj.cpp:278-377
//   collision_detection_general_SDF        simulation/redmax/CollisionDetection/CollisionDetection.cpp:124-147
//   computeForce(contacts)                 simulation/redmax/Force/ForceGeneralSDFContact.cpp:147-219
//   computeForceWithDerivative(contacts)   simulation/redmax/Force/ForceGeneralSDFContact.cpp:221-447
//   contact-scale (p6) column of dfm_dp    simulation/redmax/Force/ForceGeneralSDFContact.cpp:693-699,761-765,
//                                          815-819,869-873,953-957,1017-1021
//   math::skew / math::gamma / math::exp   simulation/redmax/Utils.h:25-43,196-201,126-159
//   constants::eps = 1e-8                  simulation/redmax/Utils.h:9
//
// Two gradient modes are kept:
//   mode 0 ("reference"): central finite differences with eps = 1e-5 exactly as
//           BodySDFObj::query_dSDF / query_ddSDF compute them;
//   mode 1 ("analytic"):  closed-form partials of the trilinear interpolant in
//           the cell that contains x (Hessian diagonal = 0). Inside a cell the
//           two agree up to FP64 rounding; they differ within eps of a cell face.
//
// Grid values are FP32 (as produced by sdfgen and as stored on the GPU) and are
// promoted to FP64 on read, which is what BodySDFObj::precompute_SDF does
// (BodySDFObj.cpp:77-83). Index order is a[i + ni*(j + nj*k)] (sdfgen array3.h:70-79).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <map>
#include <array>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// ---------------------------------------------------------------- small dense types
template <int R, int C>
struct Mat {
    double a[R * C];
    double& operator()(int r, int c) { return a[r * C + c]; }
    double operator()(int r, int c) const { return a[r * C + c]; }
    double& operator[](int i) { return a[i]; }
    double operator[](int i) const { return a[i]; }
    static Mat zero() { Mat m; for (int i = 0; i < R * C; ++i) m.a[i] = 0.0; return m; }
    static Mat identity() { Mat m = zero(); for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = 1.0; return m; }
    Mat<C, R> T() const { Mat<C, R> t; for (int r = 0; r < R; ++r) for (int c = 0; c < C; ++c) t(c, r) = (*this)(r, c); return t; }
    double norm() const { double s = 0; for (int i = 0; i < R * C; ++i) s += a[i] * a[i]; return std::sqrt(s); }
    Mat& operator+=(const Mat& o) { for (int i = 0; i < R * C; ++i) a[i] += o.a[i]; return *this; }
    Mat& operator-=(const Mat& o) { for (int i = 0; i < R * C; ++i) a[i] -= o.a[i]; return *this; }
    Mat& operator/=(double s) { for (int i = 0; i < R * C; ++i) a[i] /= s; return *this; }
    template <int RR, int CC>
    Mat<RR, CC> block(int r0, int c0) const { Mat<RR, CC> b; for (int r = 0; r < RR; ++r) for (int c = 0; c < CC; ++c) b(r, c) = (*this)(r0 + r, c0 + c); return b; }
    template <int RR, int CC>
    void set_block(int r0, int c0, const Mat<RR, CC>& b) { for (int r = 0; r < RR; ++r) for (int c = 0; c < CC; ++c) (*this)(r0 + r, c0 + c) = b(r, c); }
    template <int RR, int CC>
    void add_block(int r0, int c0, const Mat<RR, CC>& b) { for (int r = 0; r < RR; ++r) for (int c = 0; c < CC; ++c) (*this)(r0 + r, c0 + c) += b(r, c); }
    template <int RR, int CC>
    void sub_block(int r0, int c0, const Mat<RR, CC>& b) { for (int r = 0; r < RR; ++r) for (int c = 0; c < CC; ++c) (*this)(r0 + r, c0 + c) -= b(r, c); }
};

template <int R, int K, int C>
inline Mat<R, C> operator*(const Mat<R, K>& A, const Mat<K, C>& B) {
    Mat<R, C> out;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += A(r, k) * B(k, c);
            out(r, c) = s;
        }
    return out;
}
template <int R, int C> inline Mat<R, C> operator+(Mat<R, C> A, const Mat<R, C>& B) { A += B; return A; }
template <int R, int C> inline Mat<R, C> operator-(Mat<R, C> A, const Mat<R, C>& B) { A -= B; return A; }
template <int R, int C> inline Mat<R, C> operator-(Mat<R, C> A) { for (int i = 0; i < R * C; ++i) A.a[i] = -A.a[i]; return A; }
template <int R, int C> inline Mat<R, C> operator*(double s, Mat<R, C> A) { for (int i = 0; i < R * C; ++i) A.a[i] *= s; return A; }
template <int R, int C> inline Mat<R, C> operator*(Mat<R, C> A, double s) { for (int i = 0; i < R * C; ++i) A.a[i] *= s; return A; }
template <int R, int C> inline Mat<R, C> operator/(Mat<R, C> A, double s) { for (int i = 0; i < R * C; ++i) A.a[i] /= s; return A; }

using V3 = Mat<3, 1>;
using RV3 = Mat<1, 3>;
using V6 = Mat<6, 1>;
using RV6 = Mat<1, 6>;
using M3 = Mat<3, 3>;
using M36 = Mat<3, 6>;
using M63 = Mat<6, 3>;
using M6 = Mat<6, 6>;
using M4 = Mat<4, 4>;
using M12 = Mat<12, 12>;
using V12 = Mat<12, 1>;

inline V3 v3(double x, double y, double z) { V3 v; v[0] = x; v[1] = y; v[2] = z; return v; }
inline V3 unit(int k) { V3 v = V3::zero(); v[k] = 1.0; return v; }
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3& a, const V3& b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}

// math::skew (Utils.h:25-43), 3-vector case
inline M3 skew(const V3& v) {
    M3 r;
    r(0, 0) = 0; r(0, 1) = -v[2]; r(0, 2) = v[1];
    r(1, 0) = v[2]; r(1, 1) = 0; r(1, 2) = -v[0];
    r(2, 0) = -v[1]; r(2, 1) = v[0]; r(2, 2) = 0;
    return r;
}
// math::gamma (Utils.h:196-201): [skew(x)^T  I], 3x6
inline M36 gamma(const V3& x) {
    M36 G = M36::zero();
    G.set_block<3, 3>(0, 0, skew(x).T());
    for (int i = 0; i < 3; ++i) G(i, 3 + i) = 1.0;
    return G;
}
// math::exp (Utils.h:126-159): se(3) twist [w; v] -> 4x4 SE(3)
inline M4 se3_exp(const V6& q) {
    M3 I = M3::identity();
    V3 w = v3(q[0], q[1], q[2]);
    double wlen = w.norm();
    M3 R = I;
    if (wlen > 1e-12) {
        w = w / wlen;
        double wx = w[0], wy = w[1], wz = w[2];
        double c = std::cos(wlen), s = std::sin(wlen);
        double c1 = 1.0 - c;
        R(0, 0) = c + wx * wx * c1; R(0, 1) = -wz * s + wx * wy * c1; R(0, 2) = wy * s + wx * wz * c1;
        R(1, 0) = wz * s + wx * wy * c1; R(1, 1) = c + wy * wy * c1; R(1, 2) = -wx * s + wy * wz * c1;
        R(2, 0) = -wy * s + wx * wz * c1; R(2, 1) = wx * s + wy * wz * c1; R(2, 2) = c + wz * wz * c1;
    }
    M4 E = M4::identity();
    E.set_block<3, 3>(0, 0, R);
    V3 v = v3(q[3], q[4], q[5]);
    if (wlen > 1e-12) {
        v = v / wlen;
        M3 A = I - R;
        V3 cc = cross(w, v);
        V3 d = A * cc;
        double wv = dot(w, v);
        V3 p = (wv * wlen) * w + d;
        E.set_block<3, 1>(0, 3, p);
    } else {
        E.set_block<3, 1>(0, 3, v);
    }
    return E;
}

// ---------------------------------------------------------------- SDF grid
struct Grid {
    const float* v;
    int ni, nj, nk;
    double minb[3], maxb[3];
    double dx, dy, dz;
    double at(int i, int j, int k) const { return (double)v[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k)]; }
};

// BodySDFObj::query_SDF, BodySDFObj.cpp:123-174. The reference indexes the grid
// without a bound when max_box extends past the last grid node; here the cell
// index is clamped to the last node instead (grids built with
// max_box = min_box + (n-1)*dx never reach that case).
double query_SDF(const Grid& G, const V3& x, bool outside_accurate) {
    bool inside = true;
    inside = inside && x[0] >= G.minb[0] && x[0] <= G.maxb[0];
    inside = inside && x[1] >= G.minb[1] && x[1] <= G.maxb[1];
    inside = inside && x[2] >= G.minb[2] && x[2] <= G.maxb[2];
    if (inside) {
        double ci = (x[0] - G.minb[0]) / G.dx, cj = (x[1] - G.minb[1]) / G.dy, ck = (x[2] - G.minb[2]) / G.dz;
        int i = (int)std::floor(ci), j = (int)std::floor(cj), k = (int)std::floor(ck);
        ci = ci - i; cj = cj - j; ck = ck - k;
        if (i > G.ni - 1) i = G.ni - 1;
        if (j > G.nj - 1) j = G.nj - 1;
        if (k > G.nk - 1) k = G.nk - 1;
        int i_pos = (i == G.ni - 1) ? i : i + 1;
        int j_pos = (j == G.nj - 1) ? j : j + 1;
        int k_pos = (k == G.nk - 1) ? k : k + 1;
        double vi_j0k0 = (1 - ci) * G.at(i, j, k) + ci * G.at(i_pos, j, k);
        double vi_j1k0 = (1 - ci) * G.at(i, j_pos, k) + ci * G.at(i_pos, j_pos, k);
        double vi_j0k1 = (1 - ci) * G.at(i, j, k_pos) + ci * G.at(i_pos, j, k_pos);
        double vi_j1k1 = (1 - ci) * G.at(i, j_pos, k_pos) + ci * G.at(i_pos, j_pos, k_pos);
        double vij_k0 = (1 - cj) * vi_j0k0 + cj * vi_j1k0;
        double vij_k1 = (1 - cj) * vi_j0k1 + cj * vi_j1k1;
        return (1 - ck) * vij_k0 + ck * vij_k1;
    }
    if (outside_accurate) {
        V3 xb;
        for (int a = 0; a < 3; ++a) xb[a] = std::min(std::max(x[a], G.minb[a]), G.maxb[a]);
        return (x - xb).norm() + query_SDF(G, xb, false);
    }
    return (double)std::numeric_limits<float>::max();
}

// BodySDFObj::query_dSDF, BodySDFObj.cpp:176-187 (central differences, eps = 1e-5)
V3 query_dSDF_fd(const Grid& G, const V3& x) {
    const double eps = 1e-5;
    V3 ex = v3(eps, 0, 0), ey = v3(0, eps, 0), ez = v3(0, 0, eps);
    V3 v;
    v[0] = query_SDF(G, x + ex, true) - query_SDF(G, x - ex, true);
    v[1] = query_SDF(G, x + ey, true) - query_SDF(G, x - ey, true);
    v[2] = query_SDF(G, x + ez, true) - query_SDF(G, x - ez, true);
    v /= (2 * eps);
    return v;
}

// BodySDFObj::query_ddSDF, BodySDFObj.cpp:189-205 (diagonal 3-point, off-diagonal
// half-step 4-point stencils, eps = 1e-5)
M3 query_ddSDF_fd(const Grid& G, const V3& x) {
    const double eps = 1e-5;
    V3 ex = v3(eps, 0, 0), ey = v3(0, eps, 0), ez = v3(0, 0, eps);
    V3 hx = ex / 2, hy = ey / 2, hz = ez / 2;
    M3 v = M3::zero();
    double xs = query_SDF(G, x, true);
    v(0, 0) = query_SDF(G, x + ex, true) + query_SDF(G, x - ex, true) - 2 * xs;
    v(1, 1) = query_SDF(G, x + ey, true) + query_SDF(G, x - ey, true) - 2 * xs;
    v(2, 2) = query_SDF(G, x + ez, true) + query_SDF(G, x - ez, true) - 2 * xs;
    v(0, 1) = v(1, 0) = query_SDF(G, x + hx + hy, true) + query_SDF(G, x - hx - hy, true) - query_SDF(G, x + hx - hy, true) - query_SDF(G, x - hx + hy, true);
    v(0, 2) = v(2, 0) = query_SDF(G, x + hx + hz, true) + query_SDF(G, x - hx - hz, true) - query_SDF(G, x + hx - hz, true) - query_SDF(G, x - hx + hz, true);
    v(1, 2) = v(2, 1) = query_SDF(G, x + hy + hz, true) + query_SDF(G, x - hy - hz, true) - query_SDF(G, x + hy - hz, true) - query_SDF(G, x - hy + hz, true);
    v /= (eps * eps);
    return v;
}

// Closed-form partials of the trilinear interpolant of the cell holding x
// (same cell selection and clamping as query_SDF). Outside the box the FD
// stencil is used, since the interpolant is not defined there.
static bool cell_of(const Grid& G, const V3& x, int idx[3], double f[3]) {
    for (int a = 0; a < 3; ++a)
        if (!(x[a] >= G.minb[a] && x[a] <= G.maxb[a])) return false;
    const double d[3] = {G.dx, G.dy, G.dz};
    const int n[3] = {G.ni, G.nj, G.nk};
    for (int a = 0; a < 3; ++a) {
        double c = (x[a] - G.minb[a]) / d[a];
        int i = (int)std::floor(c);
        f[a] = c - i;
        if (i > n[a] - 1) i = n[a] - 1;
        idx[a] = i;
    }
    return true;
}

V3 query_dSDF_analytic(const Grid& G, const V3& x) {
    int c[3]; double f[3];
    if (!cell_of(G, x, c, f)) return query_dSDF_fd(G, x);
    int i = c[0], j = c[1], k = c[2];
    int ip = (i == G.ni - 1) ? i : i + 1, jp = (j == G.nj - 1) ? j : j + 1, kp = (k == G.nk - 1) ? k : k + 1;
    double v000 = G.at(i, j, k), v100 = G.at(ip, j, k), v010 = G.at(i, jp, k), v110 = G.at(ip, jp, k);
    double v001 = G.at(i, j, kp), v101 = G.at(ip, j, kp), v011 = G.at(i, jp, kp), v111 = G.at(ip, jp, kp);
    double fx = f[0], fy = f[1], fz = f[2];
    V3 g;
    g[0] = ((1 - fy) * (1 - fz) * (v100 - v000) + fy * (1 - fz) * (v110 - v010) + (1 - fy) * fz * (v101 - v001) + fy * fz * (v111 - v011)) / G.dx;
    g[1] = ((1 - fx) * (1 - fz) * (v010 - v000) + fx * (1 - fz) * (v110 - v100) + (1 - fx) * fz * (v011 - v001) + fx * fz * (v111 - v101)) / G.dy;
    g[2] = ((1 - fx) * (1 - fy) * (v001 - v000) + fx * (1 - fy) * (v101 - v100) + (1 - fx) * fy * (v011 - v010) + fx * fy * (v111 - v110)) / G.dz;
    if (ip == i) g[0] = 0.0;
    if (jp == j) g[1] = 0.0;
    if (kp == k) g[2] = 0.0;
    return g;
}

M3 query_ddSDF_analytic(const Grid& G, const V3& x) {
    int c[3]; double f[3];
    if (!cell_of(G, x, c, f)) return query_ddSDF_fd(G, x);
    int i = c[0], j = c[1], k = c[2];
    int ip = (i == G.ni - 1) ? i : i + 1, jp = (j == G.nj - 1) ? j : j + 1, kp = (k == G.nk - 1) ? k : k + 1;
    double v000 = G.at(i, j, k), v100 = G.at(ip, j, k), v010 = G.at(i, jp, k), v110 = G.at(ip, jp, k);
    double v001 = G.at(i, j, kp), v101 = G.at(ip, j, kp), v011 = G.at(i, jp, kp), v111 = G.at(ip, jp, kp);
    double fx = f[0], fy = f[1], fz = f[2];
    M3 H = M3::zero();
    double hxy = ((1 - fz) * (v110 - v010 - v100 + v000) + fz * (v111 - v011 - v101 + v001)) / (G.dx * G.dy);
    double hxz = ((1 - fy) * (v101 - v001 - v100 + v000) + fy * (v111 - v011 - v110 + v010)) / (G.dx * G.dz);
    double hyz = ((1 - fx) * (v011 - v001 - v010 + v000) + fx * (v111 - v101 - v110 + v100)) / (G.dy * G.dz);
    if (ip == i || jp == j) hxy = 0.0;
    if (ip == i || kp == k) hxz = 0.0;
    if (jp == j || kp == k) hyz = 0.0;
    H(0, 1) = H(1, 0) = hxy;
    H(0, 2) = H(2, 0) = hxz;
    H(1, 2) = H(2, 1) = hyz;
    return H;
}

// ---------------------------------------------------------------- SDF body state
struct SDFBody {
    Grid G;
    M4 E;        // _E_0i
    V6 phi;      // _phi = [w; v] in the body frame
    double col_th;
    int mode;    // 0 reference FD, 1 analytic
    M3 R() const { return E.block<3, 3>(0, 0); }
    V3 p() const { return E.block<3, 1>(0, 3); }
    double value(const V3& x) const { return query_SDF(G, x, false); }
    V3 dSDF(const V3& x) const { return mode == 0 ? query_dSDF_fd(G, x) : query_dSDF_analytic(G, x); }
    M3 ddSDF(const V3& x) const { return mode == 0 ? query_ddSDF_fd(G, x) : query_ddSDF_analytic(G, x); }
};

// BodySDFObj::distance, BodySDFObj.cpp:207-213
template <class Body>
double distance(const Body& B, const V3& xw) {
    V3 x = B.R().T() * (xw - B.p());
    return B.value(x) + B.col_th;
}

struct CollisionOut {
    double d, ddot;
    V3 n, tdot, xi2;
    RV3 dd_dxw; RV6 dd_dq2;
    M3 dn_dxw; M36 dn_dq2;
    RV3 dddot_dxw, dddot_dxwdot; RV6 dddot_dq2, dddot_dphi2;
    M3 dtdot_dxw, dtdot_dxwdot; M36 dtdot_dq2, dtdot_dphi2;
    M3 dxi2_dxw; M36 dxi2_dq2;
};

// BodySDFObj::collision (values only), BodySDFObj.cpp:239-274
template <class Body>
void collision(const Body& B, const V3& xw, const V3& xw_dot, CollisionOut& o) {
    M3 I = M3::identity();
    M3 R2 = B.R();
    V3 p2 = B.p();
    V3 w2_dot = B.phi.template block<3, 1>(0, 0);
    V3 v2_dot = B.phi.template block<3, 1>(3, 0);
    V3 x = R2.T() * (xw - p2);
    o.d = B.value(x) + B.col_th;
    V3 dd_dx = B.dSDF(x);
    V3 e = dd_dx / dd_dx.norm();
    o.n = R2 * e;
    o.xi2 = x - o.d * e;
    V3 xdot = skew(w2_dot).T() * x + R2.T() * xw_dot - v2_dot;
    o.ddot = (dd_dx.T() * xdot)(0, 0);
    V3 xw2_dot = R2 * (cross(w2_dot, o.xi2) + v2_dot);
    V3 vw = xw_dot - xw2_dot;
    o.tdot = (I - o.n * o.n.T()) * vw;
}

// BodySDFObj::collision (values and derivatives), BodySDFObj.cpp:278-377
template <class Body>
void collision_deriv(const Body& B, const V3& xw, const V3& xw_dot, CollisionOut& o) {
    M3 I = M3::identity();
    M3 R2 = B.R();
    V3 p2 = B.p();
    V3 w2_dot = B.phi.template block<3, 1>(0, 0);
    V3 v2_dot = B.phi.template block<3, 1>(3, 0);
    V3 x = R2.T() * (xw - p2);

    o.d = B.value(x) + B.col_th;
    V3 dd_dx = B.dSDF(x);
    V3 e = dd_dx / dd_dx.norm();
    o.n = R2 * e;
    o.xi2 = x - o.d * e;
    V3 xdot = skew(w2_dot).T() * x + R2.T() * xw_dot - v2_dot;
    o.ddot = (dd_dx.T() * xdot)(0, 0);
    V3 xw2_dot = R2 * (cross(w2_dot, o.xi2) + v2_dot);
    V3 vw = xw_dot - xw2_dot;
    o.tdot = (I - o.n * o.n.T()) * vw;
    const V3& n = o.n;
    const V3& xi2 = o.xi2;
    const double d = o.d;

    V3 e1 = unit(0), e2 = unit(1), e3 = unit(2);
    M3 dx_dxw = R2.T();
    M3 dx_dw2 = skew(x);
    M3 dx_dv2 = -I;

    o.dd_dxw = dd_dx.T() * dx_dxw;
    o.dd_dq2.set_block<1, 3>(0, 0, dd_dx.T() * dx_dw2);
    o.dd_dq2.set_block<1, 3>(0, 3, dd_dx.T() * dx_dv2);

    M3 de_ddd_dx = M3::zero();
    double dd_dx_norm = dd_dx.norm();
    for (int i = 0; i < 3; ++i) {
        V3 uv = unit(i);
        V3 row = (dd_dx_norm * uv - dd_dx[i] * dd_dx / dd_dx_norm) / (dd_dx_norm * dd_dx_norm);
        for (int c = 0; c < 3; ++c) de_ddd_dx(i, c) = row[c];
    }
    M3 ddd_dx_dx = B.ddSDF(x);
    M3 de_dx = de_ddd_dx * ddd_dx_dx;
    M3 de_dxw = de_dx * dx_dxw;
    M3 de_dw2 = de_dx * dx_dw2;
    M3 de_dv2 = de_dx * dx_dv2;
    o.dn_dxw = R2 * de_dxw;
    o.dn_dq2.set_block<3, 3>(0, 0, R2 * de_dw2);
    {
        V3 c0 = R2 * skew(e1) * e, c1 = R2 * skew(e2) * e, c2 = R2 * skew(e3) * e;
        for (int r = 0; r < 3; ++r) { o.dn_dq2(r, 0) += c0[r]; o.dn_dq2(r, 1) += c1[r]; o.dn_dq2(r, 2) += c2[r]; }
    }
    o.dn_dq2.set_block<3, 3>(0, 3, R2 * de_dv2);

    o.dxi2_dxw = dx_dxw - e * o.dd_dxw - d * de_dxw;
    o.dxi2_dq2.set_block<3, 3>(0, 0, dx_dw2 - e * o.dd_dq2.block<1, 3>(0, 0) - d * de_dw2);
    o.dxi2_dq2.set_block<3, 3>(0, 3, dx_dv2 - e * o.dd_dq2.block<1, 3>(0, 3) - d * de_dv2);

    o.dddot_dxw = (dd_dx.T() * skew(w2_dot).T() + xdot.T() * ddd_dx_dx) * R2.T();
    o.dddot_dxwdot = dd_dx.T() * R2.T();
    o.dddot_dq2.set_block<1, 3>(0, 0, dd_dx.T() * (-skew(w2_dot) * skew(x) + skew(R2.T() * xw_dot)) + xdot.T() * ddd_dx_dx * dx_dw2);
    o.dddot_dq2.set_block<1, 3>(0, 3, dd_dx.T() * skew(w2_dot) + xdot.T() * ddd_dx_dx * dx_dv2);
    o.dddot_dphi2.set_block<1, 3>(0, 0, dd_dx.T() * skew(x));
    o.dddot_dphi2.set_block<1, 3>(0, 3, dd_dx.T() * (-I));

    M3 dxw2dot_dxw = R2 * skew(w2_dot) * o.dxi2_dxw;
    M3 dxw2dot_dxwdot = M3::zero();
    M36 dxw2dot_dq2;
    dxw2dot_dq2.set_block<3, 3>(0, 0, -R2 * skew(cross(w2_dot, xi2) + v2_dot) + R2 * skew(w2_dot) * o.dxi2_dq2.block<3, 3>(0, 0));
    dxw2dot_dq2.set_block<3, 3>(0, 3, R2 * skew(w2_dot) * o.dxi2_dq2.block<3, 3>(0, 3));
    M36 dxw2dot_dphi2;
    dxw2dot_dphi2.set_block<3, 1>(0, 0, R2 * skew(e1) * xi2);
    dxw2dot_dphi2.set_block<3, 1>(0, 1, R2 * skew(e2) * xi2);
    dxw2dot_dphi2.set_block<3, 1>(0, 2, R2 * skew(e3) * xi2);
    dxw2dot_dphi2.set_block<3, 3>(0, 3, R2);

    double nTvw = (n.T() * vw)(0, 0);
    M3 nnT = n * n.T();
    o.dtdot_dxw = -dxw2dot_dxw - nTvw * o.dn_dxw - n * (vw.T() * o.dn_dxw) + nnT * dxw2dot_dxw;
    o.dtdot_dxwdot = (I - nnT) * (I - dxw2dot_dxwdot);
    {
        M3 dn_w = o.dn_dq2.block<3, 3>(0, 0), dn_v = o.dn_dq2.block<3, 3>(0, 3);
        M3 dx_w = dxw2dot_dq2.block<3, 3>(0, 0), dx_v = dxw2dot_dq2.block<3, 3>(0, 3);
        o.dtdot_dq2.set_block<3, 3>(0, 0, -dx_w - nTvw * dn_w - n * (vw.T() * dn_w - n.T() * dx_w));
        o.dtdot_dq2.set_block<3, 3>(0, 3, -dx_v - nTvw * dn_v - n * (vw.T() * dn_v - n.T() * dx_v));
    }
    o.dtdot_dphi2 = (I - nnT) * (-dxw2dot_dphi2);
}

// ---------------------------------------------------------------- force
struct ForceParams {
    double kn, kt, mu, damping, scale, contact_scale;
};

struct Contact {
    V3 xi;
    int id;
};

// collision_detection_general_SDF, CollisionDetection.cpp:124-147
void detect(const M4& E1, const SDFBody& B, const double* pts, int64_t n, std::vector<Contact>& contacts, uint8_t* active) {
    M3 R1 = E1.block<3, 3>(0, 0);
    V3 p1 = E1.block<3, 1>(0, 3);
    for (int64_t i = 0; i < n; ++i) {
        V3 xi = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        V3 xw1 = R1 * xi + p1;
        double d = distance(B, xw1);
        bool hit = d < 0.;
        if (active) active[i] = hit ? 1 : 0;
        if (hit) contacts.push_back(Contact{xi, (int)i});
    }
}

// ForceGeneralSDFContact::computeForce(contacts, fm), ForceGeneralSDFContact.cpp:147-219
template <class Body>
void compute_force(const std::vector<Contact>& contacts, const M4& E1, const V6& phi1, const Body& B,
                   const ForceParams& P, V12& fm, double* scale = nullptr) {
    const double eps = 1e-8;  // constants::eps
    M3 R1 = E1.block<3, 3>(0, 0);
    V3 p1 = E1.block<3, 1>(0, 3);
    M3 R2 = B.R();
    fm = V12::zero();
    double kn = P.kn * P.scale * P.contact_scale;
    double kt = P.kt * P.scale * P.contact_scale;
    V3 w1 = phi1.block<3, 1>(0, 0), v1 = phi1.block<3, 1>(3, 0);
    if (scale) scale[0] = scale[1] = scale[2] = 0.0;
    for (const Contact& c : contacts) {
        V12 before = fm;
        V3 xi1 = c.xi;
        V3 xw1 = R1 * xi1 + p1;
        V3 xw1_dot = R1 * (skew(xi1).T() * w1 + v1);
        CollisionOut o;
        collision(B, xw1, xw1_dot, o);
        M36 Gamma1 = gamma(xi1), Gamma2 = gamma(o.xi2);
        M63 GTRT1 = Gamma1.T() * R1.T();
        V6 ni1 = GTRT1 * o.n;
        M63 GTRT2 = Gamma2.T() * R2.T();
        V6 ni2 = GTRT2 * o.n;
        V6 fc1 = (-kn * o.d + P.damping * o.ddot * o.d) * ni1;
        fm.add_block<6, 1>(0, 0, fc1);
        V6 fc2 = (kn * o.d - P.damping * o.ddot * o.d) * ni2;
        fm.add_block<6, 1>(6, 0, fc2);
        double fc_norm = fc1.norm();
        double tdot_norm = o.tdot.norm();
        if (P.mu > eps) {
            if (P.mu * fc_norm >= kt * tdot_norm - eps) fm.add_block<6, 1>(0, 0, -kt * GTRT1 * o.tdot);
            else fm.add_block<6, 1>(0, 0, -P.mu * fc_norm / tdot_norm * GTRT1 * o.tdot);
            if (P.mu * fc_norm >= kt * tdot_norm - eps) fm.add_block<6, 1>(6, 0, kt * GTRT2 * o.tdot);
            else fm.add_block<6, 1>(6, 0, P.mu * fc_norm / tdot_norm * GTRT2 * o.tdot);
        }
        if (scale) scale[0] += (fm - before).norm();
    }
}

// d(Gamma2)/d(xi2_k) (ForceGeneralSDFContact.cpp:275-279): tmp.col(j) = sum_k dGamma2_dxi2(k)^T * y * dxi2(k, j)
// dGamma2_dxi2(k)^T y = [e_k x y; 0] with the entries set at :277-279.
template <int C>
static Mat<6, C> dGammaT_times(const Mat<3, C>& dxi2, const V3& y) {
    Mat<3, 6> dG[3];
    for (int k = 0; k < 3; ++k) dG[k] = Mat<3, 6>::zero();
    dG[0](1, 2) = 1; dG[0](2, 1) = -1;
    dG[1](0, 2) = -1; dG[1](2, 0) = 1;
    dG[2](0, 1) = 1; dG[2](1, 0) = -1;
    Mat<6, C> out = Mat<6, C>::zero();
    for (int j = 0; j < C; ++j) {
        Mat<3, 6> Gj = Mat<3, 6>::zero();
        for (int k = 0; k < 3; ++k) Gj += dxi2(k, j) * dG[k];
        V6 col = Gj.T() * y;
        for (int r = 0; r < 6; ++r) out(r, j) = col[r];
    }
    return out;
}

// ForceGeneralSDFContact::computeForceWithDerivative(contacts, fm, Km, Dm[, dfm_dp]),
// ForceGeneralSDFContact.cpp:221-447; the p6 (contact scale) column follows :693-699 and the
// friction branches' p6 lines (:761-765, :815-819, :869-873, :953-957, :1017-1021).
template <class Body>
void compute_force_deriv(const std::vector<Contact>& contacts, const M4& E1, const V6& phi1, const Body& B,
                         const ForceParams& P, V12& fm, M12& Km, M12& Dm, V12* dfm_dp6, double* scale = nullptr) {
    const double eps = 1e-8;
    fm = V12::zero(); Km = M12::zero(); Dm = M12::zero();
    if (dfm_dp6) *dfm_dp6 = V12::zero();
    M3 R1 = E1.block<3, 3>(0, 0);
    V3 p1 = E1.block<3, 1>(0, 3);
    M3 R2 = B.R();
    V3 w1 = phi1.block<3, 1>(0, 0), v1 = phi1.block<3, 1>(3, 0);
    const double kn = P.kn * P.scale * P.contact_scale;
    const double kt = P.kt * P.scale * P.contact_scale;
    const double c = P.damping, mu = P.mu, cs = P.contact_scale;
    if (scale) scale[0] = scale[1] = scale[2] = 0.0;
    for (const Contact& ct : contacts) {
        const V12 fm0 = fm;
        const M12 Km0 = Km, Dm0 = Dm;
        V3 xi1 = ct.xi;
        V3 xw1 = R1 * xi1 + p1;
        V3 xw1_dot = R1 * (skew(xi1).T() * w1 + v1);
        M36 dxw1_dq1;
        dxw1_dq1.set_block<3, 3>(0, 0, -R1 * skew(xi1));
        dxw1_dq1.set_block<3, 3>(0, 3, R1);
        M36 dxw1dot_dq1 = M36::zero();
        dxw1dot_dq1.set_block<3, 3>(0, 0, -R1 * skew(skew(xi1).T() * w1 + v1));
        M36 dxw1dot_dphi1;
        dxw1dot_dphi1.set_block<3, 3>(0, 0, R1 * skew(xi1).T());
        dxw1dot_dphi1.set_block<3, 3>(0, 3, R1);

        CollisionOut o;
        collision_deriv(B, xw1, xw1_dot, o);
        const double d = o.d, ddot = o.ddot;
        const V3& n = o.n; const V3& tdot = o.tdot; const V3& xi2 = o.xi2;

        M36 dxi2_dq1 = o.dxi2_dxw * dxw1_dq1;
        M36 Gamma1 = gamma(xi1), Gamma2 = gamma(xi2);
        M63 GTRT1 = Gamma1.T() * R1.T();
        V6 ni1 = GTRT1 * n;
        M63 GTRT2 = Gamma2.T() * R2.T();
        V6 ni2 = GTRT2 * n;

        V6 fc1 = (-kn * d + c * ddot * d) * ni1;
        fm.add_block<6, 1>(0, 0, fc1);

        RV6 dd_dq1 = o.dd_dxw * dxw1_dq1;
        RV6 dddot_dq1 = o.dddot_dxw * dxw1_dq1 + o.dddot_dxwdot * dxw1dot_dq1;
        M36 dn_dq1 = o.dn_dxw * dxw1_dq1;
        RV6 dddot_dphi1 = o.dddot_dxwdot * dxw1dot_dphi1;
        M36 dtdot_dq1 = o.dtdot_dxw * dxw1_dq1 + o.dtdot_dxwdot * dxw1dot_dq1;
        M36 dtdot_dphi1 = o.dtdot_dxwdot * dxw1dot_dphi1;

        const double lam_neg = kn * d - c * ddot * d;  // (kn*d - damping*ddot*d)
        M6 dfc1_dq1 = -ni1 * (kn * dd_dq1 - c * dddot_dq1 * d - c * ddot * dd_dq1) - lam_neg * GTRT1 * dn_dq1;
        dfc1_dq1.sub_block<6, 3>(0, 0, lam_neg * Gamma1.T() * skew(R1.T() * n));
        Km.add_block<6, 6>(0, 0, dfc1_dq1);
        M6 dfc1_dq2 = -ni1 * (kn * o.dd_dq2 - c * o.dddot_dq2 * d - c * ddot * o.dd_dq2) - lam_neg * GTRT1 * o.dn_dq2;
        Km.add_block<6, 6>(0, 6, dfc1_dq2);
        M6 dfc1_dphi1 = ni1 * c * dddot_dphi1 * d;
        Dm.add_block<6, 6>(0, 0, dfc1_dphi1);
        M6 dfc1_dphi2 = ni1 * c * o.dddot_dphi2 * d;
        Dm.add_block<6, 6>(0, 6, dfc1_dphi2);
        if (dfm_dp6) dfm_dp6->add_block<6, 1>(0, 0, -kn * d * ni1 / cs);

        V6 fc2 = (kn * d - c * ddot * d) * ni2;
        fm.add_block<6, 1>(6, 0, fc2);
        V3 n2 = R2.T() * n;
        M6 tmp1 = dGammaT_times<6>(dxi2_dq1, n2);
        Km.add_block<6, 6>(6, 0, ni2 * (kn * dd_dq1 - c * dddot_dq1 * d - c * ddot * dd_dq1) + lam_neg * (tmp1 + GTRT2 * dn_dq1));
        M6 tmp2 = dGammaT_times<6>(o.dxi2_dq2, n2);
        Km.add_block<6, 6>(6, 6, ni2 * (kn * o.dd_dq2 - c * o.dddot_dq2 * d - c * ddot * o.dd_dq2) + lam_neg * (tmp2 + GTRT2 * o.dn_dq2));
        Km.add_block<6, 3>(6, 6, lam_neg * Gamma2.T() * skew(n2));
        Dm.sub_block<6, 6>(6, 0, ni2 * c * dddot_dphi1 * d);
        Dm.sub_block<6, 6>(6, 6, ni2 * c * o.dddot_dphi2 * d);
        if (dfm_dp6) dfm_dp6->add_block<6, 1>(6, 0, kn * d * ni2 / cs);

        if (mu > eps) {
            double fc_norm = fc1.norm();
            double tdot_norm = tdot.norm();
            RV6 dfc_norm_dq1 = (1. / fc_norm) * (fc1.T() * dfc1_dq1);
            RV6 dfc_norm_dq2 = (1. / fc_norm) * (fc1.T() * dfc1_dq2);
            RV6 dfc_norm_dphi1 = (1. / fc_norm) * (fc1.T() * dfc1_dphi1);
            RV6 dfc_norm_dphi2 = (1. / fc_norm) * (fc1.T() * dfc1_dphi2);
            RV6 dtdot_norm_dq1 = (1. / tdot_norm) * (tdot.T() * dtdot_dq1);
            RV6 dtdot_norm_dq2 = (1. / tdot_norm) * (tdot.T() * o.dtdot_dq2);
            RV6 dtdot_norm_dphi1 = (1. / tdot_norm) * (tdot.T() * dtdot_dphi1);
            RV6 dtdot_norm_dphi2 = (1. / tdot_norm) * (tdot.T() * o.dtdot_dphi2);
            V3 t1 = R1.T() * tdot, t2 = R2.T() * tdot;
            bool stat = mu * fc_norm >= kt * tdot_norm - eps;
            if (stat) {
                V6 fs1 = -kt * GTRT1 * tdot;
                fm.add_block<6, 1>(0, 0, fs1);
                Km.sub_block<6, 6>(0, 0, kt * GTRT1 * dtdot_dq1);
                Km.sub_block<6, 3>(0, 0, kt * Gamma1.T() * skew(t1));
                Km.sub_block<6, 6>(0, 6, kt * GTRT1 * o.dtdot_dq2);
                Dm.sub_block<6, 6>(0, 0, kt * GTRT1 * dtdot_dphi1);
                Dm.sub_block<6, 6>(0, 6, kt * GTRT1 * o.dtdot_dphi2);
                if (dfm_dp6) dfm_dp6->add_block<6, 1>(0, 0, fs1 / cs);
            } else {
                V6 fd1 = -mu * fc_norm / tdot_norm * GTRT1 * tdot;
                fm.add_block<6, 1>(0, 0, fd1);
                Km.sub_block<6, 6>(0, 0, mu / tdot_norm * GTRT1 * (tdot * dfc_norm_dq1 - fc_norm / tdot_norm * tdot * dtdot_norm_dq1 + fc_norm * dtdot_dq1));
                Km.sub_block<6, 3>(0, 0, mu * Gamma1.T() * skew(fc_norm / tdot_norm * t1));
                Km.sub_block<6, 6>(0, 6, mu / tdot_norm * GTRT1 * (tdot * dfc_norm_dq2 - fc_norm / tdot_norm * tdot * dtdot_norm_dq2 + fc_norm * o.dtdot_dq2));
                Dm.sub_block<6, 6>(0, 0, mu / tdot_norm * GTRT1 * (tdot * dfc_norm_dphi1 - fc_norm / tdot_norm * tdot * dtdot_norm_dphi1 + fc_norm * dtdot_dphi1));
                Dm.sub_block<6, 6>(0, 6, mu / tdot_norm * GTRT1 * (tdot * dfc_norm_dphi2 - fc_norm / tdot_norm * tdot * dtdot_norm_dphi2 + fc_norm * o.dtdot_dphi2));
                if (dfm_dp6) dfm_dp6->add_block<6, 1>(0, 0, fd1 / cs);
            }
            if (stat) {
                V6 fs2 = kt * GTRT2 * tdot;
                fm.add_block<6, 1>(6, 0, fs2);
                M6 tq1 = dGammaT_times<6>(dxi2_dq1, t2);
                Km.add_block<6, 6>(6, 0, kt * GTRT2 * dtdot_dq1 + kt * tq1);
                M6 tq2 = dGammaT_times<6>(o.dxi2_dq2, t2);
                Km.add_block<6, 6>(6, 6, kt * GTRT2 * o.dtdot_dq2 + kt * tq2);
                Km.add_block<6, 3>(6, 6, kt * Gamma2.T() * skew(t2));
                Dm.add_block<6, 6>(6, 0, kt * GTRT2 * dtdot_dphi1);
                Dm.add_block<6, 6>(6, 6, kt * GTRT2 * o.dtdot_dphi2);
                if (dfm_dp6) dfm_dp6->add_block<6, 1>(6, 0, fs2 / cs);
            } else {
                V6 fd2 = mu * fc_norm / tdot_norm * GTRT2 * tdot;
                fm.add_block<6, 1>(6, 0, fd2);
                M6 tq1 = dGammaT_times<6>(dxi2_dq1, t2);
                Km.add_block<6, 6>(6, 0, mu / tdot_norm * GTRT2 * (tdot * dfc_norm_dq1 - tdot * fc_norm / tdot_norm * dtdot_norm_dq1 + fc_norm * dtdot_dq1) + mu * fc_norm / tdot_norm * tq1);
                M6 tq2 = dGammaT_times<6>(o.dxi2_dq2, t2);
                Km.add_block<6, 6>(6, 6, mu / tdot_norm * GTRT2 * (tdot * dfc_norm_dq2 - tdot * fc_norm / tdot_norm * dtdot_norm_dq2 + fc_norm * o.dtdot_dq2) + mu * fc_norm / tdot_norm * tq2);
                Km.add_block<6, 3>(6, 6, mu * Gamma2.T() * skew(fc_norm / tdot_norm * t2));
                Dm.add_block<6, 6>(6, 0, mu / tdot_norm * GTRT2 * (tdot * dfc_norm_dphi1 - tdot * fc_norm / tdot_norm * dtdot_norm_dphi1 + fc_norm * dtdot_dphi1));
                Dm.add_block<6, 6>(6, 6, mu / tdot_norm * GTRT2 * (tdot * dfc_norm_dphi2 - tdot * fc_norm / tdot_norm * dtdot_norm_dphi2 + fc_norm * o.dtdot_dphi2));
                if (dfm_dp6) dfm_dp6->add_block<6, 1>(6, 0, fd2 / cs);
            }
        }
        if (scale) {
            scale[0] += (fm - fm0).norm();
            scale[1] += (Km - Km0).norm();
            scale[2] += (Dm - Dm0).norm();
        }
    }
}

// ---------------------------------------------------------------- BVH contact variant
// BodyBVHObj's sampler (Body/BodyBVHObj.cpp:285-356 over Body/BVH/AABBTree.h:74-82):
// libigl's signed_distance_pseudonormal. libigl is not vendored in the reference;
// its published algorithm is restated here: the closest point over all triangles
// (brute force; equal distances resolve to the smallest face index), barycentric
// coordinates of that point recomputed from it (libigl barycentric_coordinates),
// the feature they name (<= 1e-12: vertex, edge or face; libigl
// pseudonormal_test) and that feature's normal: face normal, edge normal = the
// normalised sum of the adjacent face normals (per_edge_normals, uniform), vertex
// normal = the normalised angle-weighted sum (per_vertex_normals, angle; angles by
// the law of cosines as libigl's internal_angles). Sign = +1 when (p - c) . n >= 0.
struct Mesh {
    std::vector<V3> V;
    std::vector<std::array<int64_t, 3>> F;
    std::vector<V3> FN, EN, VN;
    std::vector<std::array<int64_t, 3>> FE;   // edge opposite corner c
    std::vector<char> degenerate;              // zero-area faces
    V3 bmin, bmax;
};

static V3 normalized(const V3& v) {
    const double r = std::sqrt(dot(v, v));
    return r > 0.0 ? v / r : v;
}

void mesh_prepare(Mesh& m) {
    const size_t nf = m.F.size(), nv = m.V.size();
    for (int a = 0; a < 3; ++a) { m.bmin[a] = INFINITY; m.bmax[a] = -INFINITY; }
    for (const V3& v : m.V)
        for (int a = 0; a < 3; ++a) { m.bmin[a] = std::min(m.bmin[a], v[a]); m.bmax[a] = std::max(m.bmax[a], v[a]); }
    // a face is degenerate when |ab x ac|^2 <= 1e-20 |ab|^2 |ac|^2 (zero area to
    // rounding); its normal is zero. libigl leaves this case to rounding (its
    // region tests and barycentric recomputation divide zero by zero), so the
    // rule here is a definition, shared with the GPU path.
    m.FN.assign(nf, V3::zero());
    m.degenerate.assign(nf, 0);
    for (size_t f = 0; f < nf; ++f) {
        const V3 &a = m.V[m.F[f][0]], &b = m.V[m.F[f][1]], &c = m.V[m.F[f][2]];
        const V3 ab = b - a, ac = c - a, n = cross(ab, ac);
        m.degenerate[f] = dot(n, n) <= 1e-20 * dot(ab, ab) * dot(ac, ac);
        if (!m.degenerate[f]) m.FN[f] = normalized(n);
    }
    std::map<std::pair<int64_t, int64_t>, int64_t> eid;
    m.FE.assign(nf, {0, 0, 0});
    for (size_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            int64_t u = m.F[f][(c + 1) % 3], w = m.F[f][(c + 2) % 3];
            if (u > w) std::swap(u, w);
            m.FE[f][c] = eid.emplace(std::make_pair(u, w), (int64_t)eid.size()).first->second;
        }
    m.EN.assign(eid.size(), V3::zero());
    for (size_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) m.EN[m.FE[f][c]] += m.FN[f];
    for (V3& e : m.EN) e = normalized(e);
    m.VN.assign(nv, V3::zero());
    for (size_t f = 0; f < nf; ++f) {
        double l2[3];   // squared length of the edge opposite corner c
        for (int c = 0; c < 3; ++c) {
            const V3 e = m.V[m.F[f][(c + 1) % 3]] - m.V[m.F[f][(c + 2) % 3]];
            l2[c] = dot(e, e);
        }
        for (int c = 0; c < 3; ++c) {
            const double a2 = l2[(c + 1) % 3], b2 = l2[(c + 2) % 3];
            double cs = (a2 > 0.0 && b2 > 0.0) ? (a2 + b2 - l2[c]) / (2.0 * std::sqrt(a2) * std::sqrt(b2)) : 1.0;
            cs = std::min(1.0, std::max(-1.0, cs));
            m.VN[m.F[f][c]] += std::acos(cs) * m.FN[f];
        }
    }
    for (V3& v : m.VN) v = normalized(v);
}

// closest point of triangle (a, b, c) to p by its Voronoi regions (Ericson 5.1.5)
static V3 closest_point_segment(const V3& p, const V3& u, const V3& v) {
    const V3 uv = v - u;
    const double l2 = dot(uv, uv);
    const double t = l2 > 0.0 ? std::min(std::max(dot(p - u, uv) / l2, 0.0), 1.0) : 0.0;
    return u + t * uv;
}

// Ericson's regions; a degenerate triangle is taken as its three edges (the first
// nearest of ab, bc, ca)
static V3 closest_point_triangle(const V3& p, const V3& a, const V3& b, const V3& c, bool degenerate) {
    if (degenerate) {
        V3 best = closest_point_segment(p, a, b);
        double bd = dot(p - best, p - best);
        const V3 c1 = closest_point_segment(p, b, c);
        if (dot(p - c1, p - c1) < bd) { bd = dot(p - c1, p - c1); best = c1; }
        const V3 c2 = closest_point_segment(p, c, a);
        if (dot(p - c2, p - c2) < bd) best = c2;
        return best;
    }
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const V3 bp = p - b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + (d1 / (d1 - d3)) * ab;
    const V3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + (d2 / (d2 - d6)) * ac;
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
    const double den = 1.0 / (va + vb + vc);
    return a + (vb * den) * ab + (vc * den) * ac;
}

// signed distance (without col_th), pseudonormal and face of p (body frame)
double mesh_query(const Mesh& m, const V3& p, V3& n, int64_t& face) {
    double best = INFINITY;
    int64_t bf = -1;
    V3 bc = p;
    for (size_t f = 0; f < m.F.size(); ++f) {
        const V3 c = closest_point_triangle(p, m.V[m.F[f][0]], m.V[m.F[f][1]], m.V[m.F[f][2]], m.degenerate[f]);
        const V3 pc = p - c;
        const double d2 = dot(pc, pc);
        if (d2 < best) { best = d2; bf = (int64_t)f; bc = c; }
    }
    face = bf;
    const V3 &A = m.V[m.F[bf][0]], &B = m.V[m.F[bf][1]], &Cc = m.V[m.F[bf][2]];
    const V3 v0 = B - A, v1 = Cc - A, v2 = bc - A;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double den = d00 * d11 - d01 * d01;
    const double bv = (d11 * d20 - d01 * d21) / den, bw = (d00 * d21 - d01 * d20) / den;
    const double b[3] = {1.0 - bv - bw, bv, bw};
    const double eps = 1e-12;
    const int nz = (b[0] <= eps) + (b[1] <= eps) + (b[2] <= eps);
    n = m.FN[bf];
    if (m.degenerate[bf]) {
        // the face normal (zero)
    } else if (nz >= 2) {
        for (int c = 0; c < 3; ++c)
            if (b[c] > eps) { n = m.VN[m.F[bf][c]]; break; }
    } else if (nz == 1) {
        for (int c = 0; c < 3; ++c)
            if (b[c] <= eps) { n = m.EN[m.FE[bf][c]]; break; }
    }
    const double s = dot(p - bc, n) >= 0.0 ? 1.0 : -1.0;
    return s * std::sqrt(best);
}

// BodyBVHObj seen through the collision / force routines above: value = signed
// distance, gradient = pseudonormal, Hessian = 0 (BodyBVHObj.cpp:453-566 with
// _use_sdf_gradient false)
struct MeshBody {
    const Mesh* mesh;
    M4 E;
    V6 phi;
    double col_th;
    M3 R() const { return E.block<3, 3>(0, 0); }
    V3 p() const { return E.block<3, 1>(0, 3); }
    double value(const V3& x) const { V3 n; int64_t f; return mesh_query(*mesh, x, n, f); }
    V3 dSDF(const V3& x) const { V3 n; int64_t f; mesh_query(*mesh, x, n, f); return n; }
    M3 ddSDF(const V3&) const { return M3::zero(); }
    bool in_box(const V3& x) const {
        for (int a = 0; a < 3; ++a)
            if (x[a] < mesh->bmin[a] || x[a] > mesh->bmax[a]) return false;
        return true;
    }
};

// collision_detection_general_BVH (CollisionDetection.cpp:150-185) with
// distance_normal_parallel's box filter (BodyBVHObj.cpp:322-356)
void detect_bvh(const M4& E1, const MeshBody& B, const double* pts, int64_t n, std::vector<Contact>& contacts,
                uint8_t* active) {
    M3 R1 = E1.block<3, 3>(0, 0);
    V3 p1 = E1.block<3, 1>(0, 3);
    for (int64_t i = 0; i < n; ++i) {
        V3 xi = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        V3 x = B.R().T() * ((R1 * xi + p1) - B.p());
        bool hit = false;
        if (B.in_box(x)) hit = B.value(x) + B.col_th < 0.;
        if (active) active[i] = hit ? 1 : 0;
        if (hit) contacts.push_back(Contact{xi, (int)i});
    }
}

inline M4 load_M4(const double* a) { M4 m; for (int i = 0; i < 16; ++i) m[i] = a[i]; return m; }
inline V6 load_V6(const double* a) { V6 v; for (int i = 0; i < 6; ++i) v[i] = a[i]; return v; }

}  // namespace oracle

// ====================================================================== C ABI
using namespace oracle;

extern "C" {

struct oracle_grid {
    const float* values;
    int32_t n[3];
    double min_box[3];
    double max_box[3];
    double dxyz[3];
};

struct oracle_instance {
    int32_t grid;         // index into the grid array
    int32_t mode;         // 0 reference FD stencils, 1 analytic trilinear partials
    const double* points; // contact-body points, n x 3 (body frame)
    int64_t n_points;
    double E1[16], phi1[6], E2[16], phi2[6];
    double col_th, kn, kt, mu, damping, scale, contact_scale;
};

static Grid make_grid(const oracle_grid* g) {
    Grid G;
    G.v = g->values;
    G.ni = g->n[0]; G.nj = g->n[1]; G.nk = g->n[2];
    for (int a = 0; a < 3; ++a) { G.minb[a] = g->min_box[a]; G.maxb[a] = g->max_box[a]; }
    G.dx = g->dxyz[0]; G.dy = g->dxyz[1]; G.dz = g->dxyz[2];
    return G;
}

// query_SDF for n points (x: n x 3 in the SDF body frame)
int oracle_query_sdf(const oracle_grid* g, const double* x, int64_t n, int outside_accurate, double* out) {
    Grid G = make_grid(g);
    for (int64_t i = 0; i < n; ++i) out[i] = query_SDF(G, v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]), outside_accurate != 0);
    return 0;
}

// query_dSDF (out n x 3) and query_ddSDF (out n x 9, row-major) in the chosen mode
int oracle_query_dsdf(const oracle_grid* g, const double* x, int64_t n, int mode, double* grad, double* hess) {
    Grid G = make_grid(g);
    for (int64_t i = 0; i < n; ++i) {
        V3 xi = v3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        if (grad) { V3 v = mode == 0 ? query_dSDF_fd(G, xi) : query_dSDF_analytic(G, xi); for (int a = 0; a < 3; ++a) grad[3 * i + a] = v[a]; }
        if (hess) { M3 H = mode == 0 ? query_ddSDF_fd(G, xi) : query_ddSDF_analytic(G, xi); for (int a = 0; a < 9; ++a) hess[9 * i + a] = H[a]; }
    }
    return 0;
}

// BodySDFObj::distance for n world points
int oracle_distance(const oracle_grid* g, const double* E2, double col_th, const double* xw, int64_t n, double* out) {
    SDFBody B{make_grid(g), load_M4(E2), V6::zero(), col_th, 0};
    for (int64_t i = 0; i < n; ++i) out[i] = distance(B, v3(xw[3 * i], xw[3 * i + 1], xw[3 * i + 2]));
    return 0;
}

// Number of doubles written per point by oracle_collision
//   values: d, n(3), ddot, tdot(3), xi2(3)                                  = 11
//   derivs: dd_dxw(3) dd_dq2(6) dn_dxw(9) dn_dq2(18) dddot_dxw(3) dddot_dxwdot(3)
//           dddot_dq2(6) dddot_dphi2(6) dtdot_dxw(9) dtdot_dxwdot(9) dtdot_dq2(18)
//           dtdot_dphi2(18) dxi2_dxw(9) dxi2_dq2(18)                         = 135
int oracle_collision_sizes(int* n_values, int* n_derivs) { *n_values = 11; *n_derivs = 135; return 0; }

int oracle_collision(const oracle_grid* g, const double* E2, const double* phi2, double col_th, int mode,
                     const double* xw, const double* xw_dot, int64_t n, double* values, double* derivs) {
    SDFBody B{make_grid(g), load_M4(E2), load_V6(phi2), col_th, mode};
    for (int64_t i = 0; i < n; ++i) {
        CollisionOut o;
        V3 a = v3(xw[3 * i], xw[3 * i + 1], xw[3 * i + 2]);
        V3 b = v3(xw_dot[3 * i], xw_dot[3 * i + 1], xw_dot[3 * i + 2]);
        if (derivs) collision_deriv(B, a, b, o); else collision(B, a, b, o);
        double* v = values + 11 * i;
        v[0] = o.d; for (int k = 0; k < 3; ++k) v[1 + k] = o.n[k];
        v[4] = o.ddot; for (int k = 0; k < 3; ++k) v[5 + k] = o.tdot[k];
        for (int k = 0; k < 3; ++k) v[8 + k] = o.xi2[k];
        if (derivs) {
            double* p = derivs + 135 * i;
            auto put = [&](const double* src, int cnt) { std::memcpy(p, src, sizeof(double) * cnt); p += cnt; };
            put(o.dd_dxw.a, 3); put(o.dd_dq2.a, 6); put(o.dn_dxw.a, 9); put(o.dn_dq2.a, 18);
            put(o.dddot_dxw.a, 3); put(o.dddot_dxwdot.a, 3); put(o.dddot_dq2.a, 6); put(o.dddot_dphi2.a, 6);
            put(o.dtdot_dxw.a, 9); put(o.dtdot_dxwdot.a, 9); put(o.dtdot_dq2.a, 18); put(o.dtdot_dphi2.a, 18);
            put(o.dxi2_dxw.a, 9); put(o.dxi2_dq2.a, 18);
        }
    }
    return 0;
}

// One force object (contact body A against SDF body B) evaluated as the
// reference's public wrapper does: detection, then the force over the active set.
//   flags bit0: Km/Dm (computeForceWithDerivative); bit1: p6 column (requires bit0)
// fm: 12, Km/Dm: 144 row-major, dfm_dp6: 12, active: n (0/1), n_active: 1
int oracle_eval_instance(const oracle_grid* grids, const oracle_instance* in, unsigned flags,
                         double* fm, double* Km, double* Dm, double* dfm_dp6, uint8_t* active, int64_t* n_active,
                         double* scale) {
    SDFBody B{make_grid(&grids[in->grid]), load_M4(in->E2), load_V6(in->phi2), in->col_th, in->mode};
    M4 E1 = load_M4(in->E1);
    V6 phi1 = load_V6(in->phi1);
    std::vector<Contact> contacts;
    detect(E1, B, in->points, in->n_points, contacts, active);
    ForceParams P{in->kn, in->kt, in->mu, in->damping, in->scale, in->contact_scale};
    V12 f;
    if (flags & 1u) {
        M12 K, D; V12 p6;
        compute_force_deriv(contacts, E1, phi1, B, P, f, K, D, (flags & 2u) ? &p6 : nullptr, scale);
        if (Km) std::memcpy(Km, K.a, sizeof(double) * 144);
        if (Dm) std::memcpy(Dm, D.a, sizeof(double) * 144);
        if ((flags & 2u) && dfm_dp6) std::memcpy(dfm_dp6, p6.a, sizeof(double) * 12);
    } else {
        compute_force(contacts, E1, phi1, B, P, f, scale);
    }
    std::memcpy(fm, f.a, sizeof(double) * 12);
    if (n_active) *n_active = (int64_t)contacts.size();
    return 0;
}

// Many independent force objects, OpenMP over instances (threads <= 0: runtime default).
// scale (may be null): 3 per instance, as oracle_eval_instance's.
int oracle_eval_batch(const oracle_grid* grids, const oracle_instance* inst, int64_t n_inst, unsigned flags, int threads,
                      double* fm, double* Km, double* Dm, double* dfm_dp6, int64_t* n_active, double* scale) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t i = 0; i < n_inst; ++i) {
        oracle_eval_instance(grids, &inst[i], flags, fm + 12 * i, Km ? Km + 144 * i : nullptr, Dm ? Dm + 144 * i : nullptr,
                             dfm_dp6 ? dfm_dp6 + 12 * i : nullptr, nullptr, n_active ? n_active + i : nullptr,
                             scale ? scale + 3 * i : nullptr);
    }
    return 0;
}

// Force over an explicit contact list (active set held fixed, as the reference's
// FD self-test does, ForceGeneralSDFContact.cpp:1048-1110).
int oracle_force_fixed_contacts(const oracle_grid* g, const double* xi, int64_t n, const double* E1, const double* phi1,
                                const double* E2, const double* phi2, double col_th, int mode,
                                double kn, double kt, double mu, double damping, double scale, double contact_scale,
                                unsigned flags, double* fm, double* Km, double* Dm, double* dfm_dp6) {
    SDFBody B{make_grid(g), load_M4(E2), load_V6(phi2), col_th, mode};
    std::vector<Contact> contacts;
    for (int64_t i = 0; i < n; ++i) contacts.push_back(Contact{v3(xi[3 * i], xi[3 * i + 1], xi[3 * i + 2]), (int)i});
    ForceParams P{kn, kt, mu, damping, scale, contact_scale};
    V12 f;
    if (flags & 1u) {
        M12 K, D; V12 p6;
        compute_force_deriv(contacts, load_M4(E1), load_V6(phi1), B, P, f, K, D, &p6);
        if (Km) std::memcpy(Km, K.a, sizeof(double) * 144);
        if (Dm) std::memcpy(Dm, D.a, sizeof(double) * 144);
        if (dfm_dp6) std::memcpy(dfm_dp6, p6.a, sizeof(double) * 12);
    } else {
        compute_force(contacts, load_M4(E1), load_V6(phi1), B, P, f);
    }
    std::memcpy(fm, f.a, sizeof(double) * 12);
    return 0;
}

// math::exp of an se(3) twist, 4x4 row-major out (Utils.h:126-159)
int oracle_se3_exp(const double* q, double* E) {
    M4 m = se3_exp(load_V6(q));
    std::memcpy(E, m.a, sizeof(double) * 16);
    return 0;
}

// ---- BVH contact variant (brute-force signed distance; test infrastructure only)
static Mesh make_mesh(const double* V, int64_t nv, const int64_t* F, int64_t nf) {
    Mesh m;
    for (int64_t i = 0; i < nv; ++i) m.V.push_back(v3(V[3 * i], V[3 * i + 1], V[3 * i + 2]));
    for (int64_t f = 0; f < nf; ++f) m.F.push_back({F[3 * f], F[3 * f + 1], F[3 * f + 2]});
    mesh_prepare(m);
    return m;
}

// face normals (nf x 3), edge-opposite-corner indices (nf x 3), edge normals
// (ne x 3, ne returned) and vertex normals (nv x 3); any output may be NULL
int oracle_mesh_normals(const double* V, int64_t nv, const int64_t* F, int64_t nf, double* FN, int64_t* FE,
                        double* EN, double* VN, int64_t* ne) {
    Mesh m = make_mesh(V, nv, F, nf);
    for (int64_t f = 0; f < nf; ++f)
        for (int a = 0; a < 3; ++a) {
            if (FN) FN[3 * f + a] = m.FN[f][a];
            if (FE) FE[3 * f + a] = m.FE[f][a];
        }
    if (EN)
        for (size_t e = 0; e < m.EN.size(); ++e)
            for (int a = 0; a < 3; ++a) EN[3 * e + a] = m.EN[e][a];
    if (VN)
        for (int64_t v = 0; v < nv; ++v)
            for (int a = 0; a < 3; ++a) VN[3 * v + a] = m.VN[v][a];
    if (ne) *ne = (int64_t)m.EN.size();
    return 0;
}

// BodyBVHObj::distance_normal_parallel for n world points (d = 10, normal 0 and
// face -1 outside the vertices' bounding box)
int oracle_mesh_distance_normal(const double* V, int64_t nv, const int64_t* F, int64_t nf, const double* E2,
                                double col_th, const double* xw, int64_t n, double* d, double* normal, int64_t* face) {
    Mesh m = make_mesh(V, nv, F, nf);
    MeshBody B{&m, load_M4(E2), V6::zero(), col_th};
    for (int64_t i = 0; i < n; ++i) {
        const V3 x = B.R().T() * (v3(xw[3 * i], xw[3 * i + 1], xw[3 * i + 2]) - B.p());
        V3 nn = V3::zero();
        int64_t f = -1;
        double dd = 10.0;
        if (B.in_box(x)) dd = mesh_query(m, x, nn, f) + col_th;
        d[i] = dd;
        if (normal)
            for (int a = 0; a < 3; ++a) normal[3 * i + a] = nn[a];
        if (face) face[i] = f;
    }
    return 0;
}

// One ForceGeneralBVHContact evaluation (detection, then the force over the active
// set); flags as oracle_eval_instance. in->grid and in->mode are not used.
int oracle_bvh_eval(const double* V, int64_t nv, const int64_t* F, int64_t nf, const oracle_instance* in, unsigned flags,
                    double* fm, double* Km, double* Dm, double* dfm_dp6, uint8_t* active, int64_t* n_active) {
    Mesh m = make_mesh(V, nv, F, nf);
    MeshBody B{&m, load_M4(in->E2), load_V6(in->phi2), in->col_th};
    M4 E1 = load_M4(in->E1);
    V6 phi1 = load_V6(in->phi1);
    std::vector<Contact> contacts;
    detect_bvh(E1, B, in->points, in->n_points, contacts, active);
    ForceParams P{in->kn, in->kt, in->mu, in->damping, in->scale, in->contact_scale};
    V12 f;
    if (flags & 1u) {
        M12 K, D; V12 p6;
        compute_force_deriv(contacts, E1, phi1, B, P, f, K, D, (flags & 2u) ? &p6 : nullptr);
        if (Km) std::memcpy(Km, K.a, sizeof(double) * 144);
        if (Dm) std::memcpy(Dm, D.a, sizeof(double) * 144);
        if ((flags & 2u) && dfm_dp6) std::memcpy(dfm_dp6, p6.a, sizeof(double) * 12);
    } else {
        compute_force(contacts, E1, phi1, B, P, f);
    }
    std::memcpy(fm, f.a, sizeof(double) * 12);
    if (n_active) *n_active = (int64_t)contacts.size();
    return 0;
}

}  // extern "C"
