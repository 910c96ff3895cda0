// Drop-in replacement of simulation/redmax/Body/BodySDFObj.cpp for the
// reference's own build: it defines every member the reference's unchanged
// Body/BodySDFObj.h declares, with the grid built by the GPU level-set builder,
// kept on the GPU and queried there through libatac_contact's C ABI
// (include/atac.h). The class layout is the header's: the grid also stays in
// _SDF, _min_box, _max_box, _dx/_dy/_dz as the reference fills them; the GPU
// handle lives in a table of this file keyed by the body (the header has no
// field for it).
//
// Build: compile this file in place of the reference's BodySDFObj.cpp with
// -I<this directory> -I<repo>/include and link libatac_contact.so.
#include "Body/BodySDFObj.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "atac_route_a.h"

namespace redmax {

namespace {

// the GPU grid of a body and what it was made from
struct GridEntry {
    atac_grid grid = nullptr;
    const dtype* data = nullptr;
    int n[3] = {0, 0, 0};
    dtype box[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // min_box, max_box, dx dy dz
};

std::mutex g_mu;
std::unordered_map<const BodySDFObj*, GridEntry>& grid_table() {
    static std::unordered_map<const BodySDFObj*, GridEntry> table;
    return table;
}

void pose16(const Matrix4& E, double out[16]) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[4 * r + c] = E(r, c);
}

const double kIdentity16[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

std::string sdf_path(const std::string& obj) { return obj.substr(0, obj.size() - 4) + std::string(".sdf"); }

}  // namespace

void atac_route_a_check(int rc, const std::string& what) {
    if (rc != ATAC_OK) throw_error(what + ": " + atac_last_error());
}

atac_grid atac_route_a_grid(const BodySDFObj* b) {
    std::lock_guard<std::mutex> lock(g_mu);
    GridEntry& e = grid_table()[b];
    const dtype box[9] = {b->_min_box[0], b->_min_box[1], b->_min_box[2], b->_max_box[0], b->_max_box[1],
                          b->_max_box[2], b->_dx, b->_dy, b->_dz};
    bool same = e.grid && e.data == b->_SDF.a.data() && e.n[0] == b->_SDF.ni && e.n[1] == b->_SDF.nj &&
                e.n[2] == b->_SDF.nk;
    for (int k = 0; k < 9 && same; ++k) same = e.box[k] == box[k];
    if (same) return e.grid;
    if (e.grid) atac_grid_destroy(e.grid);
    e.grid = nullptr;
    // sdfgen's Array3 keeps a[i + ni*(j + nj*k)], the library's order; the
    // values go to the GPU in FP32, as sdfgen computed them
    std::vector<float> v(b->_SDF.a.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (float)b->_SDF.a[i];
    const double mn[3] = {box[0], box[1], box[2]}, mx[3] = {box[3], box[4], box[5]}, d[3] = {box[6], box[7], box[8]};
    atac_route_a_check(atac_grid_create(v.data(), b->_SDF.ni, b->_SDF.nj, b->_SDF.nk, mn, mx, d, &e.grid),
                       "BodySDFObj: GPU grid");
    e.data = b->_SDF.a.data();
    e.n[0] = b->_SDF.ni; e.n[1] = b->_SDF.nj; e.n[2] = b->_SDF.nk;
    for (int k = 0; k < 9; ++k) e.box[k] = box[k];
    return e.grid;
}

BodySDFObj::BodySDFObj(Simulation* sim, Joint* joint, std::string filename, Matrix3 R, Vector3 p, dtype dx, int res,
                       dtype col_th, TransformType transform_type, dtype density, Vector3 scale, bool adaptive_sample,
                       bool load_sdf, bool save_sdf)
    : BodyMeshObj(sim, joint, filename, R, p, transform_type, density, scale, adaptive_sample) {
    // BodySDFObj.cpp:6-36: a saved grid if asked for and present, else built
    // from the mesh (and saved if asked for)
    const bool loaded = load_sdf && load_SDF();
    if (!loaded) {
        precompute_SDF(dx, res);
        if (save_sdf) save_SDF();
    }
    _col_th = col_th;
    atac_route_a_grid(this);
}

// The reference's grid placement (BodySDFObj.cpp:38-77: spacing min(dx, bbox / res)
// per axis, one cell of padding around the FP32 vertex bounds) and level set
// (sdfgen::make_level_set3, exact band 1), both by the library on the GPU.
void BodySDFObj::precompute_SDF(dtype dx, int res) {
    const int nv = (int)_V.cols(), nf = (int)_F.cols();
    if (nv < 1 || nf < 1) throw_error("BodySDFObj::precompute_SDF: empty mesh");
    std::vector<double> v(3 * (size_t)nv);
    std::vector<float> vf(3 * (size_t)nv);
    for (int i = 0; i < nv; ++i)
        for (int a = 0; a < 3; ++a) {
            v[3 * (size_t)i + a] = _V(a, i);
            vf[3 * (size_t)i + a] = (float)_V(a, i);
        }
    std::vector<uint32_t> f(3 * (size_t)nf);
    for (int t = 0; t < nf; ++t)
        for (int a = 0; a < 3; ++a) f[3 * (size_t)t + a] = (uint32_t)_F(a, t);
    float mn[3], mx[3];
    double d[3];
    int32_t n[3];
    atac_route_a_check(atac_sdf_mesh_grid(v.data(), nv, dx, res, mn, mx, d, n), "BodySDFObj::precompute_SDF");
    std::vector<float> phi((size_t)n[0] * n[1] * n[2]);
    atac_route_a_check(atac_sdf_from_mesh(vf.data(), nv, f.data(), nf, mn, (float)d[0], (float)d[1], (float)d[2], n[0],
                                          n[1], n[2], 1, phi.data()),
                       "BodySDFObj::precompute_SDF");
    _dx = d[0];
    _dy = d[1];
    _dz = d[2];
    _min_box = Vector3(mn[0], mn[1], mn[2]);
    _max_box = Vector3(mx[0], mx[1], mx[2]);
    _SDF.resize(n[0], n[1], n[2]);
    for (size_t e = 0; e < phi.size(); ++e) _SDF.a[e] = (dtype)phi[e];
}

// The .sdf text file beside the .obj (BodySDFObj.cpp:86-116): n, min_box,
// max_box, spacing, then the values in Array3 order.
bool BodySDFObj::load_SDF() {
    std::ifstream in(sdf_path(_filename).c_str());
    if (!in.good()) return false;
    int ni = 0, nj = 0, nk = 0;
    in >> ni >> nj >> nk;
    in >> _min_box[0] >> _min_box[1] >> _min_box[2];
    in >> _max_box[0] >> _max_box[1] >> _max_box[2];
    in >> _dx >> _dy >> _dz;
    if (!in.good() || ni < 1 || nj < 1 || nk < 1) return false;
    _SDF.resize(ni, nj, nk);
    for (size_t i = 0; i < _SDF.a.size(); ++i) in >> _SDF.a[i];
    return !in.fail();
}

void BodySDFObj::save_SDF() {
    std::ofstream out(sdf_path(_filename).c_str());
    out << _SDF.ni << " " << _SDF.nj << " " << _SDF.nk << std::endl;
    out << _min_box[0] << " " << _min_box[1] << " " << _min_box[2] << std::endl;
    out << _max_box[0] << " " << _max_box[1] << " " << _max_box[2] << std::endl;
    out << _dx << " " << _dy << " " << _dz << std::endl;
    for (size_t i = 0; i < _SDF.a.size(); ++i) out << _SDF.a[i] << std::endl;
}

void BodySDFObj::clear_saved_SDF() { std::remove(sdf_path(_filename).c_str()); }

// ------------------------------------------------------------------ queries
// BodySDFObj::distance (BodySDFObj.cpp:207-213): query_SDF(R2^T (xw - p2)) + _col_th.
dtype BodySDFObj::distance(Vector3 xw) {
    double E[16], x[3] = {xw(0), xw(1), xw(2)}, d = 0.0;
    pose16(_E_0i, E);
    atac_route_a_check(atac_sdf_distance_host(atac_route_a_grid(this), E, _col_th, x, 1, &d), "BodySDFObj::distance");
    return d;
}

// query_SDF (BodySDFObj.cpp:123-174) at body-frame points: the library
// interpolates inside the box (FLT_MAX outside); with outside_accurate a point
// outside is moved to the nearest box point and its distance to it added.
dtype BodySDFObj::query_SDF(Vector3 x, bool outside_accurate) {
    bool inside = true;
    Vector3 xb;
    for (int a = 0; a < 3; ++a) {
        inside = inside && x(a) >= _min_box[a] && x(a) <= _max_box[a];
        xb(a) = std::min(std::max(x(a), _min_box[a]), _max_box[a]);
    }
    if (!inside && !outside_accurate) return (dtype)std::numeric_limits<float>::max();
    const Vector3 q = inside ? x : xb;
    double p[3] = {q(0), q(1), q(2)}, v = 0.0;
    atac_route_a_check(atac_sdf_distance_host(atac_route_a_grid(this), kIdentity16, 0.0, p, 1, &v),
                       "BodySDFObj::query_SDF");
    return inside ? v : (x - xb).norm() + v;
}

// The FD stencils of BodySDFObj.cpp:176-205 (eps = 1e-5, outside_accurate).
Vector3 BodySDFObj::query_dSDF(Vector3 x) {
    const dtype eps = 1e-5;
    Vector3 v;
    for (int a = 0; a < 3; ++a) {
        Vector3 h = Vector3::Zero();
        h(a) = eps;
        v(a) = query_SDF(x + h, true) - query_SDF(x - h, true);
    }
    return v / (2 * eps);
}

Matrix3 BodySDFObj::query_ddSDF(Vector3 x) {
    const dtype eps = 1e-5;
    Matrix3 v = Matrix3::Zero();
    const dtype f0 = query_SDF(x, true);
    Vector3 e[3];
    for (int a = 0; a < 3; ++a) {
        e[a] = Vector3::Zero();
        e[a](a) = eps;
    }
    for (int a = 0; a < 3; ++a) v(a, a) = query_SDF(x + e[a], true) + query_SDF(x - e[a], true) - 2 * f0;
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
            const Vector3 s = (e[a] + e[b]) / 2, t = (e[a] - e[b]) / 2;
            v(a, b) = v(b, a) = query_SDF(x + s, true) + query_SDF(x - s, true) - query_SDF(x + t, true) -
                                query_SDF(x - t, true);
        }
    return v / (eps * eps);
}

Vector3i BodySDFObj::query_grid_location(Vector3 x) {
    return Vector3i((int)std::floor((x(0) - _min_box[0]) / _dx), (int)std::floor((x(1) - _min_box[1]) / _dy),
                    (int)std::floor((x(2) - _min_box[2]) / _dz));
}

namespace {
// one point through atac_sdf_collision_host (FP64 contact arithmetic):
// values d, n, ddot, tdot, xi2 and, if derivs, the 14 derivative blocks
void collide(BodySDFObj* b, const Vector3& xw, const Vector3& xw_dot, double vals[11], double* derivs) {
    double E[16], phi[6], x[3] = {xw(0), xw(1), xw(2)}, xd[3] = {xw_dot(0), xw_dot(1), xw_dot(2)};
    pose16(b->_E_0i, E);
    for (int k = 0; k < 6; ++k) phi[k] = b->_phi(k);
    atac_route_a_check(atac_sdf_collision_host(atac_route_a_grid(b), E, phi, b->_col_th, ATAC_EVAL_PRECISE, x, xd, 1,
                                               vals, derivs),
                       "BodySDFObj::collision");
}
}  // namespace

// BodySDFObj.cpp:215-237: the contact point on this body's surface, x - d e,
// with e = 0 where the gradient norm is below 1e-5.
Vector3 BodySDFObj::surface_contact_pt(Vector3 xw) {
    double vals[11], der[135];
    collide(this, xw, Vector3::Zero(), vals, der);
    const dtype gnorm = Vector3(der[0], der[1], der[2]).norm();   // |dd/dxw| = |grad phi|
    const Matrix3 R2 = _E_0i.topLeftCorner(3, 3);
    const Vector3 x = R2.transpose() * (xw - _E_0i.topRightCorner(3, 1));
    if (std::abs(gnorm) < 1e-5) return x;
    return Vector3(vals[8], vals[9], vals[10]);
}

void BodySDFObj::collision(Vector3 xw, Vector3 xw_dot, dtype& d, Vector3& n, dtype& ddot, Vector3& tdot, Vector3& xi2) {
    double v[11];
    collide(this, xw, xw_dot, v, nullptr);
    d = v[0];
    n = Vector3(v[1], v[2], v[3]);
    ddot = v[4];
    tdot = Vector3(v[5], v[6], v[7]);
    xi2 = Vector3(v[8], v[9], v[10]);
}

void BodySDFObj::collision(Vector3 xw, Vector3 xw_dot, dtype& d, Vector3& n, dtype& ddot, Vector3& tdot, Vector3& xi2,
                           RowVector3& dd_dxw, RowVector6& dd_dq2, Matrix3& dn_dxw, Matrix36& dn_dq2,
                           RowVector3& dddot_dxw, RowVector3& dddot_dxwdot, RowVector6& dddot_dq2,
                           RowVector6& dddot_dphi2, Matrix3& dtdot_dxw, Matrix3& dtdot_dxwdot, Matrix36& dtdot_dq2,
                           Matrix36& dtdot_dphi2, Matrix3& dxi2_dxw, Matrix36& dxi2_dq2) {
    double v[11], g[135];
    collide(this, xw, xw_dot, v, g);
    d = v[0];
    n = Vector3(v[1], v[2], v[3]);
    ddot = v[4];
    tdot = Vector3(v[5], v[6], v[7]);
    xi2 = Vector3(v[8], v[9], v[10]);
    // the blocks in the C ABI's order, each row-major
    const double* p = g;
    auto take = [&p](auto& M) {
        for (int r = 0; r < M.rows(); ++r)
            for (int c = 0; c < M.cols(); ++c) M(r, c) = *p++;
    };
    take(dd_dxw); take(dd_dq2); take(dn_dxw); take(dn_dq2); take(dddot_dxw); take(dddot_dxwdot); take(dddot_dq2);
    take(dddot_dphi2); take(dtdot_dxw); take(dtdot_dxwdot); take(dtdot_dq2); take(dtdot_dphi2); take(dxi2_dxw);
    take(dxi2_dq2);
}

// ------------------------------------------------------------------ self-checks
// The reference's print-only finite-difference checks of collision()'s
// derivatives (BodySDFObj.cpp:385-737), here on random points inside the body
// with a random pose, velocity and twist: forward differences (eps = 1e-7) of
// the values against the returned blocks, pose perturbed as E <- E exp(dq).
void BodySDFObj::test_collision_derivatives_runtime(Vector3 xw, Vector3 xw_dot) {
    const dtype eps = 1e-7;
    const Matrix4 E0 = _E_0i;
    const Vector6 phi0 = _phi;
    dtype d;
    Vector3 n, tdot, xi2;
    dtype ddot;
    RowVector3 dd_dxw, dddot_dxw, dddot_dxwdot;
    RowVector6 dd_dq2, dddot_dq2, dddot_dphi2;
    Matrix3 dn_dxw, dtdot_dxw, dtdot_dxwdot, dxi2_dxw;
    Matrix36 dn_dq2, dtdot_dq2, dtdot_dphi2, dxi2_dq2;
    collision(xw, xw_dot, d, n, ddot, tdot, xi2, dd_dxw, dd_dq2, dn_dxw, dn_dq2, dddot_dxw, dddot_dxwdot, dddot_dq2,
              dddot_dphi2, dtdot_dxw, dtdot_dxwdot, dtdot_dq2, dtdot_dphi2, dxi2_dxw, dxi2_dq2);
    RowVector3 fd_dd_dxw;
    Matrix3 fd_dn_dxw, fd_dtdot_dxw, fd_dxi2_dxw;
    for (int k = 0; k < 3; ++k) {
        Vector3 xp = xw;
        xp(k) += eps;
        dtype d1, ddot1;
        Vector3 n1, t1, x1;
        collision(xp, xw_dot, d1, n1, ddot1, t1, x1);
        fd_dd_dxw(k) = (d1 - d) / eps;
        fd_dn_dxw.col(k) = (n1 - n) / eps;
        fd_dtdot_dxw.col(k) = (t1 - tdot) / eps;
        fd_dxi2_dxw.col(k) = (x1 - xi2) / eps;
    }
    RowVector6 fd_dd_dq2, fd_dddot_dphi2;
    Matrix36 fd_dn_dq2, fd_dtdot_dphi2;
    for (int k = 0; k < 6; ++k) {
        Vector6 dq = Vector6::Zero();
        dq(k) = eps;
        _E_0i = E0 * math::exp(dq);
        dtype d1, ddot1;
        Vector3 n1, t1, x1;
        collision(xw, xw_dot, d1, n1, ddot1, t1, x1);
        fd_dd_dq2(k) = (d1 - d) / eps;
        fd_dn_dq2.col(k) = (n1 - n) / eps;
        _E_0i = E0;
        _phi = phi0 + dq;
        collision(xw, xw_dot, d1, n1, ddot1, t1, x1);
        fd_dddot_dphi2(k) = (ddot1 - ddot) / eps;
        fd_dtdot_dphi2.col(k) = (t1 - tdot) / eps;
        _phi = phi0;
    }
    print_error("BodySDFObj: dd_dxw", dd_dxw, fd_dd_dxw);
    print_error("BodySDFObj: dn_dxw", dn_dxw, fd_dn_dxw);
    print_error("BodySDFObj: dtdot_dxw", dtdot_dxw, fd_dtdot_dxw);
    print_error("BodySDFObj: dxi2_dxw", dxi2_dxw, fd_dxi2_dxw);
    print_error("BodySDFObj: dd_dq2", dd_dq2, fd_dd_dq2);
    print_error("BodySDFObj: dn_dq2", dn_dq2, fd_dn_dq2);
    print_error("BodySDFObj: dddot_dphi2", dddot_dphi2, fd_dddot_dphi2);
    print_error("BodySDFObj: dtdot_dphi2", dtdot_dphi2, fd_dtdot_dphi2);
}

void BodySDFObj::test_collision_derivatives() {
    srand(1000);
    const Matrix4 E0 = _E_0i;
    const Vector6 phi0 = _phi;
    Eigen::Quaternion<dtype> q(Vector4::Random());
    q.normalize();
    Matrix4 E2 = Matrix4::Identity();
    E2.topLeftCorner(3, 3) = q.toRotationMatrix();
    E2.topRightCorner(3, 1) = Vector3::Random();
    _E_0i = E2;
    _phi = Vector6::Random();
    // a point inside the body, away from the box faces
    const Vector3 lo = _min_box + Vector3::Ones() * _dx, span = _max_box - _min_box - 2 * Vector3::Ones() * _dx;
    for (int attempt = 0; attempt < 10000; ++attempt) {
        const Vector3 x = lo + (0.5 * (Vector3::Random() + Vector3::Ones())).cwiseProduct(span);
        if (query_SDF(x) >= 0.) continue;
        const Vector3 xw = E2.topLeftCorner(3, 3) * x + E2.topRightCorner(3, 1);
        test_collision_derivatives_runtime(xw, Vector3::Random());
        break;
    }
    _E_0i = E0;
    _phi = phi0;
}

}  // namespace redmax
