#include "Body/BodySDFObj.h"

#include "Utils.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>

namespace redmax {

BodySDFObj::BodySDFObj(Simulation* sim, Joint* joint, std::string filename, Matrix3 R, Vector3 p, dtype dx, int res,
                       dtype col_th, TransformType transform_type, dtype density, Vector3 scale, bool adaptive_sample,
                       bool load_sdf, bool save_sdf)
    : BodyMeshObj(sim, joint, std::move(filename), R, p, transform_type, density, scale, adaptive_sample) {
    bool loaded = load_sdf && load_SDF();
    if (!loaded) {
        precompute_SDF(_V, _F, dx, res, _SDF, _min_box, _max_box, _dx, _dy, _dz);
        if (save_sdf) save_SDF();
    }
    _col_th = col_th;
}

BodySDFObj::BodySDFObj(Simulation* sim, std::string name, const sdfgen::Array3<dtype>& sdf, Vector3 min_box,
                       Vector3 max_box, dtype dx, dtype dy, dtype dz, dtype col_th, std::vector<Vector3> contact_points,
                       int index0)
    : BodyMeshObj(sim, std::move(name), std::move(contact_points), index0),
      _min_box(min_box), _max_box(max_box), _dx(dx), _dy(dy), _dz(dz), _col_th(col_th), _SDF(sdf) {
    if ((size_t)_SDF.ni * _SDF.nj * _SDF.nk != _SDF.a.size() || _SDF.a.empty())
        throw_error("BodySDFObj: grid size does not match ni*nj*nk");
}

BodySDFObj::~BodySDFObj() { release_device_grid(); }

void BodySDFObj::precompute_SDF(const std::vector<Vector3>& V, const std::vector<std::array<int, 3>>& F, dtype dx,
                                int res, sdfgen::Array3<dtype>& sdf, Vector3& min_box, Vector3& max_box,
                                dtype& dx_out, dtype& dy_out, dtype& dz_out) {
    if (V.empty() || F.empty()) throw_error("BodySDFObj::precompute_SDF: empty mesh");
    std::vector<double> v(3 * V.size());
    std::vector<float> vf(3 * V.size());
    for (size_t i = 0; i < V.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            v[3 * i + a] = V[i](a);
            vf[3 * i + a] = (float)V[i](a);   // sdfgen's Vec3f vertices
        }
    std::vector<uint32_t> f(3 * F.size());
    for (size_t t = 0; t < F.size(); ++t)
        for (int a = 0; a < 3; ++a) {
            if (F[t][a] < 0 || (size_t)F[t][a] >= V.size()) throw_error("BodySDFObj::precompute_SDF: bad face index");
            f[3 * t + a] = (uint32_t)F[t][a];
        }
    float mn[3], mx[3];
    double d[3];
    int32_t n[3];
    if (atac_sdf_mesh_grid(v.data(), (int64_t)V.size(), dx, res, mn, mx, d, n) != ATAC_OK)
        throw_error(std::string("BodySDFObj::precompute_SDF: ") + atac_last_error());
    std::vector<float> phi((size_t)n[0] * n[1] * n[2]);
    if (atac_sdf_from_mesh(vf.data(), (int64_t)V.size(), f.data(), (int64_t)F.size(), mn, (float)d[0], (float)d[1],
                           (float)d[2], n[0], n[1], n[2], 1, phi.data()) != ATAC_OK)
        throw_error(std::string("BodySDFObj::precompute_SDF: ") + atac_last_error());
    sdf.resize(n[0], n[1], n[2]);
    for (size_t e = 0; e < phi.size(); ++e) sdf.a[e] = (dtype)phi[e];
    for (int a = 0; a < 3; ++a) {
        min_box(a) = mn[a];
        max_box(a) = mx[a];
    }
    dx_out = d[0];
    dy_out = d[1];
    dz_out = d[2];
}

bool BodySDFObj::load_SDF_file(const std::string& path, sdfgen::Array3<dtype>& sdf, Vector3& min_box, Vector3& max_box,
                               dtype& dx, dtype& dy, dtype& dz) {
    std::ifstream in(path.c_str());
    if (!in.good()) return false;
    int ni, nj, nk;
    in >> ni >> nj >> nk;
    in >> min_box(0) >> min_box(1) >> min_box(2);
    in >> max_box(0) >> max_box(1) >> max_box(2);
    in >> dx >> dy >> dz;
    if (!in.good() || ni < 1 || nj < 1 || nk < 1) return false;
    sdf.resize(ni, nj, nk);
    for (size_t i = 0; i < sdf.a.size(); ++i) in >> sdf.a[i];
    return !in.fail();
}

void BodySDFObj::save_SDF_file(const std::string& path) const {
    std::ofstream out(path.c_str());
    out << _SDF.ni << " " << _SDF.nj << " " << _SDF.nk << std::endl;
    out << _min_box(0) << " " << _min_box(1) << " " << _min_box(2) << std::endl;
    out << _max_box(0) << " " << _max_box(1) << " " << _max_box(2) << std::endl;
    out << _dx << " " << _dy << " " << _dz << std::endl;
    for (size_t i = 0; i < _SDF.a.size(); ++i) out << _SDF.a[i] << std::endl;
}

std::string BodySDFObj::sdf_filename() const {
    return _filename.size() >= 4 ? _filename.substr(0, _filename.size() - 4) + ".sdf" : _filename + ".sdf";
}

bool BodySDFObj::load_SDF() {
    if (_filename.empty()) return false;
    release_device_grid();
    return load_SDF_file(sdf_filename(), _SDF, _min_box, _max_box, _dx, _dy, _dz);
}

void BodySDFObj::save_SDF() const {
    if (!_filename.empty()) save_SDF_file(sdf_filename());
}

void BodySDFObj::clear_saved_SDF() const {
    if (!_filename.empty()) std::remove(sdf_filename().c_str());
}

atac_grid BodySDFObj::device_grid() {
    if (grid_) return grid_;
    std::vector<float> v(_SDF.a.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (float)_SDF.a[i];
    const double mn[3] = {_min_box(0), _min_box(1), _min_box(2)};
    const double mx[3] = {_max_box(0), _max_box(1), _max_box(2)};
    const double d[3] = {_dx, _dy, _dz};
    if (atac_grid_create(v.data(), _SDF.ni, _SDF.nj, _SDF.nk, mn, mx, d, &grid_) != ATAC_OK)
        throw_error(std::string("BodySDFObj: ") + atac_last_error());
    return grid_;
}

void BodySDFObj::release_device_grid() {
    if (grid_) atac_grid_destroy(grid_);
    grid_ = nullptr;
}

static void pose_arrays(const SE3& E, const se3& phi, double E2[16], double phi2[6]) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) E2[4 * r + c] = E(r, c);
    for (int k = 0; k < 6; ++k) phi2[k] = phi(k);
}

dtype BodySDFObj::distance(Vector3 xw) {
    double E2[16], phi2[6], x[3] = {xw(0), xw(1), xw(2)}, d = 0;
    pose_arrays(_E_0i, _phi, E2, phi2);
    if (atac_sdf_distance_host(device_grid(), E2, _col_th, x, 1, &d) != ATAC_OK)
        throw_error(std::string("BodySDFObj::distance: ") + atac_last_error());
    return d;
}

void BodySDFObj::collision_values(const Vector3& xw, const Vector3& xw_dot, double* values, double* derivs) {
    double E2[16], phi2[6], x[3] = {xw(0), xw(1), xw(2)}, v[3] = {xw_dot(0), xw_dot(1), xw_dot(2)};
    pose_arrays(_E_0i, _phi, E2, phi2);
    if (atac_sdf_collision_host(device_grid(), E2, phi2, _col_th, ATAC_EVAL_PRECISE, x, v, 1, values, derivs) != ATAC_OK)
        throw_error(std::string("BodySDFObj::collision: ") + atac_last_error());
}

void BodySDFObj::collision(Vector3 xw, Vector3 xw_dot, dtype& d, Vector3& n, dtype& ddot, Vector3& tdot, Vector3& xi2) {
    double v[11];
    collision_values(xw, xw_dot, v, nullptr);
    d = v[0];
    ddot = v[4];
    for (int k = 0; k < 3; ++k) {
        n(k) = v[1 + k];
        tdot(k) = v[5 + k];
        xi2(k) = v[8 + k];
    }
}

Vector3 BodySDFObj::surface_contact_pt(Vector3 xw) {
    double v[11], der[135];
    collision_values(xw, Vector3::Zero(), v, der);
    // |grad d| is |dd_dxw| (the first derivative block); below 1e-5 the direction is zero
    const double g = std::sqrt(der[0] * der[0] + der[1] * der[1] + der[2] * der[2]);
    Vector3 x;
    for (int a = 0; a < 3; ++a)
        x(a) = _E_0i(0, a) * (xw(0) - _E_0i(0, 3)) + _E_0i(1, a) * (xw(1) - _E_0i(1, 3)) +
               _E_0i(2, a) * (xw(2) - _E_0i(2, 3));
    if (g < 1e-5) return x;
    Vector3 xi2;
    for (int a = 0; a < 3; ++a) xi2(a) = v[8 + a];
    return xi2;
}

namespace {

// collision values as one vector: d, n (3), ddot, tdot (3), xi2 (3)
void values_of(BodySDFObj& b, const Vector3& xw, const Vector3& xwd, double out[11]) {
    dtype d, ddot;
    Vector3 n, tdot, xi2;
    b.collision(xw, xwd, d, n, ddot, tdot, xi2);
    out[0] = d;
    out[4] = ddot;
    for (int k = 0; k < 3; ++k) {
        out[1 + k] = n(k);
        out[5 + k] = tdot(k);
        out[8 + k] = xi2(k);
    }
}

Matrix4 pose_exp(const Matrix4& E, int k, double h) {
    // E exp([dq]) with dq = h e_k (k < 3: rotation, k >= 3: translation in the body frame)
    Vector6 dq;
    dq(k) = h;
    const Matrix4 X = math::exp(dq);
    Matrix4 out;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int l = 0; l < 4; ++l) s += E(i, l) * X(l, j);
            out(i, j) = s;
        }
    return out;
}

}  // namespace

dtype BodySDFObj::test_collision_derivatives_runtime(Vector3 xw, Vector3 xw_dot) {
    dtype d, ddot;
    Vector3 n, tdot, xi2;
    RowVector3 dd_dxw, dddot_dxw, dddot_dxwdot;
    RowVector6 dd_dq2, dddot_dq2, dddot_dphi2;
    Matrix3 dn_dxw, dtdot_dxw, dtdot_dxwdot, dxi2_dxw;
    Matrix36 dn_dq2, dtdot_dq2, dtdot_dphi2, dxi2_dq2;
    collision(xw, xw_dot, d, n, ddot, tdot, xi2, dd_dxw, dd_dq2, dn_dxw, dn_dq2, dddot_dxw, dddot_dxwdot, dddot_dq2,
              dddot_dphi2, dtdot_dxw, dtdot_dxwdot, dtdot_dq2, dtdot_dphi2, dxi2_dxw, dxi2_dq2);
    const double h = 1e-7;
    const Matrix4 E0 = _E_0i;
    const se3 phi0 = _phi;
    // FD[input][value][column]: input 0 xw, 1 xw_dot, 2 q2, 3 phi2
    double fd[4][11][6] = {};
    for (int in = 0; in < 4; ++in)
        for (int c = 0; c < (in < 2 ? 3 : 6); ++c) {
            double vp[11], vm[11];
            for (int s = 0; s < 2; ++s) {
                const double step = s == 0 ? h : -h;
                Vector3 x = xw, xd = xw_dot;
                if (in == 0) x(c) += step;
                if (in == 1) xd(c) += step;
                if (in == 2) _E_0i = pose_exp(E0, c, step);
                if (in == 3) _phi(c) += step;
                values_of(*this, x, xd, s == 0 ? vp : vm);
                _E_0i = E0;
                _phi = phi0;
            }
            for (int v = 0; v < 11; ++v) fd[in][v][c] = (vp[v] - vm[v]) / (2 * h);
        }
    double worst = 0.0;
    // block: value rows [v0, v0 + nv) against input `in`, analytic values a(r, c)
    auto check = [&](const char* name, int in, int v0, int nv, int nc, auto a) {
        double num = 0.0, den = 0.0;
        for (int r = 0; r < nv; ++r)
            for (int c = 0; c < nc; ++c) {
                const double e = a(r, c) - fd[in][v0 + r][c];
                num += e * e;
                den += a(r, c) * a(r, c);
            }
        const double rel = std::sqrt(num) / std::max(1.0, std::sqrt(den));
        std::fprintf(stderr, "Error %-14s = %.3e\n", name, rel);
        worst = std::max(worst, rel);
    };
    check("dd_dxw", 0, 0, 1, 3, [&](int, int c) { return dd_dxw(0, c); });
    check("dd_dq2", 2, 0, 1, 6, [&](int, int c) { return dd_dq2(0, c); });
    check("dn_dxw", 0, 1, 3, 3, [&](int r, int c) { return dn_dxw(r, c); });
    check("dn_dq2", 2, 1, 3, 6, [&](int r, int c) { return dn_dq2(r, c); });
    check("dddot_dxw", 0, 4, 1, 3, [&](int, int c) { return dddot_dxw(0, c); });
    check("dddot_dxwdot", 1, 4, 1, 3, [&](int, int c) { return dddot_dxwdot(0, c); });
    check("dddot_dq2", 2, 4, 1, 6, [&](int, int c) { return dddot_dq2(0, c); });
    check("dddot_dphi2", 3, 4, 1, 6, [&](int, int c) { return dddot_dphi2(0, c); });
    check("dtdot_dxw", 0, 5, 3, 3, [&](int r, int c) { return dtdot_dxw(r, c); });
    check("dtdot_dxwdot", 1, 5, 3, 3, [&](int r, int c) { return dtdot_dxwdot(r, c); });
    check("dtdot_dq2", 2, 5, 3, 6, [&](int r, int c) { return dtdot_dq2(r, c); });
    check("dtdot_dphi2", 3, 5, 3, 6, [&](int r, int c) { return dtdot_dphi2(r, c); });
    check("dxi2_dxw", 0, 8, 3, 3, [&](int r, int c) { return dxi2_dxw(r, c); });
    check("dxi2_dq2", 2, 8, 3, 6, [&](int r, int c) { return dxi2_dq2(r, c); });
    return worst;
}

dtype BodySDFObj::test_collision_derivatives() {
    // a random pose, twist and velocity, and a point inside the body whose cell, and
    // whose FD stencil's cells, the finite-difference steps do not leave
    std::mt19937 rng(1000);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    Vector6 r;
    for (int k = 0; k < 6; ++k) r(k) = U(rng);
    const Matrix4 E2 = math::exp(r);
    Vector6 phi2;
    for (int k = 0; k < 6; ++k) phi2(k) = U(rng);
    Vector3 xw_dot;
    for (int k = 0; k < 3; ++k) xw_dot(k) = U(rng);
    const double dd[3] = {_dx, _dy, _dz};
    for (int trial = 0; trial < 100000; ++trial) {
        Vector3 x;
        bool interior = true;
        for (int a = 0; a < 3; ++a) {
            const double lo = _min_box(a) + dd[a], hi = _max_box(a) - dd[a];
            x(a) = lo + 0.5 * (U(rng) + 1.0) * (hi - lo);
            const double f = (x(a) - _min_box(a)) / dd[a];
            const double fr = f - std::floor(f);
            interior = interior && fr > 0.01 && fr < 0.49;   // stencil +-1e-5 and +dx/2 stay in their cells
        }
        if (!interior) continue;
        const int i = (int)std::floor((x(0) - _min_box(0)) / _dx), j = (int)std::floor((x(1) - _min_box(1)) / _dy),
                  k = (int)std::floor((x(2) - _min_box(2)) / _dz);
        if (_SDF(i, j, k) >= 0.0) continue;   // inside the body
        _E_0i = E2;
        _phi = phi2;
        Vector3 xw;
        for (int a = 0; a < 3; ++a) xw(a) = E2(a, 0) * x(0) + E2(a, 1) * x(1) + E2(a, 2) * x(2) + E2(a, 3);
        return test_collision_derivatives_runtime(xw, xw_dot);
    }
    throw_error("BodySDFObj::test_collision_derivatives: no interior point found");
    return 0.0;
}

template <int R, int C>
static void unpack(const double*& p, FixedMatrix<R, C>& m) {
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) m(r, c) = *p++;
}

void BodySDFObj::collision(Vector3 xw, Vector3 xw_dot, dtype& d, Vector3& n, dtype& ddot, Vector3& tdot, Vector3& xi2,
                           RowVector3& dd_dxw, RowVector6& dd_dq2, Matrix3& dn_dxw, Matrix36& dn_dq2,
                           RowVector3& dddot_dxw, RowVector3& dddot_dxwdot, RowVector6& dddot_dq2,
                           RowVector6& dddot_dphi2, Matrix3& dtdot_dxw, Matrix3& dtdot_dxwdot, Matrix36& dtdot_dq2,
                           Matrix36& dtdot_dphi2, Matrix3& dxi2_dxw, Matrix36& dxi2_dq2) {
    double v[11], der[135];
    collision_values(xw, xw_dot, v, der);
    d = v[0];
    ddot = v[4];
    for (int k = 0; k < 3; ++k) {
        n(k) = v[1 + k];
        tdot(k) = v[5 + k];
        xi2(k) = v[8 + k];
    }
    const double* p = der;
    unpack(p, dd_dxw); unpack(p, dd_dq2); unpack(p, dn_dxw); unpack(p, dn_dq2);
    unpack(p, dddot_dxw); unpack(p, dddot_dxwdot); unpack(p, dddot_dq2); unpack(p, dddot_dphi2);
    unpack(p, dtdot_dxw); unpack(p, dtdot_dxwdot); unpack(p, dtdot_dq2); unpack(p, dtdot_dphi2);
    unpack(p, dxi2_dxw); unpack(p, dxi2_dq2);
}

}  // namespace redmax
