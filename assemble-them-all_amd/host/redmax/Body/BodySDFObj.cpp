#include "Body/BodySDFObj.h"

#include <cfloat>
#include <fstream>

namespace redmax {

BodySDFObj::BodySDFObj(Simulation* sim, std::string name, const sdfgen::Array3<dtype>& sdf, Vector3 min_box,
                       Vector3 max_box, dtype dx, dtype dy, dtype dz, dtype col_th, std::vector<Vector3> contact_points,
                       int index0)
    : Body(sim, std::move(name), std::move(contact_points), index0),
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
