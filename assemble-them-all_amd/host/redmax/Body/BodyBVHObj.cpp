// BodyBVHObj over the atac C ABI (atac_mesh_create, atac_mesh_distance_normal_host).
#include "Body/BodyBVHObj.h"

#include <algorithm>
#include <cstring>
#include <string>

namespace redmax {

namespace {
void check(int rc, const char* what) {
    if (rc != ATAC_OK) throw_error(std::string("BodyBVHObj: ") + what + ": " + atac_last_error());
}
}  // namespace

BodyBVHObj::BodyBVHObj(Simulation* sim, std::string name, std::vector<Vector3> V, std::vector<std::array<int, 3>> F,
                       dtype col_th, std::vector<Vector3> contact_points, int index0)
    : Body(sim, std::move(name), std::move(contact_points), index0), _BVH_V(std::move(V)), _BVH_F(std::move(F)),
      _col_th(col_th) {}

BodyBVHObj::~BodyBVHObj() { release_device_mesh(); }

std::pair<Vector3, Vector3> BodyBVHObj::get_AABB() const {
    Vector3 lo, hi;
    for (int a = 0; a < 3; ++a) { lo(a) = 1e300; hi(a) = -1e300; }
    for (const Vector3& v : _BVH_V)
        for (int a = 0; a < 3; ++a) {
            lo(a) = std::min(lo(a), v(a));
            hi(a) = std::max(hi(a), v(a));
        }
    return {lo, hi};
}

atac_mesh BodyBVHObj::device_mesh() {
    if (mesh_) return mesh_;
    std::vector<double> V(3 * _BVH_V.size());
    std::vector<int64_t> F(3 * _BVH_F.size());
    for (size_t i = 0; i < _BVH_V.size(); ++i)
        for (int a = 0; a < 3; ++a) V[3 * i + a] = _BVH_V[i](a);
    for (size_t f = 0; f < _BVH_F.size(); ++f)
        for (int c = 0; c < 3; ++c) F[3 * f + c] = _BVH_F[f][c];
    check(atac_mesh_create(V.data(), (int64_t)_BVH_V.size(), F.data(), (int64_t)_BVH_F.size(), &mesh_), "mesh");
    return mesh_;
}

void BodyBVHObj::release_device_mesh() {
    if (mesh_) atac_mesh_destroy(mesh_);
    mesh_ = nullptr;
}

void BodyBVHObj::distance_normal_parallel(const std::vector<Vector3>& xw, std::vector<dtype>& d,
                                          std::vector<Vector3>& dd_dx) {
    const size_t n = xw.size();
    d.assign(n, 0.0);
    dd_dx.assign(n, Vector3::Zero());
    if (n == 0) return;
    std::vector<double> x(3 * n), nrm(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) x[3 * i + a] = xw[i](a);
    double E[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) E[4 * r + c] = _E_0i(r, c);
    check(atac_mesh_distance_normal_host(device_mesh(), E, _col_th, x.data(), (int64_t)n, d.data(), nrm.data(), nullptr),
          "distance");
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) dd_dx[i](a) = nrm[3 * i + a];
}

void BodyBVHObj::distance_parallel(const std::vector<Vector3>& xw, std::vector<dtype>& d) {
    std::vector<Vector3> g;
    distance_normal_parallel(xw, d, g);
}

void BodyBVHObj::distance(Vector3 xw, dtype& d) {
    std::vector<dtype> dv;
    distance_parallel({xw}, dv);
    d = dv[0];
}

void BodyBVHObj::distance_normal(Vector3 xw, dtype& d, Vector3& dd_dx) {
    std::vector<dtype> dv;
    std::vector<Vector3> g;
    distance_normal_parallel({xw}, dv, g);
    d = dv[0];
    dd_dx = g[0];
}

}  // namespace redmax
