// GPU-backed ForceGeneralBVHContact: the public wrappers accumulate with += into the
// global maximal-coordinate arrays at each body's _index, as the reference's
// (ForceGeneralBVHContact.cpp:46-144); libatac_contact does the arithmetic.
#include "Force/ForceGeneralBVHContact.h"

#include <chrono>
#include <cstring>
#include <string>

#include "Body/Body.h"
#include "Body/BodyBVHObj.h"
#include "atac.h"

namespace redmax {

namespace {
void check(int rc, const char* what) {
    if (rc != ATAC_OK) throw_error(std::string("ForceGeneralBVHContact: ") + what + ": " + atac_last_error());
}

long long now_us() {
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

void scatter_wrench(const double* sub, VectorX& fm, const Body* a, const Body* b) {
    for (int k = 0; k < 6; ++k) {
        fm(a->_index[k]) += sub[k];
        fm(b->_index[k]) += sub[6 + k];
    }
}

void scatter_blocks(const double* sub, MatrixX& M, const Body* a, const Body* b) {
    const Body* bodies[2] = {a, b};
    for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c)
                    M(bodies[bi]->_index[r], bodies[bj]->_index[c]) += sub[12 * (6 * bi + r) + 6 * bj + c];
}
}  // namespace

ForceGeneralBVHContact::ForceGeneralBVHContact(Simulation* sim, Body* contact_body, Body* BVH_body, dtype kn, dtype kt,
                                               dtype mu, dtype damping)
    : Force(sim) {
    _contact_body = contact_body;
    _BVH_body = dynamic_cast<BodyBVHObj*>(BVH_body);
    if (_BVH_body == nullptr) throw_error("The second body in Contact should be BVH body.");
    _kn = kn;
    _kt = kt;
    _mu = mu;
    _damping = damping;
    _scale = 1.;
}

void ForceGeneralBVHContact::set_stiffness(dtype kn, dtype kt) { _kn = kn; _kt = kt; }
void ForceGeneralBVHContact::set_friction(dtype mu) { _mu = mu; }
void ForceGeneralBVHContact::set_damping(dtype damping) { _damping = damping; }
void ForceGeneralBVHContact::set_scale(dtype scale) { _scale = scale; }

void ForceGeneralBVHContact::evaluate(uint32_t flags, double* fm, double* Km, double* Dm, double* p6) {
    atac_instance in;
    std::memset(&in, 0, sizeof in);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            in.E1[4 * r + c] = _contact_body->_E_0i(r, c);
            in.E2[4 * r + c] = _BVH_body->_E_0i(r, c);
        }
    for (int k = 0; k < 6; ++k) {
        in.phi1[k] = _contact_body->_phi(k);
        in.phi2[k] = _BVH_body->_phi(k);
    }
    in.col_th = _BVH_body->_col_th;
    in.kn = _kn;
    in.kt = _kt;
    in.mu = _mu;
    in.damping = _damping;
    in.scale = _scale;
    in.contact_scale = _contact_body->_contact_scale;
    const std::vector<Vector3> pts = _contact_body->get_contact_points();
    std::vector<double> xyz(3 * pts.size());
    for (size_t i = 0; i < pts.size(); ++i)
        for (int k = 0; k < 3; ++k) xyz[3 * i + k] = pts[i](k);
    int64_t n_active = 0;
    check(atac_bvh_force_host(_BVH_body->device_mesh(), &in, xyz.data(), (int64_t)pts.size(), flags, fm, Km, Dm, p6,
                              &n_active),
          "eval");
    _last_active = (int)n_active;
    if (n_active > 0) {   // CollisionDetection.cpp:181-184
        _contact_body->add_contact_body(_BVH_body->_name);
        _BVH_body->add_contact_body(_contact_body->_name);
    }
}

void ForceGeneralBVHContact::computeForce(VectorX& fm, VectorX& fr, bool verbose) {
    const long long t0 = now_us();
    double sub[12];
    evaluate(ATAC_EVAL_FORCE, sub, nullptr, nullptr, nullptr);
    scatter_wrench(sub, fm, _contact_body, _BVH_body);
    if (verbose) std::cerr << "General BVH Contact (GPU): " << _last_active << " contacts" << std::endl;
    _time.add("collision detection", 0);
    _time.add("compute force", now_us() - t0);
}

void ForceGeneralBVHContact::computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr,
                                                        MatrixX& Dr, bool verbose) {
    const long long t0 = now_us();
    double sub[12], K[144], D[144];
    evaluate(ATAC_EVAL_DERIV, sub, K, D, nullptr);
    scatter_wrench(sub, fm, _contact_body, _BVH_body);
    scatter_blocks(K, Km, _contact_body, _BVH_body);
    scatter_blocks(D, Dm, _contact_body, _BVH_body);
    if (verbose) std::cerr << "General BVH Contact (GPU): " << _last_active << " contacts" << std::endl;
    _time.add("collision detection", 0);
    _time.add("compute force", now_us() - t0);
}

void ForceGeneralBVHContact::computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr,
                                                        MatrixX& Dr, MatrixX& dfm_dp, MatrixX& dfr_dp, bool verbose) {
    const long long t0 = now_us();
    double sub[12], K[144], D[144], p6[12];
    evaluate(ATAC_EVAL_DERIV | ATAC_EVAL_P6, sub, K, D, p6);
    scatter_wrench(sub, fm, _contact_body, _BVH_body);
    scatter_blocks(K, Km, _contact_body, _BVH_body);
    scatter_blocks(D, Dm, _contact_body, _BVH_body);
    const int col = _contact_body->_p6_column;
    if (col >= 0)
        for (int k = 0; k < 6; ++k) {
            dfm_dp(_contact_body->_index[k], col) += p6[k];
            dfm_dp(_BVH_body->_index[k], col) += p6[6 + k];
        }
    if (verbose) std::cerr << "General BVH Contact (GPU): " << _last_active << " contacts" << std::endl;
    _time.add("collision detection", 0);
    _time.add("compute force", now_us() - t0);
}

}  // namespace redmax
