// Drop-in replacement of simulation/redmax/Force/ForceGeneralSDFContact.cpp for the
// reference's own build: every member its unchanged Force/ForceGeneralSDFContact.h
// declares, with detection and the contact wrench (Km, Dm, the contact-scale
// design column) computed on the GPU by libatac_contact (include/atac.h) and
// scattered into the caller's global arrays with += at each body's _index, as
// ForceGeneralSDFContact.cpp:46-144 does. The contact points go to the GPU once
// (again when the body's point set changes); each evaluation uploads the two
// poses and twists. The GPU handles live in a table of this file keyed by the
// force object (the header has no field for them).
//
// Design-parameter columns: the contact-scale column (type 6,
// ForceGeneralSDFContact.cpp:693-699 and the friction lines after it) is
// computed; a simulation with joint (type 1) or contact-body shape (types 2, 3)
// design parameters is refused by throw_error instead of being given partial
// derivatives.
//
// Build: compile this file in place of the reference's ForceGeneralSDFContact.cpp
// with -I<this directory> -I<repo>/include and link libatac_contact.so.
#include "Force/ForceGeneralSDFContact.h"

#include <cstring>
#include <ctime>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "Body/Body.h"
#include "Body/BodySDFObj.h"
#include "CollisionDetection/Contact.h"
#include "Simulation.h"
#include "atac_route_a.h"

namespace redmax {

namespace {

struct ForceState {
    atac_points points = nullptr;
    atac_batch batch = nullptr;
    atac_grid grid = nullptr;
    std::vector<Vector3> pts;   // the contact points the GPU copy holds
};

std::mutex g_mu;
std::unordered_map<const ForceGeneralSDFContact*, ForceState>& state_table() {
    static std::unordered_map<const ForceGeneralSDFContact*, ForceState> table;
    return table;
}

void fill_instance(const ForceGeneralSDFContact* f, atac_instance& in) {
    std::memset(&in, 0, sizeof in);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            in.E1[4 * r + c] = f->_contact_body->_E_0i(r, c);
            in.E2[4 * r + c] = f->_SDF_body->_E_0i(r, c);
        }
    for (int k = 0; k < 6; ++k) {
        in.phi1[k] = f->_contact_body->_phi(k);
        in.phi2[k] = f->_SDF_body->_phi(k);
    }
    in.col_th = f->_SDF_body->_col_th;
    in.kn = f->_kn;
    in.kt = f->_kt;
    in.mu = f->_mu;
    in.damping = f->_damping;
    in.scale = f->_scale;
    in.contact_scale = f->_contact_body->_contact_scale;
}

std::vector<double> xyz_of(const std::vector<Vector3>& pts) {
    std::vector<double> xyz(3 * pts.size());
    for (size_t i = 0; i < pts.size(); ++i)
        for (int k = 0; k < 3; ++k) xyz[3 * i + k] = pts[i](k);
    return xyz;
}

// Detection over the contact body's points and the contact wrench (with the
// flags' derivatives) of one force object: fm 12, Km/Dm 144, p6 12 (each may be
// null). Returns the contact count; stage times (ms) of detection and force.
int evaluate(ForceGeneralSDFContact* f, uint32_t flags, double* fm, double* Km, double* Dm, double* p6,
             double& ms_detect, double& ms_force) {
    std::lock_guard<std::mutex> lock(g_mu);
    ForceState& S = state_table()[f];
    const atac_grid grid = atac_route_a_grid(f->_SDF_body);
    std::vector<Vector3> pts = f->_contact_body->get_contact_points();
    atac_instance in;
    fill_instance(f, in);
    if (!S.batch || S.grid != grid || pts != S.pts) {
        if (S.batch) atac_batch_destroy(S.batch);
        if (S.points) atac_points_destroy(S.points);
        S.batch = nullptr;
        S.points = nullptr;
        const std::vector<double> xyz = xyz_of(pts);
        atac_route_a_check(atac_points_create(xyz.data(), (int64_t)pts.size(), nullptr, 1, 0.0, &S.points),
                           "ForceGeneralSDFContact: contact points");
        atac_route_a_check(atac_batch_create(&grid, 1, &S.points, 1, &in, 1, nullptr, &S.batch),
                           "ForceGeneralSDFContact: batch");
        atac_route_a_check(atac_batch_timing(S.batch, 1, nullptr, nullptr), "ForceGeneralSDFContact: timing");
        S.grid = grid;
        S.pts.swap(pts);
    }
    atac_route_a_check(atac_batch_set_instances(S.batch, &in, 1, nullptr), "ForceGeneralSDFContact: poses");
    int32_t n_active = 0;
    atac_route_a_check(atac_batch_eval_host(S.batch, flags, fm, Km, Dm, p6, &n_active), "ForceGeneralSDFContact");
    double ms[6] = {0, 0, 0, 0, 0, 0};
    int32_t launches[6];
    atac_route_a_check(atac_batch_timing(S.batch, 1, ms, launches), "ForceGeneralSDFContact: timing");
    ms_detect = ms[0] + ms[1];   // detection, slow-path samples
    ms_force = ms[2] + ms[3];    // contact wrench, reduction
    if (n_active > 0) {          // CollisionDetection.cpp:143-146
        f->_contact_body->add_contact_body(f->_SDF_body->_name);
        f->_SDF_body->add_contact_body(f->_contact_body->_name);
    }
    return n_active;
}

// The same over a given contact list: every listed point is a contact (the
// reference's contact-list overloads, ForceGeneralSDFContact.cpp:147, :221, :449).
void evaluate_list(ForceGeneralSDFContact* f, const std::vector<Contact>& contacts, uint32_t flags, double* fm,
                   double* Km, double* Dm, double* p6) {
    std::memset(fm, 0, 12 * sizeof(double));
    if (Km) std::memset(Km, 0, 144 * sizeof(double));
    if (Dm) std::memset(Dm, 0, 144 * sizeof(double));
    if (p6) std::memset(p6, 0, 12 * sizeof(double));
    if (contacts.empty()) return;
    std::vector<Vector3> pts;
    for (const Contact& c : contacts) pts.push_back(c._xi);
    const std::vector<double> xyz = xyz_of(pts);
    atac_instance in;
    fill_instance(f, in);
    atac_grid grid = atac_route_a_grid(f->_SDF_body);
    atac_tuning tu;
    std::memset(&tu, 0, sizeof tu);
    tu.max_active = -1;   // a record per listed contact
    atac_points P = nullptr;
    atac_batch B = nullptr;
    int rc = atac_points_create(xyz.data(), (int64_t)pts.size(), nullptr, 1, 0.0, &P);
    if (rc == ATAC_OK) rc = atac_batch_create(&grid, 1, &P, 1, &in, 1, &tu, &B);
    if (rc == ATAC_OK) rc = atac_batch_eval_host(B, flags | ATAC_EVAL_ALL_ACTIVE, fm, Km, Dm, p6, nullptr);
    const std::string err = rc == ATAC_OK ? "" : atac_last_error();
    if (B) atac_batch_destroy(B);
    if (P) atac_points_destroy(P);
    if (rc != ATAC_OK) throw_error("ForceGeneralSDFContact: contact list: " + err);
}

// fm (12), Km, Dm (12 x 12 row-major) of the pair into the global arrays at the
// two bodies' _index (ForceGeneralSDFContact.cpp:88-99)
void scatter(const ForceGeneralSDFContact* f, const double* fm, const double* Km, const double* Dm, VectorX& Fm,
             MatrixX* KM, MatrixX* DM) {
    const int ix[2] = {f->_contact_body->_index[0], f->_SDF_body->_index[0]};
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 6; ++k) Fm(ix[b] + k) += fm[6 * b + k];
    for (int rb = 0; rb < 2; ++rb)
        for (int cb = 0; cb < 2; ++cb)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    const int e = 12 * (6 * rb + r) + 6 * cb + c;
                    if (KM) (*KM)(ix[rb] + r, ix[cb] + c) += Km[e];
                    if (DM) (*DM)(ix[rb] + r, ix[cb] + c) += Dm[e];
                }
}

long long to_clock(double ms) { return (long long)(ms * 1e-3 * CLOCKS_PER_SEC); }

}  // namespace

ForceGeneralSDFContact::ForceGeneralSDFContact(Simulation* sim, Body* contact_body, Body* SDF_body, dtype kn, dtype kt,
                                               dtype mu, dtype damping)
    : Force(sim) {
    _contact_body = contact_body;
    _SDF_body = dynamic_cast<BodySDFObj*>(SDF_body);
    if (_SDF_body == nullptr) throw_error("The second body in Contact should be SDF body.");
    _kn = kn;
    _kt = kt;
    _mu = mu;
    _damping = damping;
    _scale = 1.;
}

void ForceGeneralSDFContact::set_stiffness(dtype kn, dtype kt) { _kn = kn; _kt = kt; }
void ForceGeneralSDFContact::set_friction(dtype mu) { _mu = mu; }
void ForceGeneralSDFContact::set_damping(dtype damping) { _damping = damping; }
void ForceGeneralSDFContact::set_scale(dtype scale) { _scale = scale; }

void ForceGeneralSDFContact::computeForce(VectorX& fm, VectorX& fr, bool verbose) {
    double f[12], msd = 0, msf = 0;
    evaluate(this, ATAC_EVAL_FORCE, f, nullptr, nullptr, nullptr, msd, msf);
    scatter(this, f, nullptr, nullptr, fm, nullptr, nullptr);
    _time.add("collision detection", to_clock(msd));
    _time.add("compute force", to_clock(msf));
}

void ForceGeneralSDFContact::computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm,
                                                        MatrixX& Kr, MatrixX& Dr, bool verbose) {
    double f[12], K[144], D[144], msd = 0, msf = 0;
    evaluate(this, ATAC_EVAL_DERIV, f, K, D, nullptr, msd, msf);
    scatter(this, f, K, D, fm, &Km, &Dm);
    _time.add("collision detection", to_clock(msd));
    _time.add("compute force", to_clock(msf));
}

void ForceGeneralSDFContact::computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm,
                                                        MatrixX& Kr, MatrixX& Dr, MatrixX& dfm_dp, MatrixX& dfr_dp,
                                                        bool verbose) {
    if (_sim->_ndof_p1 > 0 || _contact_body->_design_params_2._active || _contact_body->_design_params_3._active)
        throw_error("ForceGeneralSDFContact: only the contact-scale design parameter (type 6) is supported");
    double f[12], K[144], D[144], p6[12], msd = 0, msf = 0;
    evaluate(this, ATAC_EVAL_DERIV | ATAC_EVAL_P6, f, K, D, p6, msd, msf);
    scatter(this, f, K, D, fm, &Km, &Dm);
    if (_contact_body->_design_params_6._active) {   // ForceGeneralSDFContact.cpp:694-699
        const int p6_offset = _sim->_ndof_p1 + _sim->_ndof_p2 + _sim->_ndof_p3 + _sim->_ndof_p4 + _sim->_ndof_p5;
        const int col = _contact_body->_design_params_6._param_index(0) + p6_offset;
        for (int k = 0; k < 6; ++k) {
            dfm_dp(_contact_body->_index[0] + k, col) += p6[k];
            dfm_dp(_SDF_body->_index[0] + k, col) += p6[6 + k];
        }
    }
    _time.add("collision detection", to_clock(msd));
    _time.add("compute force", to_clock(msf));
}

// ------------------------------------------------------------------ contact-list overloads
void ForceGeneralSDFContact::computeForce(std::vector<Contact>& contacts, VectorX& fm, bool verbose) {
    double f[12];
    evaluate_list(this, contacts, ATAC_EVAL_FORCE, f, nullptr, nullptr, nullptr);
    fm = Eigen::Map<VectorX>(f, 12);
}

void ForceGeneralSDFContact::computeForceWithDerivative(std::vector<Contact>& contacts, VectorX& fm, MatrixX& Km,
                                                        MatrixX& Dm, bool verbose) {
    double f[12], K[144], D[144];
    evaluate_list(this, contacts, ATAC_EVAL_DERIV, f, K, D, nullptr);
    fm = Eigen::Map<VectorX>(f, 12);
    Km = Eigen::Map<Eigen::Matrix<double, 12, 12, Eigen::RowMajor>>(K);
    Dm = Eigen::Map<Eigen::Matrix<double, 12, 12, Eigen::RowMajor>>(D);
}

void ForceGeneralSDFContact::computeForceWithDerivative(std::vector<Contact>& contacts, VectorX& fm, MatrixX& Km,
                                                        MatrixX& Dm, MatrixX& dfm_dp, bool verbose) {
    double f[12], K[144], D[144], p6[12];
    evaluate_list(this, contacts, ATAC_EVAL_DERIV | ATAC_EVAL_P6, f, K, D, p6);
    fm = Eigen::Map<VectorX>(f, 12);
    Km = Eigen::Map<Eigen::Matrix<double, 12, 12, Eigen::RowMajor>>(K);
    Dm = Eigen::Map<Eigen::Matrix<double, 12, 12, Eigen::RowMajor>>(D);
    if (_contact_body->_design_params_6._active) {
        const int p6_offset = _sim->_ndof_p1 + _sim->_ndof_p2 + _sim->_ndof_p3 + _sim->_ndof_p4 + _sim->_ndof_p5;
        const int col = _contact_body->_design_params_6._param_index(0) + p6_offset;
        for (int k = 0; k < 6; ++k) {
            dfm_dp(_contact_body->_index[0] + k, col) += p6[k];
            dfm_dp(_SDF_body->_index[0] + k, col) += p6[6 + k];
        }
    }
}

// ------------------------------------------------------------------ self-checks
// Print-only finite-difference checks in the manner of the reference's
// (ForceGeneralSDFContact.cpp:1027-1243): the wrench of the current contacts,
// held fixed, against forward differences (eps = 1e-7) in each body's pose
// (E <- E exp(dq)) and twist, scored by print_error.
void ForceGeneralSDFContact::test_derivatives_runtime() {
    // the current contacts, as the reference's test holds them fixed
    std::vector<Contact> contacts;
    {
        const std::vector<Vector3> pts = _contact_body->get_contact_points();
        const Matrix3 R1 = _contact_body->_E_0i.topLeftCorner(3, 3);
        const Vector3 p1 = _contact_body->_E_0i.topRightCorner(3, 1);
        for (size_t i = 0; i < pts.size(); ++i) {
            const Vector3 xw = R1 * pts[i] + p1;
            const dtype d = _SDF_body->distance(xw);
            if (d < 0.) contacts.push_back(Contact(pts[i], xw, d, Vector3::Zero(), (int)i));
        }
    }
    if (contacts.empty()) return;
    const dtype eps = 1e-7;
    VectorX fm;
    MatrixX Km, Dm;
    computeForceWithDerivative(contacts, fm, Km, Dm);
    MatrixX Km_fd(12, 12), Dm_fd(12, 12);
    Body* bodies[2] = {_contact_body, _SDF_body};
    for (int b = 0; b < 2; ++b) {
        const Matrix4 E0 = bodies[b]->_E_0i;
        const Vector6 phi0 = bodies[b]->_phi;
        for (int k = 0; k < 6; ++k) {
            Vector6 dq = Vector6::Zero();
            dq(k) = eps;
            VectorX fp;
            bodies[b]->_E_0i = E0 * math::exp(dq);
            computeForce(contacts, fp);
            Km_fd.col(6 * b + k) = (fp - fm) / eps;
            bodies[b]->_E_0i = E0;
            bodies[b]->_phi = phi0 + dq;
            computeForce(contacts, fp);
            Dm_fd.col(6 * b + k) = (fp - fm) / eps;
            bodies[b]->_phi = phi0;
        }
    }
    print_error("ForceGeneralSDFContact: Km", Km, Km_fd);
    print_error("ForceGeneralSDFContact: Dm", Dm, Dm_fd);
}

void ForceGeneralSDFContact::test_design_derivatives_runtime() {
    std::vector<Contact> contacts;
    const std::vector<Vector3> pts = _contact_body->get_contact_points();
    const Matrix3 R1 = _contact_body->_E_0i.topLeftCorner(3, 3);
    const Vector3 p1 = _contact_body->_E_0i.topRightCorner(3, 1);
    for (size_t i = 0; i < pts.size(); ++i) {
        const Vector3 xw = R1 * pts[i] + p1;
        const dtype d = _SDF_body->distance(xw);
        if (d < 0.) contacts.push_back(Contact(pts[i], xw, d, Vector3::Zero(), (int)i));
    }
    if (contacts.empty()) return;
    // d fm / d contact_scale against a forward difference in the contact body's _contact_scale
    const dtype eps = 1e-7, cs0 = _contact_body->_contact_scale;
    double f0[12], f1[12], K[144], D[144], p6[12];
    evaluate_list(this, contacts, ATAC_EVAL_DERIV | ATAC_EVAL_P6, f0, K, D, p6);
    _contact_body->_contact_scale = cs0 + eps;
    evaluate_list(this, contacts, ATAC_EVAL_FORCE, f1, nullptr, nullptr, nullptr);
    _contact_body->_contact_scale = cs0;
    VectorX an = Eigen::Map<VectorX>(p6, 12), fd(12);
    for (int k = 0; k < 12; ++k) fd(k) = (f1[k] - f0[k]) / eps;
    print_error("ForceGeneralSDFContact: dfm_dp6", an, fd);
}

void ForceGeneralSDFContact::test_derivatives() {
    // the runtime check at random twists of both bodies (their poses as they are)
    srand(1000);
    const Vector6 phi1 = _contact_body->_phi, phi2 = _SDF_body->_phi;
    _contact_body->_phi = Vector6::Random();
    _SDF_body->_phi = Vector6::Random();
    test_derivatives_runtime();
    _contact_body->_phi = phi1;
    _SDF_body->_phi = phi2;
}

}  // namespace redmax
