// GPU-backed ForceGeneralSDFContact (interface of the reference's
// simulation/redmax/Force/ForceGeneralSDFContact.cpp). The public wrappers keep the
// reference's contract: detection of the contact body's points against the SDF
// body, then the wrench (and Km, Dm, the contact-scale design column) accumulated
// with += into the global maximal-coordinate arrays at each body's _index
// (ForceGeneralSDFContact.cpp:46-144). The arithmetic is done by libatac_contact
// through the C ABI (include/atac.h); this file only moves data.
#include "Force/ForceGeneralSDFContact.h"

#include <chrono>
#include <cstring>
#include <random>
#include <string>

#include "Body/Body.h"
#include "Body/BodySDFObj.h"
#include "atac.h"

namespace redmax {

namespace {
void check(int rc, const char* what) {
    if (rc != ATAC_OK) throw_error(std::string("ForceGeneralSDFContact: ") + what + ": " + atac_last_error());
}

long long now_us() {
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// column of the contact-scale design parameter of the contact body in dfm_dp,
// -1 when that parameter is not active
int contact_scale_column(const Body* b) { return b->_p6_column; }
}  // namespace

struct ForceGeneralSDFContact::Device {
    atac_points points = nullptr;
    atac_batch batch = nullptr;
    atac_grid grid = nullptr;
    size_t n_points = 0;
    ~Device() {
        if (batch) atac_batch_destroy(batch);
        if (points) atac_points_destroy(points);
    }
};

ForceGeneralSDFContact::ForceGeneralSDFContact(Simulation* sim, Body* contact_body, Body* SDF_body, dtype kn, dtype kt,
                                               dtype mu, dtype damping)
    : Force(sim), _dev(new Device()) {
    _contact_body = contact_body;
    _SDF_body = dynamic_cast<BodySDFObj*>(SDF_body);
    if (_SDF_body == nullptr) throw_error("The second body in Contact should be SDF body.");
    _kn = kn;
    _kt = kt;
    _mu = mu;
    _damping = damping;
    _scale = 1.;
}

ForceGeneralSDFContact::~ForceGeneralSDFContact() = default;

void ForceGeneralSDFContact::set_stiffness(dtype kn, dtype kt) { _kn = kn; _kt = kt; }
void ForceGeneralSDFContact::set_friction(dtype mu) { _mu = mu; }
void ForceGeneralSDFContact::set_damping(dtype damping) { _damping = damping; }
void ForceGeneralSDFContact::set_scale(dtype scale) { _scale = scale; }

void ForceGeneralSDFContact::fill_instance(void* p) const {
    atac_instance& in = *static_cast<atac_instance*>(p);
    std::memset(&in, 0, sizeof in);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            in.E1[4 * r + c] = _contact_body->_E_0i(r, c);
            in.E2[4 * r + c] = _SDF_body->_E_0i(r, c);
        }
    for (int k = 0; k < 6; ++k) {
        in.phi1[k] = _contact_body->_phi(k);
        in.phi2[k] = _SDF_body->_phi(k);
    }
    in.col_th = _SDF_body->_col_th;
    in.kn = _kn;
    in.kt = _kt;
    in.mu = _mu;
    in.damping = _damping;
    in.scale = _scale;
    in.contact_scale = _contact_body->_contact_scale;
}

// detection + contact over all points of the contact body (points uploaded once)
void ForceGeneralSDFContact::evaluate(uint32_t flags, double* fm, double* Km, double* Dm, double* p6) {
    Device& D = *_dev;
    atac_grid grid = _SDF_body->device_grid();
    if (!D.points || grid != D.grid) {
        std::vector<Vector3> pts = _contact_body->get_contact_points();
        std::vector<double> xyz(3 * pts.size());
        for (size_t i = 0; i < pts.size(); ++i)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = pts[i](k);
        if (D.batch) { atac_batch_destroy(D.batch); D.batch = nullptr; }
        if (D.points) { atac_points_destroy(D.points); D.points = nullptr; }
        check(atac_points_create(xyz.data(), (int64_t)pts.size(), nullptr, 1, 0.0, &D.points), "points");
        D.n_points = pts.size();
        D.grid = grid;
        atac_instance in;
        fill_instance(&in);
        check(atac_batch_create(&grid, 1, &D.points, 1, &in, 1, nullptr, &D.batch), "batch");
        check(atac_batch_timing(D.batch, 1, nullptr, nullptr), "timing");   // stage times for _time
    }
    atac_instance in;
    fill_instance(&in);
    check(atac_batch_set_instances(D.batch, &in, 1, nullptr), "set_instances");
    int32_t n_active = 0;
    check(atac_batch_eval_host(D.batch, flags, fm, Km, Dm, p6, &n_active), "eval");
    _last_active = n_active;
    double ms[6];
    int32_t launches[6];
    check(atac_batch_timing(D.batch, 1, ms, launches), "timing");
    _last_detect_us = (long long)((ms[0] + ms[1]) * 1e3);   // detection and the slow-path samples
    if (n_active > 0) {   // CollisionDetection.cpp:143-146
        _contact_body->add_contact_body(_SDF_body->_name);
        _SDF_body->add_contact_body(_contact_body->_name);
    }
}

// the same over a fixed contact list: every listed point is a contact
void ForceGeneralSDFContact::evaluate_list(const std::vector<Contact>& contacts, uint32_t flags, double* fm, double* Km,
                                           double* Dm, double* p6) {
    std::memset(fm, 0, 12 * sizeof(double));
    if (Km) std::memset(Km, 0, 144 * sizeof(double));
    if (Dm) std::memset(Dm, 0, 144 * sizeof(double));
    if (p6) std::memset(p6, 0, 12 * sizeof(double));
    if (contacts.empty()) return;
    std::vector<double> xyz(3 * contacts.size());
    for (size_t i = 0; i < contacts.size(); ++i)
        for (int k = 0; k < 3; ++k) xyz[3 * i + k] = contacts[i]._xi(k);
    atac_points pts = nullptr;
    atac_batch batch = nullptr;
    atac_grid grid = _SDF_body->device_grid();
    atac_instance in;
    fill_instance(&in);
    int rc = atac_points_create(xyz.data(), (int64_t)contacts.size(), nullptr, 1, 0.0, &pts);
    if (rc == ATAC_OK) rc = atac_batch_create(&grid, 1, &pts, 1, &in, 1, nullptr, &batch);
    if (rc == ATAC_OK) rc = atac_batch_eval_host(batch, flags | ATAC_EVAL_ALL_ACTIVE, fm, Km, Dm, p6, nullptr);
    const std::string err = rc == ATAC_OK ? "" : atac_last_error();
    if (batch) atac_batch_destroy(batch);
    if (pts) atac_points_destroy(pts);
    if (rc != ATAC_OK) throw_error("ForceGeneralSDFContact: fixed contact list: " + err);
}

static void scatter_wrench(const double* sub, VectorX& fm, const Body* a, const Body* b) {
    for (int k = 0; k < 6; ++k) {
        fm(a->_index[k]) += sub[k];
        fm(b->_index[k]) += sub[6 + k];
    }
}

// the four 6x6 blocks of a 12x12 row-major block into the global matrix (ForceGeneralSDFContact.cpp:91-99)
static void scatter_blocks(const double* sub, MatrixX& M, const Body* a, const Body* b) {
    const Body* bodies[2] = {a, b};
    for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c)
                    M(bodies[bi]->_index[r], bodies[bj]->_index[c]) += sub[12 * (6 * bi + r) + 6 * bj + c];
}

void ForceGeneralSDFContact::computeForce(VectorX& fm, VectorX& fr, bool verbose) {
    const long long t0 = now_us();
    double sub[12];
    evaluate(ATAC_EVAL_FORCE, sub, nullptr, nullptr, nullptr);
    scatter_wrench(sub, fm, _contact_body, _SDF_body);
    const long long t1 = now_us();
    if (verbose)
        std::cerr << "General SDF Contact (GPU): " << _last_active << " contacts, contact body = " << _contact_body->_name
                  << ", SDF body = " << _SDF_body->_name << std::endl;
    _time.add("collision detection", _last_detect_us);
    _time.add("compute force", t1 - t0 - _last_detect_us);
}

void ForceGeneralSDFContact::computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr,
                                                        MatrixX& Dr, bool verbose) {
    const long long t0 = now_us();
    double sub[12], K[144], D[144];
    evaluate(ATAC_EVAL_DERIV, sub, K, D, nullptr);
    scatter_wrench(sub, fm, _contact_body, _SDF_body);
    scatter_blocks(K, Km, _contact_body, _SDF_body);
    scatter_blocks(D, Dm, _contact_body, _SDF_body);
    const long long t1 = now_us();
    if (verbose)
        std::cerr << "General SDF Contact (GPU): " << _last_active << " contacts" << std::endl;
    _time.add("collision detection", _last_detect_us);
    _time.add("compute force", t1 - t0 - _last_detect_us);
}

void ForceGeneralSDFContact::computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr,
                                                        MatrixX& Dr, MatrixX& dfm_dp, MatrixX& dfr_dp, bool verbose) {
    const long long t0 = now_us();
    double sub[12], K[144], D[144], p6[12];
    evaluate(ATAC_EVAL_DERIV | ATAC_EVAL_P6, sub, K, D, p6);
    scatter_wrench(sub, fm, _contact_body, _SDF_body);
    scatter_blocks(K, Km, _contact_body, _SDF_body);
    scatter_blocks(D, Dm, _contact_body, _SDF_body);
    const int col = contact_scale_column(_contact_body);
    if (col >= 0)
        for (int k = 0; k < 6; ++k) {
            dfm_dp(_contact_body->_index[k], col) += p6[k];
            dfm_dp(_SDF_body->_index[k], col) += p6[6 + k];
        }
    const long long t1 = now_us();
    if (verbose) std::cerr << "General SDF Contact (GPU): " << _last_active << " contacts" << std::endl;
    _time.add("collision detection", _last_detect_us);
    _time.add("compute force", t1 - t0 - _last_detect_us);
}

void ForceGeneralSDFContact::computeForce(std::vector<Contact>& contacts, VectorX& fm, bool verbose) {
    fm = VectorX::Zero(12);
    evaluate_list(contacts, ATAC_EVAL_PRECISE, fm.data(), nullptr, nullptr, nullptr);
}

static void row_major_to(const double* src, MatrixX& M) {
    M = MatrixX::Zero(12, 12);
    for (int r = 0; r < 12; ++r)
        for (int c = 0; c < 12; ++c) M(r, c) = src[12 * r + c];
}

void ForceGeneralSDFContact::computeForceWithDerivative(std::vector<Contact>& contacts, VectorX& fm, MatrixX& Km,
                                                        MatrixX& Dm, bool verbose) {
    double K[144], D[144];
    fm = VectorX::Zero(12);
    evaluate_list(contacts, ATAC_EVAL_PRECISE | ATAC_EVAL_DERIV, fm.data(), K, D, nullptr);
    row_major_to(K, Km);
    row_major_to(D, Dm);
}

void ForceGeneralSDFContact::computeForceWithDerivative(std::vector<Contact>& contacts, VectorX& fm, MatrixX& Km,
                                                        MatrixX& Dm, MatrixX& dfm_dp, bool verbose) {
    double K[144], D[144], p6[12];
    fm = VectorX::Zero(12);
    evaluate_list(contacts, ATAC_EVAL_PRECISE | ATAC_EVAL_DERIV | ATAC_EVAL_P6, fm.data(), K, D, p6);
    row_major_to(K, Km);
    row_major_to(D, Dm);
    const int col = contact_scale_column(_contact_body);
    if (col >= 0)
        for (int k = 0; k < 6; ++k) {
            dfm_dp(_contact_body->_index[k], col) += p6[k];
            dfm_dp(_SDF_body->_index[k], col) += p6[6 + k];
        }
}

// ---------------------------------------------------------------- self-checks
static Matrix4 right_mul_exp(const Matrix4& E, int k, dtype eps) {
    Vector6 dq;
    dq(k) = eps;
    return math::mul(E, math::exp(dq));
}

static MatrixX as_matrix(const VectorX& v) {
    MatrixX m(v.size(), 1);
    for (int i = 0; i < v.size(); ++i) m(i, 0) = v(i);
    return m;
}

// Km, Dm of the fixed contact list against forward differences (eps = 1e-7) with
// E <- E exp(eps e_k) right-multiplied and phi <- phi + eps e_k
// (ForceGeneralSDFContact.cpp:1062-1110)
static dtype fd_check(ForceGeneralSDFContact& f, std::vector<Contact>& contacts,
                      void (ForceGeneralSDFContact::*force)(std::vector<Contact>&, VectorX&, bool),
                      void (ForceGeneralSDFContact::*deriv)(std::vector<Contact>&, VectorX&, MatrixX&, MatrixX&, bool)) {
    const dtype eps = 1e-7;
    VectorX fm;
    MatrixX Km, Dm;
    (f.*deriv)(contacts, fm, Km, Dm, false);
    MatrixX Kfd(12, 12), Dfd(12, 12);
    Body* bodies[2] = {f._contact_body, (Body*)f._SDF_body};
    for (int b = 0; b < 2; ++b) {
        const Matrix4 E = bodies[b]->_E_0i;
        const Vector6 phi = bodies[b]->_phi;
        for (int k = 0; k < 6; ++k) {
            VectorX fp;
            bodies[b]->_E_0i = right_mul_exp(E, k, eps);
            (f.*force)(contacts, fp, false);
            bodies[b]->_E_0i = E;
            for (int r = 0; r < 12; ++r) Kfd(r, 6 * b + k) = (fp(r) - fm(r)) / eps;
            bodies[b]->_phi(k) = phi(k) + eps;
            (f.*force)(contacts, fp, false);
            bodies[b]->_phi = phi;
            for (int r = 0; r < 12; ++r) Dfd(r, 6 * b + k) = (fp(r) - fm(r)) / eps;
        }
    }
    dtype worst = print_error("Km ", Km, Kfd);
    worst = std::max(worst, print_error("Dm ", Dm, Dfd));
    (void)as_matrix;
    return worst;
}

dtype ForceGeneralSDFContact::test_derivatives() {
    std::cerr << "**************************** General-SDF Collision Derivatives ***************************" << std::endl;
    // one synthetic contact in the SDF body's box, random poses and twists (seed 1000)
    std::mt19937_64 rng(1000);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const Matrix4 E1 = _contact_body->_E_0i, E2 = _SDF_body->_E_0i;
    const Vector6 phi1 = _contact_body->_phi, phi2 = _SDF_body->_phi;
    Vector6 q1, q2;
    for (int k = 0; k < 6; ++k) { q1(k) = 0.1 * U(rng); q2(k) = 0.1 * U(rng); _contact_body->_phi(k) = U(rng); _SDF_body->_phi(k) = U(rng); }
    _contact_body->_E_0i = math::mul(E1, math::exp(q1));
    _SDF_body->_E_0i = math::mul(E2, math::exp(q2));
    std::vector<Contact> contacts;
    for (int i = 0; i < 8; ++i) {
        Vector3 x;   // a point inside the grid box, in the SDF body frame, mapped to body 1
        for (int k = 0; k < 3; ++k) {
            const dtype t = 0.5 + 0.4 * U(rng);
            x(k) = _SDF_body->_min_box(k) + t * (_SDF_body->_max_box(k) - _SDF_body->_min_box(k));
        }
        Vector3 xw, xi;
        for (int r = 0; r < 3; ++r) {
            xw(r) = _SDF_body->_E_0i(r, 3);
            for (int c = 0; c < 3; ++c) xw(r) += _SDF_body->_E_0i(r, c) * x(c);
        }
        for (int r = 0; r < 3; ++r) {
            xi(r) = 0;
            for (int c = 0; c < 3; ++c) xi(r) += _contact_body->_E_0i(c, r) * (xw(c) - _contact_body->_E_0i(c, 3));
        }
        contacts.push_back(Contact(xi, xw, 0.0, Vector3(), i));
    }
    void (ForceGeneralSDFContact::*force)(std::vector<Contact>&, VectorX&, bool) = &ForceGeneralSDFContact::computeForce;
    void (ForceGeneralSDFContact::*deriv)(std::vector<Contact>&, VectorX&, MatrixX&, MatrixX&, bool) =
        &ForceGeneralSDFContact::computeForceWithDerivative;
    const dtype worst = fd_check(*this, contacts, force, deriv);
    _contact_body->_E_0i = E1; _SDF_body->_E_0i = E2;
    _contact_body->_phi = phi1; _SDF_body->_phi = phi2;
    return worst;
}

std::vector<Contact> ForceGeneralSDFContact::active_contacts() {
    double sub[12];
    evaluate(ATAC_EVAL_CHECKPOINT, sub, nullptr, nullptr, nullptr);
    int64_t n = 0;
    check(atac_batch_active_ids(_dev->batch, 0, nullptr, 0, &n), "active ids");
    std::vector<int64_t> ids((size_t)n);
    if (n) check(atac_batch_active_ids(_dev->batch, 0, ids.data(), n, &n), "active ids");
    std::vector<Vector3> pts = _contact_body->get_contact_points();
    std::vector<Contact> contacts;
    for (int64_t id : ids) contacts.push_back(Contact(pts[(size_t)id], Vector3(), 0.0, Vector3(), (int)id));
    return contacts;
}

dtype ForceGeneralSDFContact::test_derivatives_runtime_error() {
    std::vector<Contact> contacts = active_contacts();
    if (contacts.empty()) return 0.0;
    void (ForceGeneralSDFContact::*force)(std::vector<Contact>&, VectorX&, bool) = &ForceGeneralSDFContact::computeForce;
    void (ForceGeneralSDFContact::*deriv)(std::vector<Contact>&, VectorX&, MatrixX&, MatrixX&, bool) =
        &ForceGeneralSDFContact::computeForceWithDerivative;
    return fd_check(*this, contacts, force, deriv);
}

void ForceGeneralSDFContact::test_derivatives_runtime() { (void)test_derivatives_runtime_error(); }

// d fm / d contact_scale against a forward difference (ForceGeneralSDFContact.cpp:1192-1243, p6 column)
void ForceGeneralSDFContact::test_design_derivatives_runtime() {
    std::vector<Contact> contacts = active_contacts();
    if (contacts.empty()) return;
    const dtype eps = 1e-7;
    VectorX fm, fp;
    MatrixX Km, Dm, P(12, 1), Pfd(12, 1);
    double K[144], D[144], p6[12];
    fm = VectorX::Zero(12);
    evaluate_list(contacts, ATAC_EVAL_PRECISE | ATAC_EVAL_DERIV | ATAC_EVAL_P6, fm.data(), K, D, p6);
    const dtype cs = _contact_body->_contact_scale;
    _contact_body->_contact_scale = cs + eps;
    computeForce(contacts, fp, false);
    _contact_body->_contact_scale = cs;
    for (int r = 0; r < 12; ++r) {
        P(r, 0) = p6[r];
        Pfd(r, 0) = (fp(r) - fm(r)) / eps;
    }
    print_error("dfm_dp6 ", P, Pfd);
}

}  // namespace redmax
