#include "Simulation/Simulation.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

#include "Body/ContactPoints.h"
#include "Simulation/SceneXml.h"

namespace redmax {

namespace {

constexpr dtype kNewtonTol = 1e-8;      // Simulation.cpp:1007
constexpr int kMaxIterLS = 20;          // Simulation.cpp:1011
constexpr int kMaxFailStrike = 3;       // Simulation.cpp:1013
constexpr size_t kMaxCachedBatches = 16;

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// attribute readers with pugixml's conversions (as_float rounds to float, as_bool
// looks at the first character, vectors are read as doubles: str_to_eigen,
// Simulation_Constructor.cpp:40-54)
double as_float(const std::string& s) { return (double)(float)std::strtod(s.c_str(), nullptr); }
int as_int(const std::string& s) { return (int)std::strtol(s.c_str(), nullptr, 0); }
bool as_bool(const std::string& s) {
    return !s.empty() && (s[0] == '1' || s[0] == 't' || s[0] == 'T' || s[0] == 'y' || s[0] == 'Y');
}
std::vector<double> as_vector(const std::string& s) {
    std::istringstream in(s);
    std::vector<double> v;
    double x;
    while (in >> x) v.push_back(x);
    return v;
}
Vector3 as_vector3(const std::string& s, const char* what) {
    const auto v = as_vector(s);
    if (v.size() != 3) throw_error(std::string("scene XML: ") + what + " needs 3 numbers: '" + s + "'");
    return make_vector3(v[0], v[1], v[2]);
}
// math::quat2mat (Utils.h:232-240): the rotation of the normalised quaternion (w, x, y, z)
Matrix3 quat2mat(const std::string& s) {
    auto q = as_vector(s);
    if (q.size() != 4) throw_error("scene XML: quat needs 4 numbers: '" + s + "'");
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(n > 0.0)) throw_error("scene XML: zero quaternion");
    const double w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
    Matrix3 R;
    R(0, 0) = 1 - 2 * (y * y + z * z); R(0, 1) = 2 * (x * y - z * w);     R(0, 2) = 2 * (x * z + y * w);
    R(1, 0) = 2 * (x * y + z * w);     R(1, 1) = 1 - 2 * (x * x + z * z); R(1, 2) = 2 * (y * z - x * w);
    R(2, 0) = 2 * (x * z - y * w);     R(2, 1) = 2 * (y * z + x * w);     R(2, 2) = 1 - 2 * (x * x + y * y);
    return R;
}

Matrix4 mul4(const Matrix4& A, const Matrix4& B) {
    Matrix4 C;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += A(i, k) * B(k, j);
            C(i, j) = s;
        }
    return C;
}
Matrix4 einv(const Matrix4& E) {
    Matrix4 F = Matrix4::Identity();
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) F(i, j) = E(j, i);
        F(i, 3) = -(E(0, i) * E(0, 3) + E(1, i) * E(1, 3) + E(2, i) * E(2, 3));
    }
    return F;
}
// Ad(E) for twists [w; v]: [[R, 0], [[p]x R, R]], row-major 6 x 6
void adjoint(const Matrix4& E, double A[36]) {
    for (int e = 0; e < 36; ++e) A[e] = 0.0;
    const double p[3] = {E(0, 3), E(1, 3), E(2, 3)};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[6 * i + j] = E(i, j);
            A[6 * (3 + i) + 3 + j] = E(i, j);
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;   // ([p]x R)(i, j) = p[i1] R(i2, j) - p[i2] R(i1, j)
            A[6 * (3 + i) + j] = p[i1] * E(i2, j) - p[i2] * E(i1, j);
        }
}
void skew(const double v[3], double S[9]) {
    S[0] = 0;     S[1] = -v[2]; S[2] = v[1];
    S[3] = v[2];  S[4] = 0;     S[5] = -v[0];
    S[6] = -v[1]; S[7] = v[0];  S[8] = 0;
}
void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
void to_row_major(const Matrix4& E, double out[16]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = E(i, j);
}

// A body's joint Jacobian blocks (Robot::computeJointJacobianWithDerivative,
// Robot.cpp:656-760, for a joint attached to the world): J = Ad(E_ij) S_j, its time
// derivative and both partial derivatives in q. Row-major 6 x n.
struct BodyJac {
    int n = 0;
    std::vector<double> J, Jd;
    std::vector<std::vector<double>> dJ, dJd;
    std::vector<char> dzero;   // per q_k: dJ[k] and dJd[k] are zero (S_j does not depend on q_k)
};
bool all_zero(const MatrixX& S) {
    for (int i = 0; i < S.size(); ++i)
        if (S.data()[i] != 0.0) return false;
    return true;
}
BodyJac body_jacobian(const Simulation::SimBody& b, bool derivs) {
    BodyJac o;
    const Joint& j = *b.joint;
    o.n = j._ndof;
    double A[36];
    adjoint(b.body->_E_ij, A);
    auto apply = [&](const MatrixX& S, std::vector<double>& out) {
        out.assign(6 * o.n, 0.0);
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < o.n; ++c) {
                double s = 0.0;
                for (int k = 0; k < 6; ++k) s += A[6 * r + k] * S(k, c);
                out[o.n * r + c] = s;
            }
    };
    apply(j._S_j, o.J);
    apply(j._S_j_dot, o.Jd);
    if (derivs) {
        o.dJ.resize(o.n);
        o.dJd.resize(o.n);
        o.dzero.assign(o.n, 0);
        for (int k = 0; k < o.n; ++k) {
            if (all_zero(j._dSj_dq[k]) && all_zero(j._dSjdot_dq[k])) {
                o.dJ[k].assign(6 * o.n, 0.0);
                o.dJd[k].assign(6 * o.n, 0.0);
                o.dzero[k] = 1;
                continue;
            }
            apply(j._dSj_dq[k], o.dJ[k]);
            apply(j._dSjdot_dq[k], o.dJd[k]);
        }
    }
    return o;
}

void check_rc(int rc, const char* what) {
    if (rc != ATAC_OK) throw_error(std::string("Simulation: ") + what + ": " + atac_last_error());
}

// x = A^-1 b, A n x n (column-major MatrixX): LU with partial pivoting
// (the reference's partialPivLu().solve)
VectorX lu_solve(const MatrixX& A_in, const VectorX& b_in) {
    const int n = A_in.rows();
    std::vector<double> A(A_in.data(), A_in.data() + (size_t)n * n);
    auto a = [&](int r, int c) -> double& { return A[(size_t)r + (size_t)n * c]; };
    VectorX b = b_in;
    std::vector<std::pair<int, double>> rows;
    rows.reserve(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(a(r, k)) > std::fabs(a(p, k))) p = r;
        if (p != k) {
            for (int c = 0; c < n; ++c) std::swap(a(k, c), a(p, c));
            std::swap(b(k), b(p));
        }
        // the rows below with a non-zero in column k and their multipliers, then the
        // update column by column over the pivot row's non-zeros (the bodies' blocks
        // leave most of both zero; each entry gets the same single update as a plain
        // row-by-row elimination gives it)
        const double piv = a(k, k);
        rows.clear();
        for (int r = k + 1; r < n; ++r) {
            const double ark = a(r, k);
            if (ark == 0.0) continue;
            const double l = ark / piv;
            rows.push_back({r, l});
            b(r) -= l * b(k);
        }
        if (rows.empty()) continue;
        for (int c = k + 1; c < n; ++c) {
            const double akc = a(k, c);
            if (akc == 0.0) continue;
            for (const auto& rl : rows) a(rl.first, c) -= rl.second * akc;
        }
    }
    VectorX x(n);
    for (int k = n - 1; k >= 0; --k) {
        double s = b(k);
        for (int c = k + 1; c < n; ++c) s -= a(k, c) * x(c);
        x(k) = s / a(k, k);
    }
    return x;
}

double norm(const VectorX& v) {
    double s = 0.0;
    for (int i = 0; i < v.size(); ++i) s += v(i) * v(i);
    return std::sqrt(s);
}

}  // namespace

// ------------------------------------------------------------------ construction

Simulation::Simulation(const std::string& xml_string, const std::string& asset_folder, bool verbose, bool precise)
    : _options(new Options()), _asset_folder(asset_folder), _verbose(verbose), _precise(precise) {
    parse(xml_string);
    build_contact_data();
    init();
}

std::unique_ptr<Simulation> Simulation::from_file(const std::string& xml_file_path, bool verbose) {
    std::ifstream in(xml_file_path.c_str());
    if (!in.good()) throw_error("Simulation: cannot read " + xml_file_path);
    std::stringstream ss;
    ss << in.rdbuf();
    const size_t slash = xml_file_path.find_last_of('/');
    const std::string folder = slash == std::string::npos ? "." : xml_file_path.substr(0, slash);
    return std::unique_ptr<Simulation>(new Simulation(ss.str(), folder, verbose));
}

Simulation::~Simulation() {
    for (auto& kv : _batches)
        if (kv.second.batch) atac_batch_destroy(kv.second.batch);
    if (_points) atac_points_destroy(_points);
    delete _options;
}

void Simulation::parse(const std::string& xml_string) {
    const xml::Node doc = xml::parse(xml_string);
    const xml::Node* root = doc.child("redmax");
    if (!root) throw_error("Input xml string is incorrect.");
    if (const auto* m = root->attr("model")) _name = *m;
    if (const xml::Node* opt = root->child("option")) {
        if (const auto* a = opt->attr("gravity")) _options->_gravity = as_vector3(*a, "gravity");
        if (const auto* a = opt->attr("timestep")) _options->_h = as_float(*a);
        if (const auto* a = opt->attr("integrator")) _options->_integrator = *a;
        if (const auto* a = opt->attr("unit")) _options->_unit = *a;
    }
    if (root->child("ground")) throw_error("Simulation: <ground> is not in this build (SDF contact scenes only)");
    const xml::Node* dflt = root->child("default");
    auto default_attr = [&](const std::string& elem, const std::string& key) -> const std::string* {
        const xml::Node* d = dflt ? dflt->child(elem) : nullptr;
        return d ? d->attr(key) : nullptr;
    };
    for (const auto& robot : root->children) {
        if (robot.name != "robot") continue;
        for (const auto& link : robot.children) {
            if (link.name != "link") continue;
            if (link.child("link"))
                throw_error("Simulation: nested links are not in this build (each body on its own joint to the world)");
            // joint (Simulation_Constructor.cpp:406-473)
            const xml::Node* jn = link.child("joint");
            if (!jn) throw_error("Simulation: a link without a joint");
            const std::string* jt = jn->attr("type");
            const std::string type = jt ? *jt : "";
            Joint::Type t;
            if (type == "fixed") t = Joint::Type::Fixed;
            else if (type == "translational") t = Joint::Type::Translational;
            else if (type == "free3d-exp") t = Joint::Type::Free3DExp;
            else throw_error("Joint type error: " + type + " (this build has fixed, translational and free3d-exp)");
            if (const auto* fr = jn->attr("frame"))
                if (*fr != "LOCAL" && *fr != "WORLD") throw_error("Frame type error: " + *fr);
            const Vector3 jpos = as_vector3(jn->attr("pos") ? *jn->attr("pos") : std::string("0 0 0"), "joint pos");
            const Matrix3 jR = quat2mat(jn->attr("quat") ? *jn->attr("quat") : std::string("1 0 0 0"));
            std::unique_ptr<Joint> joint(new Joint(jn->attr("name") ? *jn->attr("name") : "", (int)_joints.size(), t, jR, jpos));
            if (const auto* d = jn->attr("damping")) joint->_Dr = as_float(*d);
            else if (const auto* d2 = default_attr("joint", "damping")) joint->_Dr = as_float(*d2);
            if (const auto* lim = jn->attr("lim")) {
                const auto l = as_vector(*lim);
                if (l.size() != 2) throw_error("scene XML: lim needs 2 numbers");
                joint->_joint_limit_lower = l[0];
                joint->_joint_limit_upper = l[1];
                if (const auto* s = jn->attr("lim_stiffness")) joint->_joint_limit_stiffness = as_float(*s);
                else if (const auto* s2 = default_attr("joint", "lim_stiffness")) joint->_joint_limit_stiffness = as_float(*s2);
            }
            // body (Simulation_Constructor.cpp:498-590)
            const xml::Node* bn = link.child("body");
            if (!bn) throw_error("Simulation: a link without a body");
            const std::string btype = bn->attr("type") ? *bn->attr("type") : "";
            if (btype != "mesh" && btype != "SDF")
                throw_error("Simulation: body type " + btype + " is not in this build (mesh and SDF bodies)");
            const Vector3 bpos = as_vector3(bn->attr("pos") ? *bn->attr("pos") : std::string("0 0 0"), "body pos");
            const Matrix3 bR = quat2mat(bn->attr("quat") ? *bn->attr("quat") : std::string("1 0 0 0"));
            dtype density = 1.0;
            if (const auto* d = bn->attr("density")) density = as_float(*d);
            else if (const auto* d2 = default_attr("body", "density")) density = as_float(*d2);
            std::string ttype = "BODY_TO_JOINT";
            if (const auto* tt = bn->attr("transform_type")) {
                ttype = *tt;
                if (ttype != "OBJ_TO_JOINT" && ttype != "OBJ_TO_WORLD" && ttype != "BODY_TO_JOINT")
                    throw_error("Transform type error: " + ttype);
            }
            const Vector3 scale = bn->attr("scale") ? as_vector3(*bn->attr("scale"), "scale") : Vector3::Ones();
            const bool adaptive = bn->attr("adaptive_sample") ? as_bool(*bn->attr("adaptive_sample")) : false;
            std::string filename = bn->attr("filename") ? *bn->attr("filename") : "";
            if (filename.find(_asset_folder) == std::string::npos) filename = _asset_folder + "//" + filename;

            SimBody b;
            b.name = bn->attr("name") ? *bn->attr("name") : "";
            b.joint = joint.get();
            const BodyMeshObj::TransformType tt = ttype == "BODY_TO_JOINT" ? BodyMeshObj::BODY_TO_JOINT
                                                  : ttype == "OBJ_TO_WORLD" ? BodyMeshObj::OBJ_TO_WOLRD
                                                                            : BodyMeshObj::OBJ_TO_JOINT;
            if (btype == "SDF") {
                dtype dx = 0.05, col_th = 0.01;
                int res = 20;
                if (const auto* a = bn->attr("dx")) dx = as_float(*a);
                if (const auto* a = bn->attr("res")) res = as_int(*a);
                if (const auto* a = bn->attr("col_th")) col_th = as_float(*a);
                const bool load_sdf = bn->attr("load_sdf") ? as_bool(*bn->attr("load_sdf")) : false;
                const bool save_sdf = bn->attr("save_sdf") ? as_bool(*bn->attr("save_sdf")) : false;
                auto* s = new BodySDFObj(this, joint.get(), filename, bR, bpos, dx, res, col_th, tt, density, scale,
                                         adaptive, load_sdf, save_sdf);
                b.body.reset(s);
                b.sdf = s;
            } else {
                b.body.reset(new BodyMeshObj(this, joint.get(), filename, bR, bpos, tt, density, scale, adaptive));
            }
            b.body->_name = b.name;
            const int index0 = 6 * (int)_bodies.size();
            for (int k = 0; k < 6; ++k) b.body->_index[k] = index0 + k;
            for (int a = 0; a < 3; ++a) {
                b.pts_lo(a) = INFINITY;
                b.pts_hi(a) = -INFINITY;
            }
            for (const auto& v : b.body->_contact_points)
                for (int a = 0; a < 3; ++a) {
                    b.pts_lo(a) = std::min(b.pts_lo(a), v(a));
                    b.pts_hi(a) = std::max(b.pts_hi(a), v(a));
                }
            _bodies.push_back(std::move(b));
            _joints.push_back(std::move(joint));
        }
    }
    if (root->child("actuator")) throw_error("Simulation: actuators are not in this build");
    // contact (Simulation_Constructor.cpp:251-330)
    for (const auto& contact : root->children) {
        if (contact.name != "contact") continue;
        for (const auto& c : contact.children) {
            if (c.name != "general_SDF_contact")
                throw_error("Simulation: contact type " + c.name + " is not in this build (general_SDF_contact)");
            auto param = [&](const char* key) -> dtype {
                if (const auto* a = c.attr(key)) return as_float(*a);
                if (const auto* a = default_attr(c.name, key)) return as_float(*a);
                return 0.0;
            };
            const std::string g = c.attr("general_body") ? *c.attr("general_body") : "";
            const std::string s = c.attr("SDF_body") ? *c.attr("SDF_body") : "";
            const int gi = body_index(g), si = body_index(s);
            if (gi < 0) throw_error("General contact body name error: " + g);
            if (si < 0) throw_error("SDF contact body name error: " + s);
            if (!_bodies[si].sdf) throw_error("ForceGeneralSDFContact: SDF_body " + s + " is not an SDF body");
            _contacts.push_back({gi, si, param("kn"), param("kt"), param("mu"), param("damping"), 1.0});
        }
    }
}

void Simulation::build_contact_data() {
    std::vector<double> xyz;
    std::vector<int64_t> seg{0};
    for (const auto& b : _bodies) {
        for (const auto& v : b.body->_contact_points)
            for (int a = 0; a < 3; ++a) xyz.push_back(v(a));
        seg.push_back((int64_t)(xyz.size() / 3));
    }
    _grid_of_body.assign(_bodies.size(), -1);
    for (size_t i = 0; i < _bodies.size(); ++i)
        if (_bodies[i].sdf) {
            _grid_of_body[i] = (int)_grids.size();
            _grids.push_back(_bodies[i].sdf->device_grid());
        }
    if (!_contacts.empty() && !_bodies.empty())
        check_rc(atac_points_create(xyz.data(), (int64_t)(xyz.size() / 3), seg.data(), (int32_t)_bodies.size(), 0.0,
                                    &_points),
                 "atac_points_create");
}

void Simulation::init() {
    int off = 0;
    for (auto& j : _joints) {
        j->_index.clear();
        for (int k = 0; k < j->_ndof; ++k) j->_index.push_back(off + k);
        off += j->_ndof;
    }
    _ndof_r = off;
    _ndof_m = 6 * (int)_bodies.size();
    update_robot();
    _q_init = get_q();
    _qdot_init = get_qdot();
}

int Simulation::body_index(const std::string& name) const {
    for (size_t i = 0; i < _bodies.size(); ++i)
        if (_bodies[i].name == name) return (int)i;
    return -1;
}

const Joint* Simulation::joint_by_name(const std::string& name) const {
    for (const auto& j : _joints)
        if (j->_name == name) return j.get();
    throw_error("Simulation: no joint named " + name);
    return nullptr;
}

Joint* Simulation::joint_by_name(const std::string& name) {
    return const_cast<Joint*>(static_cast<const Simulation*>(this)->joint_by_name(name));
}

// ------------------------------------------------------------------ state

VectorX Simulation::get_q() const {
    VectorX q(_ndof_r);
    for (const auto& j : _joints)
        for (int k = 0; k < j->_ndof; ++k) q(j->_index[k]) = j->_q(k);
    return q;
}

VectorX Simulation::get_qdot() const {
    VectorX q(_ndof_r);
    for (const auto& j : _joints)
        for (int k = 0; k < j->_ndof; ++k) q(j->_index[k]) = j->_qdot(k);
    return q;
}

void Simulation::set_q(const VectorX& q) {
    if (q.size() != _ndof_r) throw_error("set_q: q.size() != ndof_r");
    for (auto& j : _joints)
        for (int k = 0; k < j->_ndof; ++k) j->_q(k) = q(j->_index[k]);
}

void Simulation::set_qdot(const VectorX& qdot) {
    if (qdot.size() != _ndof_r) throw_error("set_qdot: qdot.size() != ndof_r");
    for (auto& j : _joints)
        for (int k = 0; k < j->_ndof; ++k) j->_qdot(k) = qdot(j->_index[k]);
}

void Simulation::set_state(const VectorX& q, const VectorX& qdot) {
    set_q(q);
    set_qdot(qdot);
}

void Simulation::set_state_init(const VectorX& q, const VectorX& qdot) {
    set_q_init(q);
    set_qdot_init(qdot);
}

void Simulation::set_q_init(const VectorX& q) {
    if (q.size() != _ndof_r) throw_error("set_q_init: q_init.size() != _ndof_r.");
    _q_init = q;
}

void Simulation::set_qdot_init(const VectorX& qdot) {
    if (qdot.size() != _ndof_r) throw_error("set_qdot_init: qdot_init.size() != _ndof_r.");
    _qdot_init = qdot;
}

// Joint::update then Body::update (Joint.cpp:98-140, Body.cpp:138-160): E_0i = E_0j E_ji,
// phi_i = Ad(E_ij) phi_j
void Simulation::update_robot(bool) { update_kinematics(true); }

// the joints (with the q-derivatives of S_j when derivs), then the bodies' poses and twists
void Simulation::update_kinematics(bool derivs) {
    for (auto& j : _joints) j->update(derivs);
    for (auto& b : _bodies) {
        b.body->_E_0i = mul4(b.joint->_E_0j, b.body->_E_ji);
        double A[36];
        adjoint(b.body->_E_ij, A);
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += A[6 * r + k] * b.joint->_phi(k);
            b.body->_phi(r) = s;
        }
    }
}

void Simulation::reparam() {
    for (auto& j : _joints) j->reparam();
}

VectorX Simulation::get_joint_q(const std::string& name) const { return joint_by_name(name)->_q; }
VectorX Simulation::get_joint_qdot(const std::string& name) const { return joint_by_name(name)->_qdot; }

void Simulation::set_joint_q(const std::string& name, const VectorX& q) {
    Joint* j = joint_by_name(name);
    if (q.size() != j->_ndof) throw_error("set_joint_q: size mismatch for joint " + name);
    j->_q = q;
    update_robot();
}

void Simulation::zero_joint_q(const std::string& name) {
    joint_by_name(name)->_q.setZero();
    update_robot();
}

void Simulation::zero_joint_qdot(const std::string& name) {
    joint_by_name(name)->_qdot.setZero();
    update_robot();
}

dtype Simulation::get_body_mass(const std::string& name) const {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    return _bodies[i].body->_mass;
}

Matrix4 Simulation::get_body_E0i(const std::string& name) const {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    return _bodies[i].body->_E_0i;
}

Matrix4 Simulation::get_body_Ei0(const std::string& name) const { return einv(get_body_E0i(name)); }

std::vector<Vector3> Simulation::get_body_vertices(const std::string& name, bool world_frame) const {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    std::vector<Vector3> V = _bodies[i].body->_V;
    if (world_frame) {
        const Matrix4& E = _bodies[i].body->_E_0i;
        for (auto& v : V) {
            const Vector3 x = v;
            for (int r = 0; r < 3; ++r) v(r) = E(r, 0) * x(0) + E(r, 1) * x(1) + E(r, 2) * x(2) + E(r, 3);
        }
    }
    return V;
}

std::vector<std::array<int, 3>> Simulation::get_body_faces(const std::string& name) const {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    return _bodies[i].body->_F;
}

std::vector<Vector3> Simulation::get_body_contact_points(const std::string& name) const {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    return _bodies[i].body->_contact_points;
}

void Simulation::set_body_external_force(const std::string& name, const Vector6& force) {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    _bodies[i].body->set_external_force(force);
}

dtype Simulation::get_body_distance(const std::string& name_from, const std::string& name_to) {
    const int t = body_index(name_to);
    if (t < 0 || body_index(name_from) < 0) throw_error("Simulation: no body named " + (t < 0 ? name_to : name_from));
    if (!_bodies[t].sdf) throw_error("get_body_distance: " + name_to + " is not an SDF body");
    const auto V = get_body_vertices(name_from, true);
    std::vector<double> xw(3 * V.size()), d(V.size());
    for (size_t i = 0; i < V.size(); ++i)
        for (int a = 0; a < 3; ++a) xw[3 * i + a] = V[i](a);
    double E2[16];
    to_row_major(_bodies[t].body->_E_0i, E2);
    check_rc(atac_sdf_distance_host(_bodies[t].sdf->device_grid(), E2, _bodies[t].sdf->_col_th, xw.data(),
                                    (int64_t)V.size(), d.data()),
             "atac_sdf_distance_host");
    double m = INFINITY;
    for (double x : d) m = std::min(m, x);
    return m;
}

bool Simulation::body_in_contact(const std::string& a, const std::string& b, dtype eps) {
    return std::min(get_body_distance(a, b), get_body_distance(b, a)) < eps;
}

std::vector<std::string> Simulation::get_contact_bodies(const std::string& name) const {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    return _bodies[i].body->get_contact_bodies();
}

void Simulation::clear_contact_bodies(const std::string& name) {
    const int i = body_index(name);
    if (i < 0) throw_error("Simulation: no body named " + name);
    _bodies[i].body->clear_contact_bodies();
}

void Simulation::clear_all_contact_bodies() {
    for (auto& b : _bodies) b.body->clear_contact_bodies();
}

void Simulation::clear_all_saved_sdfs() {
    for (const auto& b : _bodies)
        if (b.sdf) b.sdf->clear_saved_SDF();
}

void Simulation::set_contact_scale(dtype scale) {
    for (auto& c : _contacts) c.scale = scale;
}

// ------------------------------------------------------------------ forces

// Body::computeMaximalForce(WithDerivative) (Body.cpp:250-330): Coriolis ad(phi)^T (I phi),
// gravity R^T (m g), the external force scaled by the inertia entries as the reference
// applies it, and their derivatives in the body's pose and twist.
void Simulation::body_forces(bool derivs, VectorX& fm, MatrixX* Km, MatrixX* Dm) const {
    for (size_t bi = 0; bi < _bodies.size(); ++bi) {
        const SimBody& b = _bodies[bi];
        const int o = 6 * (int)bi;
        const Matrix4& E = b.body->_E_0i;
        const double w[3] = {b.body->_phi(0), b.body->_phi(1), b.body->_phi(2)};
        const double v[3] = {b.body->_phi(3), b.body->_phi(4), b.body->_phi(5)};
        double Iw[3], mv[3];
        for (int k = 0; k < 3; ++k) {
            Iw[k] = b.body->_Inertia(k) * w[k];
            mv[k] = b.body->_Inertia(3 + k) * v[k];
        }
        // ad(phi)^T (I phi) = [Iw x w + mv x v; mv x w]
        double c1[3], c2[3], c3[3];
        cross(Iw, w, c1);
        cross(mv, v, c2);
        cross(mv, w, c3);
        auto RT = [&](const double x[3], double out[3]) {
            for (int i = 0; i < 3; ++i) out[i] = E(0, i) * x[0] + E(1, i) * x[1] + E(2, i) * x[2];
        };
        const double mg[3] = {b.body->_mass * _options->_gravity(0), b.body->_mass * _options->_gravity(1),
                              b.body->_mass * _options->_gravity(2)};
        const double te[3] = {b.body->_f_external(0), b.body->_f_external(1), b.body->_f_external(2)};
        const double fe[3] = {b.body->_f_external(3), b.body->_f_external(4), b.body->_f_external(5)};
        double g_b[3], t_b[3], f_b[3];
        RT(mg, g_b);
        RT(te, t_b);
        RT(fe, f_b);
        for (int k = 0; k < 3; ++k) {
            fm(o + k) += c1[k] + c2[k] + b.body->_Inertia(k) * t_b[k];
            fm(o + 3 + k) += c3[k] + g_b[k] + b.body->_Inertia(3 + k) * f_b[k];
        }
        if (!derivs) continue;
        MatrixX& K = *Km;
        MatrixX& D = *Dm;
        double S[9];
        // gravity: d(R^T m g)/d(rotation) = [R^T m g]x
        skew(g_b, S);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) K(o + 3 + i, o + j) += S[3 * i + j];
        // external: rows scaled by the inertia entries
        skew(t_b, S);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) K(o + i, o + j) += S[3 * i + j] * b.body->_Inertia(i);
        skew(f_b, S);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) K(o + 3 + i, o + j) += S[3 * i + j] * b.body->_Inertia(3 + i);
        // Coriolis: ad(phi)^T diag(I) + [[ [Iw]x, [mv]x ], [ [mv]x, 0 ]]
        double adT[36] = {0};
        double Sw[9], Sv[9];
        skew(w, Sw);
        skew(v, Sv);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                adT[6 * i + j] = -Sw[3 * i + j];
                adT[6 * i + 3 + j] = -Sv[3 * i + j];
                adT[6 * (3 + i) + 3 + j] = -Sw[3 * i + j];
            }
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 6; ++k) D(o + i, o + k) += adT[6 * i + k] * b.body->_Inertia(k);
        double SI[9], SM[9];
        skew(Iw, SI);
        skew(mv, SM);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                D(o + i, o + j) += SI[3 * i + j];
                D(o + i, o + 3 + j) += SM[3 * i + j];
                D(o + 3 + i, o + j) += SM[3 * i + j];
            }
    }
}

// Pairs whose contact points can reach the SDF body's grid: the world AABB of the
// contact points' body-frame box against the world AABB of the grid box. A point
// outside the grid box is never active (query_SDF gives FLT_MAX there), so a pair
// whose boxes do not meet has no contact; the boxes are widened by 1e-6 of their size.
std::vector<int> Simulation::broad_phase() const {
    auto world_box = [](const Matrix4& E, const Vector3& lo, const Vector3& hi, double out_lo[3], double out_hi[3]) {
        for (int a = 0; a < 3; ++a) {
            out_lo[a] = INFINITY;
            out_hi[a] = -INFINITY;
        }
        for (int c = 0; c < 8; ++c) {
            const double x[3] = {(c & 1) ? hi(0) : lo(0), (c & 2) ? hi(1) : lo(1), (c & 4) ? hi(2) : lo(2)};
            for (int a = 0; a < 3; ++a) {
                const double y = E(a, 0) * x[0] + E(a, 1) * x[1] + E(a, 2) * x[2] + E(a, 3);
                out_lo[a] = std::min(out_lo[a], y);
                out_hi[a] = std::max(out_hi[a], y);
            }
        }
        double ext = 0.0;
        for (int a = 0; a < 3; ++a) ext = std::max(ext, out_hi[a] - out_lo[a]);
        const double m = 1e-6 * ext + 1e-12;
        for (int a = 0; a < 3; ++a) {
            out_lo[a] -= m;
            out_hi[a] += m;
        }
    };
    std::vector<int> keep;
    for (size_t c = 0; c < _contacts.size(); ++c) {
        const SimBody& g = _bodies[_contacts[c].general_body];
        const SimBody& s = _bodies[_contacts[c].sdf_body];
        if (!_broad_phase) {
            keep.push_back((int)c);
            continue;
        }
        if (g.body->_contact_points.empty()) continue;
        double l1[3], h1[3], l2[3], h2[3];
        world_box(g.body->_E_0i, g.pts_lo, g.pts_hi, l1, h1);
        world_box(s.body->_E_0i, s.sdf->_min_box, s.sdf->_max_box, l2, h2);
        bool meet = true;
        for (int a = 0; a < 3; ++a) meet = meet && h1[a] >= l2[a] && h2[a] >= l1[a];
        if (meet) keep.push_back((int)c);
    }
    return keep;
}

// ForceGeneralSDFContact::computeForce(WithDerivative) of every pair the broad phase
// keeps, in one batched GPU evaluation; the wrench, Km and Dm scattered at the two
// bodies' maximal indices (ForceGeneralSDFContact.cpp:46-144) and the contact-body
// history updated (CollisionDetection.cpp:143-146).
void Simulation::collect_instances(std::vector<int>& pairs, std::vector<atac_instance>& inst) const {
    pairs = broad_phase();
    for (int p : pairs) {
        const SdfContact& c = _contacts[p];
        const SimBody& g = _bodies[c.general_body];
        const SimBody& s = _bodies[c.sdf_body];
        atac_instance in{};
        in.grid = _grid_of_body[c.sdf_body];
        in.points = 0;
        in.segment = c.general_body;
        to_row_major(g.body->_E_0i, in.E1);
        to_row_major(s.body->_E_0i, in.E2);
        for (int k = 0; k < 6; ++k) {
            in.phi1[k] = g.body->_phi(k);
            in.phi2[k] = s.body->_phi(k);
        }
        in.col_th = s.sdf->_col_th;
        in.kn = c.kn;
        in.kt = c.kt;
        in.mu = c.mu;
        in.damping = c.damping;
        in.scale = c.scale;
        in.contact_scale = g.body->_contact_scale;
        inst.push_back(in);
    }
}

void Simulation::evaluate_instances(const std::vector<int>& key, std::vector<atac_instance>& inst, bool derivs,
                                    ContactEval& out) {
    const auto t0 = Clock::now();
    const size_t n = inst.size();
    out.derivs = derivs;
    out.f.assign(12 * n, 0.0);
    out.K.assign(derivs ? 144 * n : 0, 0.0);
    out.D.assign(derivs ? 144 * n : 0, 0.0);
    out.na.assign(n, 0);
    if (n == 0) return;
    auto it = _batches.find(key);
    if (it == _batches.end()) {
        if (_batches.size() >= kMaxCachedBatches) {   // drop the least recently used
            auto old = _batches.begin();
            for (auto jt = _batches.begin(); jt != _batches.end(); ++jt)
                if (jt->second.last_use < old->second.last_use) old = jt;
            atac_batch_destroy(old->second.batch);
            _batches.erase(old);
        }
        CachedBatch cb;
        check_rc(atac_batch_create(_grids.data(), (int32_t)_grids.size(), &_points, 1, inst.data(), (int32_t)n, nullptr,
                                   &cb.batch),
                 "atac_batch_create");
        ++_stats.batches_built;
        it = _batches.emplace(key, cb).first;
    }
    it->second.last_use = ++_use_clock;
    const uint32_t flags = (derivs ? ATAC_EVAL_DERIV : ATAC_EVAL_FORCE) | (_precise ? ATAC_EVAL_PRECISE : 0u);
    check_rc(atac_batch_eval_instances_host(it->second.batch, inst.data(), (int32_t)n, flags, out.f.data(),
                                            derivs ? out.K.data() : nullptr, derivs ? out.D.data() : nullptr, nullptr,
                                            out.na.data()),
             "atac_batch_eval_instances_host");
    ++_stats.evaluations;
    if (derivs) ++_stats.derivative_evaluations;
    _last_batch = it->second.batch;
    _last_eval.pairs.clear();
    for (int k : key)
        if (k >= 0) _last_eval.pairs.push_back({_contacts[k].general_body, _contacts[k].sdf_body});
    _last_eval.f = out.f;
    _last_eval.K = out.K;
    _last_eval.D = out.D;
    _last_eval.na = out.na;
    _last_eval.derivs = derivs;
    _stats.eval_ms += ms_since(t0);
}

std::vector<double> Simulation::contact_vjp(const std::vector<double>& g_fm) {
    const size_t n = _last_eval.pairs.size();
    if (!_last_batch || n == 0) throw_error("contact_vjp: no contact evaluation to differentiate");
    if (g_fm.size() != 12 * n)
        throw_error("contact_vjp: g_fm needs 12 values per instance of the last evaluation (" + std::to_string(n) + ")");
    std::vector<double> q1(6 * n), q2(6 * n), p1(6 * n), p2(6 * n), kn(n), kt(n);
    check_rc(atac_batch_vjp_host(_last_batch, g_fm.data(), q1.data(), q2.data(), p1.data(), p2.data(), kn.data(),
                                 kt.data()),
             "atac_batch_vjp_host");
    std::vector<double> out(26 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 6; ++c) {
            out[26 * i + c] = q1[6 * i + c];
            out[26 * i + 6 + c] = q2[6 * i + c];
            out[26 * i + 12 + c] = p1[6 * i + c];
            out[26 * i + 18 + c] = p2[6 * i + c];
        }
        out[26 * i + 24] = kn[i];
        out[26 * i + 25] = kt[i];
    }
    return out;
}

void Simulation::scatter_contacts(const ContactEval& ce, size_t st, VectorX& fm, MatrixX* Km, MatrixX* Dm) {
    const std::vector<int>& pairs = ce.pairs[st];
    for (size_t k = 0; k < pairs.size(); ++k) {
        const size_t t = ce.first[st] + k;
        if (ce.na[t] < 0) throw_error("Simulation: the contact evaluation overflowed its record pool");
        const SdfContact& c = _contacts[pairs[k]];
        const int o[2] = {6 * c.general_body, 6 * c.sdf_body};
        for (int r = 0; r < 12; ++r) fm(o[r / 6] + r % 6) += ce.f[12 * t + r];
        if (Km && ce.derivs)
            for (int r = 0; r < 12; ++r)
                for (int q = 0; q < 12; ++q) {
                    (*Km)(o[r / 6] + r % 6, o[q / 6] + q % 6) += ce.K[144 * t + 12 * r + q];
                    (*Dm)(o[r / 6] + r % 6, o[q / 6] + q % 6) += ce.D[144 * t + 12 * r + q];
                }
        if (ce.na[t] > 0) {
            _bodies[c.general_body].body->add_contact_body(_bodies[c.sdf_body].name);
            _bodies[c.sdf_body].body->add_contact_body(_bodies[c.general_body].name);
        }
    }
}

// ForceGeneralSDFContact::computeForce(WithDerivative) of every pair the broad phase
// keeps, in one batched GPU evaluation; the wrench, Km and Dm scattered at the two
// bodies' maximal indices (ForceGeneralSDFContact.cpp:46-144) and the contact-body
// history updated (CollisionDetection.cpp:143-146).
void Simulation::contact_forces(bool derivs, VectorX& fm, MatrixX* Km, MatrixX* Dm) {
    if (_contacts.empty()) return;
    if (_override) {
        scatter_contacts(*_override, _override_state, fm, Km, Dm);
        return;
    }
    ContactEval ce;
    ce.pairs.resize(1);
    ce.first = {0};
    std::vector<atac_instance> inst;
    collect_instances(ce.pairs[0], inst);
    _stats.pairs_culled += (long long)(_contacts.size() - ce.pairs[0].size());
    _stats.pairs_evaluated += (long long)ce.pairs[0].size();
    if (inst.empty()) return;
    evaluate_instances(ce.pairs[0], inst, derivs, ce);
    scatter_contacts(ce, 0, fm, Km, Dm);
}

// ------------------------------------------------------------------ dynamics

void Simulation::computeMatrices(MatrixX& M, VectorX& f) {
    update_kinematics(false);   // M and f need no q-derivatives of S_j
    VectorX fm = VectorX::Zero(_ndof_m);
    body_forces(false, fm, nullptr, nullptr);
    contact_forces(false, fm, nullptr, nullptr);
    f = VectorX::Zero(_ndof_r);
    for (const auto& j : _joints) j->computeJointForce(f);
    M = MatrixX::Zero(_ndof_r, _ndof_r);
    // M = J^T Mm J, f = J^T (fm - Mm Jdot qdot) + fr (Simulation.cpp:379-419)
    for (size_t bi = 0; bi < _bodies.size(); ++bi) {
        const SimBody& b = _bodies[bi];
        const Joint& j = *b.joint;
        if (j._ndof == 0) continue;
        const BodyJac J = body_jacobian(b, false);
        const int n = J.n, o = j._index[0], mo = 6 * (int)bi;
        double v[6];
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int c = 0; c < n; ++c) s += J.Jd[n * r + c] * j._qdot(c);
            v[r] = b.body->_Inertia(r) * s;
        }
        for (int a = 0; a < n; ++a) {
            for (int c = 0; c < n; ++c) {
                double s = 0.0;
                for (int r = 0; r < 6; ++r) s += J.J[n * r + a] * b.body->_Inertia(r) * J.J[n * r + c];
                M(o + a, o + c) = s;
            }
            double s = 0.0;
            for (int r = 0; r < 6; ++r) s += J.J[n * r + a] * (fm(mo + r) - v[r]);
            f(o + a) += s;
        }
    }
}

void Simulation::computeMatrices(MatrixX& M, VectorX& f, std::vector<MatrixX>& dM_dq, MatrixX& K, MatrixX& D) {
    std::vector<std::vector<MatrixX>> dMj;
    assemble(M, f, K, D, dMj);
    dM_dq.assign(_ndof_r, MatrixX::Zero(_ndof_r, _ndof_r));
    for (size_t ji = 0; ji < _joints.size(); ++ji) {
        const Joint& j = *_joints[ji];
        for (int k = 0; k < j._ndof; ++k)
            for (int a = 0; a < j._ndof; ++a)
                for (int c = 0; c < j._ndof; ++c) dM_dq[j._index[k]](j._index[a], j._index[c]) = dMj[ji][k](a, c);
    }
}

void Simulation::assemble(MatrixX& M, VectorX& f, MatrixX& K, MatrixX& D, std::vector<std::vector<MatrixX>>& dMj) {
    update_robot();
    VectorX fm = VectorX::Zero(_ndof_m);
    MatrixX Km = MatrixX::Zero(_ndof_m, _ndof_m), Dm = MatrixX::Zero(_ndof_m, _ndof_m);
    body_forces(true, fm, &Km, &Dm);
    contact_forces(true, fm, &Km, &Dm);
    f = VectorX::Zero(_ndof_r);
    MatrixX Kr = MatrixX::Zero(_ndof_r, _ndof_r), Dr = MatrixX::Zero(_ndof_r, _ndof_r);
    for (const auto& j : _joints) j->computeJointForceWithDerivative(f, Kr, Dr);
    M = MatrixX::Zero(_ndof_r, _ndof_r);
    K = Kr;
    D = Dr;
    dMj.assign(_joints.size(), std::vector<MatrixX>());
    const int nb = (int)_bodies.size();
    std::vector<BodyJac> Js(nb);
    std::vector<std::vector<double>> dJq(nb);   // dJ_dq(k) qdot as columns: 6 x n
    for (int bi = 0; bi < nb; ++bi) {
        const SimBody& b = _bodies[bi];
        const Joint& j = *b.joint;
        if (j._ndof == 0) continue;
        Js[bi] = body_jacobian(b, true);
        const BodyJac& J = Js[bi];
        const int n = J.n, o = j._index[0], mo = 6 * bi;
        dJq[bi].assign(6 * n, 0.0);
        for (int k = 0; k < n; ++k)
            for (int r = 0; r < 6 && !J.dzero[k]; ++r) {
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += J.dJ[k][n * r + c] * j._qdot(c);
                dJq[bi][n * r + k] = s;
            }
        const double* Mm = b.body->_Inertia.data();
        double MJdq[6];   // Mm Jdot qdot
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int c = 0; c < n; ++c) s += J.Jd[n * r + c] * j._qdot(c);
            MJdq[r] = Mm[r] * s;
        }
        for (int a = 0; a < n; ++a) {
            double s = 0.0;
            for (int r = 0; r < 6; ++r) s += J.J[n * r + a] * (fm(mo + r) - MJdq[r]);
            f(o + a) += s;   // J^T fm + fqvv (fqvv = -J^T Mm Jdot qdot)
            for (int c = 0; c < n; ++c) {
                double m = 0.0, dq = 0.0;
                for (int r = 0; r < 6; ++r) {
                    m += J.J[n * r + a] * Mm[r] * J.J[n * r + c];
                    // Dqvv = -J^T Mm Jdot - J^T Mm dJdq_qdot
                    dq -= J.J[n * r + a] * Mm[r] * (J.Jd[n * r + c] + dJq[bi][n * r + c]);
                }
                M(o + a, o + c) = m;
                D(o + a, o + c) += dq;
            }
        }
        const size_t ji = (size_t)j._id;
        dMj[ji].assign(n, MatrixX::Zero(n, n));
        for (int k = 0; k < n; ++k) {
            if (J.dzero[k]) continue;   // dM_dq(k) and the K column below are zero
            // dM_dq(k) = dJ_k^T Mm J + (dJ_k^T Mm J)^T (non-zero only in the joint's block)
            MatrixX& dM = dMj[ji][k];
            for (int a = 0; a < n; ++a)
                for (int c = 0; c < n; ++c) {
                    double s = 0.0;
                    for (int r = 0; r < 6; ++r) s += J.dJ[k][n * r + a] * Mm[r] * J.J[n * r + c];
                    dM(a, c) += s;
                    dM(c, a) += s;
                }
            // Kqvv col k = -dJ_k^T (Mm Jdot qdot) - J^T Mm (dJdot_k qdot); + dJ_k^T fm
            double dJdq[6];
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += J.dJd[k][n * r + c] * j._qdot(c);
                dJdq[r] = s;
            }
            for (int a = 0; a < n; ++a) {
                double s = 0.0;
                for (int r = 0; r < 6; ++r)
                    s += -J.dJ[k][n * r + a] * MJdq[r] - J.J[n * r + a] * Mm[r] * dJdq[r] + J.dJ[k][n * r + a] * fm(mo + r);
                K(o + a, o + k) += s;
            }
        }
    }
    // J^T Km J + J^T Dm dJdq_qdot, J^T Dm J over the body pairs with non-zero blocks:
    // each body's own, and those of the contact pairs this evaluation gave a force
    auto nonzero = [&](int a, int b) {
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c)
                if (Km(6 * a + r, 6 * b + c) != 0.0 || Dm(6 * a + r, 6 * b + c) != 0.0) return true;
        return false;
    };
    std::set<std::pair<int, int>> blocks;
    for (int bi = 0; bi < nb; ++bi) blocks.insert({bi, bi});
    for (const auto& c : _contacts) {
        if (nonzero(c.general_body, c.sdf_body)) blocks.insert({c.general_body, c.sdf_body});
        if (nonzero(c.sdf_body, c.general_body)) blocks.insert({c.sdf_body, c.general_body});
    }
    for (const auto& ab : blocks) {
        const int a = ab.first, b = ab.second;
        if (_bodies[a].joint->_ndof == 0 || _bodies[b].joint->_ndof == 0) continue;
        const BodyJac &Ja = Js[a], &Jb = Js[b];
        const int na = Ja.n, nbb = Jb.n, oa = _bodies[a].joint->_index[0], ob = _bodies[b].joint->_index[0];
        double KJ[6][6], DJ[6][6], DdJ[6][6];   // Km_ab Jb, Dm_ab Jb, Dm_ab dJq_b
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < nbb; ++c) {
                double k = 0.0, d = 0.0, dd = 0.0;
                for (int s = 0; s < 6; ++s) {
                    k += Km(6 * a + r, 6 * b + s) * Jb.J[nbb * s + c];
                    d += Dm(6 * a + r, 6 * b + s) * Jb.J[nbb * s + c];
                    dd += Dm(6 * a + r, 6 * b + s) * dJq[b][nbb * s + c];
                }
                KJ[r][c] = k;
                DJ[r][c] = d;
                DdJ[r][c] = dd;
            }
        for (int i = 0; i < na; ++i)
            for (int c = 0; c < nbb; ++c) {
                double k = 0.0, d = 0.0;
                for (int r = 0; r < 6; ++r) {
                    k += Ja.J[na * r + i] * (KJ[r][c] + DdJ[r][c]);
                    d += Ja.J[na * r + i] * DJ[r][c];
                }
                K(oa + i, ob + c) += k;
                D(oa + i, ob + c) += d;
            }
    }
}

// ------------------------------------------------------------------ BDF1

// g = M (q1 - q0 - h qdot0) - h^2 f at qdot1 = (q1 - q0)/h (Simulation.cpp:1081-1091)
void Simulation::evaluate_g(const VectorX& q1, VectorX& g) {
    VectorX qdot1(_ndof_r);
    for (int i = 0; i < _ndof_r; ++i) qdot1(i) = (q1(i) - _q0(i)) / _h;
    set_state(q1, qdot1);
    MatrixX M;
    VectorX f;
    computeMatrices(M, f);
    g = VectorX(_ndof_r);
    for (int i = 0; i < _ndof_r; ++i) {
        double s = 0.0;
        for (int c = 0; c < _ndof_r; ++c) s += M(i, c) * (q1(c) - _q0(c) - _h * _qdot0(c));
        g(i) = s - _h * _h * f(i);
    }
}

// and H = M - h^2 K - h D + [dM/dq_k (q1 - q0 - h qdot0)]_k (Simulation.cpp:1093-1117)
void Simulation::evaluate_g_with_derivatives(const VectorX& q1, VectorX& g, MatrixX& H) {
    VectorX qdot1(_ndof_r);
    for (int i = 0; i < _ndof_r; ++i) qdot1(i) = (q1(i) - _q0(i)) / _h;
    set_state(q1, qdot1);
    MatrixX M, K, D;
    VectorX f;
    std::vector<std::vector<MatrixX>> dMj;
    assemble(M, f, K, D, dMj);
    VectorX dq(_ndof_r);
    for (int i = 0; i < _ndof_r; ++i) dq(i) = q1(i) - _q0(i) - _h * _qdot0(i);
    g = VectorX(_ndof_r);
    H = MatrixX(_ndof_r, _ndof_r);
    for (int i = 0; i < _ndof_r; ++i) {
        double s = 0.0;
        for (int c = 0; c < _ndof_r; ++c) s += M(i, c) * dq(c);
        g(i) = s - _h * _h * f(i);
        for (int c = 0; c < _ndof_r; ++c) H(i, c) = M(i, c) - _h * _h * K(i, c) - _h * D(i, c);
    }
    for (size_t ji = 0; ji < _joints.size(); ++ji) {
        const Joint& j = *_joints[ji];
        for (int k = 0; k < j._ndof; ++k)
            for (int a = 0; a < j._ndof; ++a) {
                double s = 0.0;
                for (int c = 0; c < j._ndof; ++c) s += dMj[ji][k](a, c) * dq(j._index[c]);
                H(j._index[a], j._index[k]) += s;
            }
    }
}

// Simulation::newton (Simulation.cpp:999-1079): full Newton steps with a halving line
// search on |g|, three failed line searches in a row end it
void Simulation::newton(VectorX& x) {
    if (_ndof_r == 0) return;
    const int max_iter = std::max(20 * _ndof_r, 100);
    bool success = false;
    int fail_strike = 0;
    dtype g_last = 0.0;
    for (int it = 0; it < max_iter; ++it) {
        VectorX g;
        MatrixX H;
        evaluate_g_with_derivatives(x, g, H);
        ++_stats.newton_iterations;
        VectorX mg(_ndof_r);
        for (int i = 0; i < _ndof_r; ++i) mg(i) = -g(i);
        const VectorX dx = lu_solve(H, mg);
        const dtype gnorm = norm(g);
        dtype alpha = 1.0;
        bool ok = false;
        VectorX g_new, xt(_ndof_r);
        for (int trial = 0; trial < kMaxIterLS; ++trial, alpha *= 0.5) {
            for (int i = 0; i < _ndof_r; ++i) xt(i) = x(i) + alpha * dx(i);
            evaluate_g(xt, g_new);
            if (norm(g_new) < gnorm) {
                ok = true;
                break;
            }
        }
        if (ok) {
            fail_strike = 0;
        } else if (++fail_strike >= kMaxFailStrike) {
            break;
        }
        for (int i = 0; i < _ndof_r; ++i) x(i) += alpha * dx(i);
        if (norm(g_new) < kNewtonTol) {
            success = true;
            break;
        }
        g_last = norm(g_new);
    }
    _newton_converge = success;
    if (!success && _verbose) std::fprintf(stderr, "g = %g\nNewton method did not converge\n", g_last);
}

void Simulation::integration_BDF1(const VectorX& q0, const VectorX& qdot0, dtype h, VectorX& q1, VectorX& qdot1) {
    _q0 = q0;
    _qdot0 = qdot0;
    _h = h;
    q1 = VectorX(_ndof_r);
    for (int i = 0; i < _ndof_r; ++i) q1(i) = q0(i) + h * qdot0(i);
    newton(q1);
    qdot1 = VectorX(_ndof_r);
    for (int i = 0; i < _ndof_r; ++i) qdot1(i) = (q1(i) - q0(i)) / h;
}

void Simulation::reset(bool backward_flag, bool) {
    if (backward_flag) throw_error("Simulation: backward (the time adjoint) is not in this build");
    set_state(_q_init, _qdot_init);
    clear_all_contact_bodies();
    update_robot();
    for (auto& b : _bodies) b.body->reset_external_force();
    _q_his.clear();
    _qdot_his.clear();
    _q_his.push_back(_q_init);
    _qdot_his.push_back(_qdot_init);
}

void Simulation::forward_candidates(std::vector<VectorX>& q, std::vector<VectorX>& qdot,
                                    const std::vector<std::vector<Vector6>>& f_external, int num_steps,
                                    std::vector<int>* converged) {
    const size_t nc = q.size();
    if (qdot.size() != nc || f_external.size() != nc)
        throw_error("forward_candidates: q, qdot and f_external need one entry per candidate");
    for (size_t c = 0; c < nc; ++c) {
        if (q[c].size() != _ndof_r || qdot[c].size() != _ndof_r) throw_error("forward_candidates: state size != ndof_r");
        if (f_external[c].size() != _bodies.size()) throw_error("forward_candidates: one external force per body");
    }
    if (_options->_integrator != "BDF1")
        throw_error("Integrator " + _options->_integrator + " is not in this build (BDF1)");
    // the scene's own state, restored at the end
    const VectorX q_save = get_q(), qd_save = get_qdot();
    std::vector<Vector6> f_save;
    for (const auto& b : _bodies) f_save.push_back(b.body->_f_external);
    const dtype h = _options->_h;
    const int max_iter = std::max(20 * _ndof_r, 100);
    // per candidate: the reference newton's state (Simulation.cpp:999-1079)
    struct Cand {
        VectorX q0, qd0, x, dx, g, gnew;
        dtype gnorm = 0, alpha = 1;
        int trial = 0, fail_strike = 0, iter = 0;
        bool active = false, ls_done = false, ls_ok = false, conv = false;
    };
    std::vector<Cand> C(nc);
    auto load = [&](size_t c, const VectorX& x) {
        VectorX xd(_ndof_r);
        for (int i = 0; i < _ndof_r; ++i) xd(i) = (x(i) - C[c].q0(i)) / h;
        set_state(x, xd);
        for (size_t b = 0; b < _bodies.size(); ++b) _bodies[b].body->_f_external = f_external[c][b];
        update_robot();
        _q0 = C[c].q0;
        _qdot0 = C[c].qd0;
        _h = h;
    };
    // one round: the states xs[k] of candidates cs[k] evaluated together; g (and H)
    auto round = [&](const std::vector<size_t>& cs, const std::vector<VectorX>& xs, bool derivs,
                     std::vector<VectorX>& g, std::vector<MatrixX>* H) {
        ContactEval ce;
        std::vector<atac_instance> inst;
        std::vector<int> key;
        ce.pairs.resize(cs.size());
        for (size_t k = 0; k < cs.size(); ++k) {
            load(cs[k], xs[k]);
            ce.first.push_back(inst.size());
            if (!_contacts.empty()) collect_instances(ce.pairs[k], inst);
            _stats.pairs_culled += (long long)(_contacts.size() - ce.pairs[k].size());
            _stats.pairs_evaluated += (long long)ce.pairs[k].size();
            key.insert(key.end(), ce.pairs[k].begin(), ce.pairs[k].end());
            key.push_back(-1);
        }
        evaluate_instances(key, inst, derivs, ce);
        g.assign(cs.size(), VectorX());
        if (H) H->assign(cs.size(), MatrixX());
        for (size_t k = 0; k < cs.size(); ++k) {
            load(cs[k], xs[k]);
            _override = &ce;
            _override_state = k;
            try {
                if (derivs) evaluate_g_with_derivatives(xs[k], g[k], (*H)[k]);
                else evaluate_g(xs[k], g[k]);
            } catch (...) {
                _override = nullptr;
                throw;
            }
            _override = nullptr;
        }
    };
    for (int step = 0; step < num_steps; ++step) {
        for (size_t c = 0; c < nc; ++c) {
            C[c] = Cand();
            C[c].q0 = q[c];
            C[c].qd0 = qdot[c];
            C[c].x = VectorX(_ndof_r);
            for (int i = 0; i < _ndof_r; ++i) C[c].x(i) = q[c](i) + h * qdot[c](i);
            C[c].active = _ndof_r > 0;
        }
        while (true) {
            std::vector<size_t> act;
            for (size_t c = 0; c < nc; ++c)
                if (C[c].active) act.push_back(c);
            if (act.empty()) break;
            // derivative round: g, H at x; the Newton directions
            std::vector<VectorX> xs, g;
            std::vector<MatrixX> H;
            for (size_t c : act) xs.push_back(C[c].x);
            round(act, xs, true, g, &H);
            for (size_t k = 0; k < act.size(); ++k) {
                Cand& cd = C[act[k]];
                cd.g = g[k];
                VectorX mg(_ndof_r);
                for (int i = 0; i < _ndof_r; ++i) mg(i) = -cd.g(i);
                cd.dx = lu_solve(H[k], mg);
                cd.gnorm = norm(cd.g);
                cd.alpha = 1.0;
                cd.trial = 0;
                cd.ls_done = false;
                cd.ls_ok = false;
                ++_stats.newton_iterations;
            }
            // line-search rounds: halving alpha until |g| decreases, 20 trials at most
            while (true) {
                std::vector<size_t> ls;
                for (size_t c : act)
                    if (!C[c].ls_done) ls.push_back(c);
                if (ls.empty()) break;
                std::vector<VectorX> xt;
                for (size_t c : ls) {
                    VectorX x(_ndof_r);
                    for (int i = 0; i < _ndof_r; ++i) x(i) = C[c].x(i) + C[c].alpha * C[c].dx(i);
                    xt.push_back(x);
                }
                std::vector<VectorX> gn;
                round(ls, xt, false, gn, nullptr);
                for (size_t k = 0; k < ls.size(); ++k) {
                    Cand& cd = C[ls[k]];
                    cd.gnew = gn[k];
                    if (norm(cd.gnew) < cd.gnorm) {
                        cd.ls_ok = true;
                        cd.ls_done = true;
                    } else {
                        cd.alpha *= 0.5;
                        if (++cd.trial >= kMaxIterLS) cd.ls_done = true;
                    }
                }
            }
            for (size_t c : act) {
                Cand& cd = C[c];
                if (cd.ls_ok) {
                    cd.fail_strike = 0;
                } else if (++cd.fail_strike >= kMaxFailStrike) {
                    cd.active = false;
                    continue;
                }
                for (int i = 0; i < _ndof_r; ++i) cd.x(i) += cd.alpha * cd.dx(i);
                if (norm(cd.gnew) < kNewtonTol) {
                    cd.conv = true;
                    cd.active = false;
                    continue;
                }
                if (++cd.iter >= max_iter) cd.active = false;
            }
        }
        // qdot1 = (q1 - q0)/h, then the joints' reparametrisation
        for (size_t c = 0; c < nc; ++c) {
            VectorX qd1(_ndof_r);
            for (int i = 0; i < _ndof_r; ++i) qd1(i) = (C[c].x(i) - C[c].q0(i)) / h;
            set_state(C[c].x, qd1);
            reparam();
            update_robot();
            q[c] = get_q();
            qdot[c] = get_qdot();
        }
        ++_stats.steps;
    }
    if (converged) {
        converged->assign(nc, 0);
        for (size_t c = 0; c < nc; ++c) (*converged)[c] = C[c].conv ? 1 : 0;
    }
    set_state(q_save, qd_save);
    for (size_t b = 0; b < _bodies.size(); ++b) _bodies[b].body->_f_external = f_save[b];
    update_robot();
}

void Simulation::forward(int num_steps, bool verbose, bool) {
    _verbose = verbose;
    if (_q_his.empty()) throw_error("Please call simulation.reset() before simulation.forward().");
    if (_options->_integrator != "BDF1")
        throw_error("Integrator " + _options->_integrator + " is not in this build (BDF1)");
    for (int s = 0; s < num_steps; ++s) {
        const auto t0 = Clock::now();
        const VectorX q = get_q(), qdot = get_qdot();
        VectorX q1, qdot1;
        integration_BDF1(q, qdot, _options->_h, q1, qdot1);
        set_state(q1, qdot1);
        reparam();
        update_robot();
        _q_his.push_back(get_q());
        _qdot_his.push_back(get_qdot());
        ++_stats.steps;
        _stats.step_ms += ms_since(t0);
    }
}

}  // namespace redmax
