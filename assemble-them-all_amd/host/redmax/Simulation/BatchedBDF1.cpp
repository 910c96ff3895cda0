#include "Simulation/BatchedBDF1.h"

#include <cmath>
#include <string>

namespace redmax {

namespace {

constexpr dtype kNewtonTol = 1e-8;     // Simulation.cpp:1007
constexpr int kMaxIterLS = 20;         // Simulation.cpp:1011
constexpr int kMaxFailStrike = 3;      // Simulation.cpp:1013

using V3 = std::array<double, 3>;

double norm3(const V3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// x = A^-1 b for a 3 x 3 row-major A: LU with partial (row) pivoting
V3 solve3(const double* A_in, const V3& b_in) {
    double A[9];
    V3 b = b_in;
    for (int i = 0; i < 9; ++i) A[i] = A_in[i];
    int perm[3] = {0, 1, 2};
    for (int k = 0; k < 3; ++k) {
        int p = k;
        for (int r = k + 1; r < 3; ++r)
            if (std::fabs(A[3 * r + k]) > std::fabs(A[3 * p + k])) p = r;
        if (p != k) {
            for (int c = 0; c < 3; ++c) std::swap(A[3 * k + c], A[3 * p + c]);
            std::swap(perm[k], perm[p]);
            std::swap(b[k], b[p]);
        }
        for (int r = k + 1; r < 3; ++r) {
            const double l = A[3 * r + k] / A[3 * k + k];
            A[3 * r + k] = l;
            for (int c = k + 1; c < 3; ++c) A[3 * r + c] -= l * A[3 * k + c];
            b[r] -= l * b[k];
        }
    }
    V3 x{};
    for (int k = 2; k >= 0; --k) {
        double s = b[k];
        for (int c = k + 1; c < 3; ++c) s -= A[3 * k + c] * x[c];
        x[k] = s / A[3 * k + k];
    }
    return x;
}

void check_rc(int rc, const char* what) {
    if (rc != ATAC_OK) throw_error(std::string("BatchedBDF1: ") + what + ": " + atac_last_error());
}

}  // namespace

BatchedBDF1::BatchedBDF1(std::vector<BDF1Part> parts, std::vector<std::pair<int, int>> contacts,
                         BDF1ContactParams params, dtype h, Vector3 gravity, bool precise)
    : parts_(std::move(parts)), contacts_(std::move(contacts)), params_(params), h_(h), precise_(precise) {
    if (!(h_ > 0.0)) throw_error("BatchedBDF1: the time step must be positive");
    for (int a = 0; a < 3; ++a) grav_[a] = gravity(a);
    grid_of_part_.assign(parts_.size(), -1);
    for (size_t p = 0; p < parts_.size(); ++p)
        if (parts_[p].grid) {
            grid_of_part_[p] = (int)grids_.size();
            grids_.push_back(parts_[p].grid);
        }
    for (const auto& k : contacts_) {
        if (k.first < 0 || k.second < 0 || k.first >= (int)parts_.size() || k.second >= (int)parts_.size() ||
            k.first == k.second)
            throw_error("BatchedBDF1: bad contact pair");
        if (grid_of_part_[k.second] < 0) throw_error("BatchedBDF1: the SDF body of a contact has no grid");
    }
}

BatchedBDF1::~BatchedBDF1() {
    if (batch_) atac_batch_destroy(batch_);
    if (points_) atac_points_destroy(points_);
}

int BatchedBDF1::add_candidate(int move, dtype mass, Vector3 q, Vector3 qdot, Vector3 u) {
    if (move < 0 || move >= (int)parts_.size()) throw_error("BatchedBDF1: bad moving part");
    if (!(mass > 0.0)) throw_error("BatchedBDF1: the mass must be positive");
    Cand c;
    c.move = move;
    c.mass = mass;
    for (int a = 0; a < 3; ++a) {
        c.q0[a] = q(a);
        c.qd0[a] = qdot(a);
        c.u[a] = u(a);
    }
    cand_.push_back(c);
    if (batch_) {   // the instance table changes: rebuilt at the next step
        atac_batch_destroy(batch_);
        batch_ = nullptr;
    }
    return (int)cand_.size() - 1;
}

void BatchedBDF1::set_state(int c, Vector3 q, Vector3 qdot) {
    for (int a = 0; a < 3; ++a) {
        cand_.at(c).q0[a] = q(a);
        cand_.at(c).qd0[a] = qdot(a);
    }
}

void BatchedBDF1::set_force(int c, Vector3 u) {
    for (int a = 0; a < 3; ++a) cand_.at(c).u[a] = u(a);
}

Vector3 BatchedBDF1::q(int c) const {
    Vector3 v;
    for (int a = 0; a < 3; ++a) v(a) = cand_.at(c).q0[a];
    return v;
}

Vector3 BatchedBDF1::qdot(int c) const {
    Vector3 v;
    for (int a = 0; a < 3; ++a) v(a) = cand_.at(c).qd0[a];
    return v;
}

void BatchedBDF1::build() {
    if (!points_) {
        std::vector<double> xyz;
        std::vector<int64_t> seg{0};
        for (const auto& p : parts_) {
            for (const auto& v : p.points)
                for (int a = 0; a < 3; ++a) xyz.push_back(v(a));
            seg.push_back((int64_t)(xyz.size() / 3));
        }
        check_rc(atac_points_create(xyz.data(), (int64_t)(xyz.size() / 3), seg.data(), (int32_t)parts_.size(), 0.0,
                                    &points_),
                 "atac_points_create");
    }
    inst_.clear();
    for (auto& c : cand_) {
        c.inst.clear();
        c.side.clear();
        for (const auto& k : contacts_) {
            if (k.first != c.move && k.second != c.move) continue;   // fixed pairs exert nothing on the joint
            atac_instance in{};
            in.grid = grid_of_part_[k.second];
            in.points = 0;
            in.segment = k.first;
            for (int e = 0; e < 16; ++e) {
                in.E1[e] = parts_[k.first].E0[e];
                in.E2[e] = parts_[k.second].E0[e];
            }
            in.col_th = params_.col_th;
            in.kn = params_.kn;
            in.kt = params_.kt;
            in.mu = params_.mu;
            in.damping = params_.damping;
            in.scale = 1.0;
            in.contact_scale = 1.0;
            c.inst.push_back((int)inst_.size());
            c.side.push_back(k.first == c.move ? 1 : 2);
            inst_.push_back(in);
        }
    }
    if (!inst_.empty()) {
        check_rc(atac_batch_create(grids_.data(), (int32_t)grids_.size(), &points_, 1, inst_.data(), (int32_t)inst_.size(),
                                   nullptr, &batch_),
                 "atac_batch_create");
        fm_.assign(inst_.size() * 12, 0.0);
        Km_.assign(inst_.size() * 144, 0.0);
        Dm_.assign(inst_.size() * 144, 0.0);
    }
}

// E = E0 translated by x (world); phi = [0; R0^T xdot] (body frame)
void BatchedBDF1::set_poses(const std::vector<V3>& x, const std::vector<V3>& xdot) {
    for (size_t ci = 0; ci < cand_.size(); ++ci) {
        const Cand& c = cand_[ci];
        const auto& E0 = parts_[c.move].E0;
        double phi[6] = {0, 0, 0, 0, 0, 0};
        for (int r = 0; r < 3; ++r)
            for (int a = 0; a < 3; ++a) phi[3 + r] += E0[4 * a + r] * xdot[ci][a];
        for (size_t k = 0; k < c.inst.size(); ++k) {
            atac_instance& in = inst_[c.inst[k]];
            double* E = c.side[k] == 1 ? in.E1 : in.E2;
            double* ph = c.side[k] == 1 ? in.phi1 : in.phi2;
            for (int e = 0; e < 16; ++e) E[e] = E0[e];
            for (int a = 0; a < 3; ++a) E[4 * a + 3] = E0[4 * a + 3] + x[ci][a];
            for (int a = 0; a < 6; ++a) ph[a] = phi[a];
        }
    }
}

void BatchedBDF1::evaluate(bool deriv) {
    if (!batch_) return;
    const uint32_t flags = (deriv ? ATAC_EVAL_DERIV : ATAC_EVAL_FORCE) | (precise_ ? ATAC_EVAL_PRECISE : 0u);
    check_rc(atac_batch_eval_instances_host(batch_, inst_.data(), (int32_t)inst_.size(), flags, fm_.data(),
                                            deriv ? Km_.data() : nullptr, deriv ? Dm_.data() : nullptr, nullptr, nullptr),
             "atac_batch_eval_instances_host");
    ++n_evals_;
}

// f = u + m g + sum over the pairs of R f_body of the moving part (Jacobian^T of the
// translational joint); K, D: R (d f_body / d v_body) R^T from Km / Dm
void BatchedBDF1::joint_force(const Cand& c, V3& f, double* K, double* D) const {
    for (int a = 0; a < 3; ++a) f[a] = c.u[a] + c.mass * grav_[a];
    if (K)
        for (int e = 0; e < 9; ++e) K[e] = D[e] = 0.0;
    const auto& E0 = parts_[c.move].E0;
    for (size_t k = 0; k < c.inst.size(); ++k) {
        const int i = c.inst[k];
        const int o = c.side[k] == 1 ? 3 : 9;   // rows (f of the moving body) = columns (its v)
        for (int r = 0; r < 3; ++r)
            for (int a = 0; a < 3; ++a) f[r] += E0[4 * r + a] * fm_[12 * i + o + a];
        if (!K) continue;
        for (int r = 0; r < 3; ++r)
            for (int s = 0; s < 3; ++s) {
                double kk = 0.0, dd = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        const double w = E0[4 * r + a] * E0[4 * s + b];
                        kk += w * Km_[144 * i + 12 * (o + a) + (o + b)];
                        dd += w * Dm_[144 * i + 12 * (o + a) + (o + b)];
                    }
                K[3 * r + s] += kk;
                D[3 * r + s] += dd;
            }
    }
}

// g = M (x - q0 - h qdot0) - h^2 f, H = M - h^2 K - h D (evaluate_g_with_derivatives_BDF1,
// Simulation.cpp:1093-1110; M = m I and dM/dq = 0 for a translational joint)
void BatchedBDF1::residual(const Cand& c, const V3& x, V3& g, double* H) const {
    V3 f;
    double K[9], D[9];
    joint_force(c, f, H ? K : nullptr, H ? D : nullptr);
    for (int a = 0; a < 3; ++a) g[a] = c.mass * (x[a] - c.q0[a] - h_ * c.qd0[a]) - h_ * h_ * f[a];
    if (!H) return;
    for (int e = 0; e < 9; ++e) H[e] = -h_ * h_ * K[e] - h_ * D[e];
    for (int a = 0; a < 3; ++a) H[4 * a] += c.mass;
}

void BatchedBDF1::step(int n) {
    if (cand_.empty()) return;
    if (!batch_) build();
    const int max_iter = std::max(20 * 3, 100);   // Simulation.cpp:1008
    std::vector<V3> xs(cand_.size()), xd(cand_.size());
    auto eval_at = [&](bool deriv, bool trial) {
        for (size_t i = 0; i < cand_.size(); ++i) {
            const Cand& c = cand_[i];
            V3 x = c.x;
            if (trial && c.active && !c.ls_done)
                for (int a = 0; a < 3; ++a) x[a] = c.x[a] + c.alpha * c.dx[a];
            xs[i] = x;
            for (int a = 0; a < 3; ++a) xd[i][a] = (x[a] - c.q0[a]) / h_;
        }
        set_poses(xs, xd);
        evaluate(deriv);
    };
    for (int s = 0; s < n; ++s) {
        for (auto& c : cand_) {
            for (int a = 0; a < 3; ++a) c.x[a] = c.q0[a] + h_ * c.qd0[a];   // initial guess
            c.active = true;
            c.converged = false;
            c.fail_strike = 0;
            c.iters = 0;
        }
        for (int it = 0; it < max_iter; ++it) {
            bool any = false;
            for (const auto& c : cand_) any = any || c.active;
            if (!any) break;
            eval_at(true, false);
            for (size_t i = 0; i < cand_.size(); ++i) {
                Cand& c = cand_[i];
                if (!c.active) continue;
                double H[9];
                residual(c, c.x, c.g, H);
                const V3 rhs{-c.g[0], -c.g[1], -c.g[2]};
                c.dx = solve3(H, rhs);
                c.gnorm = norm3(c.g);
                c.alpha = 1.0;
                c.ls_done = c.ls_ok = false;
                ++c.iters;
            }
            // line search x + alpha dx, alpha halved per failed trial (Simulation.cpp:1034-1045)
            for (int trial = 0; trial < kMaxIterLS; ++trial) {
                bool open = false;
                for (const auto& c : cand_) open = open || (c.active && !c.ls_done);
                if (!open) break;
                eval_at(false, true);
                for (size_t i = 0; i < cand_.size(); ++i) {
                    Cand& c = cand_[i];
                    if (!c.active || c.ls_done) continue;
                    residual(c, xs[i], c.gnew, nullptr);
                    if (norm3(c.gnew) < c.gnorm) c.ls_done = c.ls_ok = true;
                    else c.alpha *= 0.5;
                }
            }
            for (auto& c : cand_) {
                if (!c.active) continue;
                if (c.ls_ok) {
                    c.fail_strike = 0;
                } else if (++c.fail_strike >= kMaxFailStrike) {
                    c.active = false;
                    continue;
                }
                for (int a = 0; a < 3; ++a) c.x[a] += c.alpha * c.dx[a];
                if (norm3(c.gnew) < kNewtonTol) {
                    c.converged = true;
                    c.active = false;
                }
            }
        }
        for (auto& c : cand_) {
            for (int a = 0; a < 3; ++a) {
                const double q1 = c.x[a];
                c.qd0[a] = (q1 - c.q0[a]) / h_;
                c.q0[a] = q1;
            }
        }
    }
}

}  // namespace redmax
