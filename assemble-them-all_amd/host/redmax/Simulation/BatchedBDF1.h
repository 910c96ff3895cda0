// Implicit (BDF1) time stepping of parts on translational joints against fixed
// parts, with SDF contact forces, for many candidates at once: the reference's
// Simulation::integration_BDF1 / newton / evaluate_g(_with_derivatives)_BDF1
// (simulation/redmax/Simulation.cpp:999-1098, 1248-1270) for the joint type its
// disassembly planner moves parts with (examples/run_joint_plan.py:161, 216:
// 'translational'), every candidate (moving part, start state, applied force) run
// by its own Newton iteration in FP64 on the host. Each Newton iteration and each
// line-search trial of all candidates is one batched evaluation of the
// ForceGeneralSDFContact pairs on the GPU (atac_batch_eval_host).
#pragma once
#include <array>
#include <utility>
#include <vector>

#include "Common.h"
#include "atac.h"

namespace redmax {

// One part: its contact points (its own frame; may be empty), its SDF grid (may be
// null when no contact uses it as the SDF body) and its pose when its joint is at
// q = 0 (E = [R p; 0 1], row-major 4 x 4 in `E0`).
struct BDF1Part {
    std::vector<Vector3> points;
    atac_grid grid = nullptr;
    std::array<double, 16> E0{};
};

// ForceGeneralSDFContact parameters shared by every contact pair.
struct BDF1ContactParams {
    dtype kn = 1e6, kt = 1e3, mu = 0.0, damping = 0.0;
    dtype col_th = 0.01;
};

class BatchedBDF1 {
public:
    // contacts: (general_body, SDF_body) part pairs (Simulation_Constructor's
    // <general_SDF_contact general_body= SDF_body=>); h: time step; gravity: world.
    BatchedBDF1(std::vector<BDF1Part> parts, std::vector<std::pair<int, int>> contacts, BDF1ContactParams params,
                dtype h, Vector3 gravity, bool precise = false);
    ~BatchedBDF1();
    BatchedBDF1(const BatchedBDF1&) = delete;
    BatchedBDF1& operator=(const BatchedBDF1&) = delete;

    // A candidate: part `move` on a translational joint (mass `mass`, M = mass I),
    // joint state q, qdot and a constant applied world force u. Returns its index.
    int add_candidate(int move, dtype mass, Vector3 q, Vector3 qdot, Vector3 u);
    void set_state(int c, Vector3 q, Vector3 qdot);
    void set_force(int c, Vector3 u);

    // n BDF1 steps of every candidate (Simulation::integration_BDF1).
    void step(int n = 1);

    Vector3 q(int c) const;
    Vector3 qdot(int c) const;
    bool converged(int c) const { return cand_[c].converged; }   // last step's Newton converged
    int newton_iterations(int c) const { return cand_[c].iters; }
    int n_candidates() const { return (int)cand_.size(); }
    int n_evaluations() const { return n_evals_; }             // batched GPU evaluations so far

private:
    struct Cand {
        int move;
        dtype mass;
        std::array<double, 3> q0, qd0, u;
        std::array<double, 3> x, dx, g, gnew;
        dtype alpha = 1.0, gnorm = 0.0;
        int fail_strike = 0, iters = 0;
        bool active = false, converged = false, ls_done = false, ls_ok = false;
        std::vector<int> inst;     // its instances
        std::vector<int> side;     // per instance: 1 = the moving part is the contact body, 2 = the SDF body
    };
    std::vector<BDF1Part> parts_;
    std::vector<std::pair<int, int>> contacts_;
    BDF1ContactParams params_;
    dtype h_;
    std::array<double, 3> grav_;
    bool precise_;
    std::vector<Cand> cand_;
    std::vector<atac_grid> grids_;
    std::vector<int> grid_of_part_;
    atac_points points_ = nullptr;
    atac_batch batch_ = nullptr;
    std::vector<atac_instance> inst_;
    std::vector<double> fm_, Km_, Dm_;
    int n_evals_ = 0;

    void build();
    void set_poses(const std::vector<std::array<double, 3>>& x, const std::vector<std::array<double, 3>>& xdot);
    void evaluate(bool deriv);
    // the joint-space force on candidate c from the last evaluation, and (deriv) its
    // derivatives K = df/dq, D = df/dqdot
    void joint_force(const Cand& c, std::array<double, 3>& f, double* K, double* D) const;
    void residual(const Cand& c, const std::array<double, 3>& x, std::array<double, 3>& g, double* H) const;
};

}  // namespace redmax
