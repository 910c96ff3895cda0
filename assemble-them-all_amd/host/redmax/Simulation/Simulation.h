// The simulator subset the SDF contact path runs under: redmax's Simulation
// (simulation/redmax/Simulation.{h,cpp}, Simulation_Constructor.cpp, Robot.cpp) for
// scenes of mesh and SDF bodies, each on its own joint to the world (fixed,
// translational or free3d-exp), with <general_SDF_contact> forces between them, as the
// disassembly planner builds them (examples/run_joint_plan.py:140-180). The scene comes
// from the reference's XML format; the state is stepped by its BDF1 integrator with the
// Newton solve in FP64 on the host (Simulation.cpp:999-1098, 1248-1270), and every force
// evaluation of a Newton iteration or line-search trial is one batched evaluation of all
// contact pairs on the GPU, after an AABB broad phase that leaves out pairs whose
// contact points cannot reach the other body's grid.
#pragma once
#include <array>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "Body/BodySDFObj.h"
#include "Joint/Joint.h"
#include "atac.h"

namespace redmax {

class Simulation {
public:
    // Simulation::Options (Simulation.h:83-92), with the reference's defaults
    struct Options {
        Vector3 _gravity = make_vector3(0.0, 0.0, -980.0);
        dtype _h = 0.02;
        std::string _integrator = "BDF1";
        std::string _unit = "cm-g";
    };

    // A mesh or SDF body on its joint
    struct SimBody {
        std::string name;
        Joint* joint = nullptr;
        std::unique_ptr<BodyMeshObj> body;   // a BodySDFObj for SDF bodies
        BodySDFObj* sdf = nullptr;           // the same object, for SDF bodies
        Vector3 pts_lo, pts_hi;              // box of the contact points, body frame
    };

    // ForceGeneralSDFContact between two bodies (indices into _bodies)
    struct SdfContact {
        int general_body, sdf_body;
        dtype kn, kt, mu, damping;
        dtype scale = 1.0;
    };

    Options* _options;
    std::string _name, _asset_folder;
    std::vector<std::unique_ptr<Joint>> _joints;
    std::vector<SimBody> _bodies;
    std::vector<SdfContact> _contacts;
    int _ndof_r = 0, _ndof_m = 0;
    bool _verbose = false;

    // Simulation(xml_string, asset_folder, verbose) (Simulation_Constructor.cpp:123-141);
    // precise: FP64 contact arithmetic on the GPU (detection stays FP32)
    Simulation(const std::string& xml_string, const std::string& asset_folder, bool verbose = false,
               bool precise = false);
    // Simulation(xml_file_path, verbose): the file's folder is the asset folder
    static std::unique_ptr<Simulation> from_file(const std::string& xml_file_path, bool verbose = false);
    ~Simulation();
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    // state (Simulation.cpp:44-160)
    void init();
    VectorX get_q() const;
    VectorX get_qdot() const;
    VectorX get_q_init() const { return _q_init; }
    VectorX get_qdot_init() const { return _qdot_init; }
    void set_state_init(const VectorX& q, const VectorX& qdot);
    void set_q_init(const VectorX& q);
    void set_qdot_init(const VectorX& qdot);
    void set_state(const VectorX& q, const VectorX& qdot);
    void set_q(const VectorX& q);
    void set_qdot(const VectorX& qdot);
    void update_robot(bool design_gradient = false);
    void reparam();
    const std::vector<VectorX>& get_q_his() const { return _q_his; }
    const std::vector<VectorX>& get_qdot_his() const { return _qdot_his; }

    VectorX get_joint_q(const std::string& name) const;
    VectorX get_joint_qdot(const std::string& name) const;
    void set_joint_q(const std::string& name, const VectorX& q);
    void zero_joint_q(const std::string& name);
    void zero_joint_qdot(const std::string& name);

    dtype get_body_mass(const std::string& name) const;
    Matrix4 get_body_E0i(const std::string& name) const;
    Matrix4 get_body_Ei0(const std::string& name) const;
    // 3 x n, body frame or world frame (Simulation.cpp:255-266)
    std::vector<Vector3> get_body_vertices(const std::string& name, bool world_frame = false) const;
    std::vector<std::array<int, 3>> get_body_faces(const std::string& name) const;
    std::vector<Vector3> get_body_contact_points(const std::string& name) const;
    void set_body_external_force(const std::string& name, const Vector6& force);
    // min over name_from's world vertices of name_to's SDF distance (Simulation.cpp:276-296)
    dtype get_body_distance(const std::string& name_from, const std::string& name_to);
    bool body_in_contact(const std::string& name_a, const std::string& name_b, dtype eps = 1e-5);
    std::vector<std::string> get_contact_bodies(const std::string& name) const;
    void clear_contact_bodies(const std::string& name);
    void clear_all_contact_bodies();
    void clear_all_saved_sdfs();
    // Robot::set_contact_scale (Robot.cpp:462-472)
    void set_contact_scale(dtype scale);
    // the AABB broad phase (on by default); off, every contact pair is evaluated
    void set_broad_phase(bool on) { _broad_phase = on; }

    // Simulation::reset / forward (Simulation.cpp:894-993); BDF1 only
    void reset(bool backward_flag = false, bool backward_design_params_flag = true);
    void forward(int num_steps, bool verbose = false, bool test_derivatives = false);
    bool is_converged() const { return _newton_converge; }

    // Candidates of a planner's search (examples/run_joint_plan.py:593-660 steps each
    // action from a state with its own forward calls): each candidate is this scene
    // with its own state q, qdot and external forces (one [torque; force] per body),
    // stepped num_steps BDF1 steps. All candidates advance in lockstep, each with its
    // own Newton decisions, so each ends as forward would leave it; every derivative
    // round and every line-search round of all of them is one batched contact
    // evaluation. The scene's own state is left unchanged. converged (may be null)
    // receives each candidate's last-step Newton convergence.
    void forward_candidates(std::vector<VectorX>& q, std::vector<VectorX>& qdot,
                            const std::vector<std::vector<Vector6>>& f_external, int num_steps,
                            std::vector<int>* converged = nullptr);

    // The last contact evaluation (of contact_forces, forward or forward_candidates):
    // per evaluated instance its (general_body, SDF_body) pair, wrench [tau1 f1 tau2 f2],
    // active count and, when it had derivatives, Km and Dm (12 x 12 row-major, columns
    // [q1 q2] and [phi1 phi2]).
    struct LastContactEval {
        std::vector<std::pair<int, int>> pairs;
        std::vector<double> f, K, D;
        std::vector<int32_t> na;
        bool derivs = false;
    };
    const LastContactEval& last_contact_eval() const { return _last_eval; }
    // The adjoint of that evaluation (ForceGeneralSDFContact's dfm/dq, dfm/dphi and
    // dfm/dkn contracted with a cotangent, without forming them): g_fm holds 12 values per
    // evaluated instance; returns 26 per instance: d/dq1 (6), d/dq2 (6), d/dphi1 (6),
    // d/dphi2 (6), d/dkn, d/dkt (the conventions of atac_batch_vjp)
    std::vector<double> contact_vjp(const std::vector<double>& g_fm);

    // M qddot = f in reduced coordinates, with K = df/dq, D = df/dqdot and dM/dq
    // (Simulation::computeMatrices, Simulation.cpp:379-506) at the current state
    void computeMatrices(MatrixX& M, VectorX& f);
    void computeMatrices(MatrixX& M, VectorX& f, std::vector<MatrixX>& dM_dq, MatrixX& K, MatrixX& D);

    // counters of the step loop
    struct Stats {
        long long evaluations = 0, derivative_evaluations = 0, newton_iterations = 0, steps = 0;
        long long pairs_evaluated = 0, pairs_culled = 0, batches_built = 0;
        double eval_ms = 0.0, step_ms = 0.0;
    };
    const Stats& stats() const { return _stats; }
    void reset_stats() { _stats = Stats(); }

private:
    VectorX _q_init, _qdot_init;
    std::vector<VectorX> _q_his, _qdot_his;
    VectorX _q0, _qdot0;
    dtype _h = 0.0;
    bool _newton_converge = true;
    bool _precise = false;
    bool _broad_phase = true;
    Stats _stats;

    // contact evaluation on the GPU
    atac_points _points = nullptr;
    std::vector<atac_grid> _grids;    // per SDF body (owned by the body objects)
    std::vector<int> _grid_of_body;
    struct CachedBatch {
        atac_batch batch = nullptr;
        long long last_use = 0;
    };
    std::map<std::vector<int>, CachedBatch> _batches;
    long long _use_clock = 0;
    LastContactEval _last_eval;
    atac_batch _last_batch = nullptr;   // the batch of the last evaluation (its records are the adjoint's)

    void parse(const std::string& xml_string);
    void update_kinematics(bool derivs);
    void build_contact_data();
    int body_index(const std::string& name) const;
    const Joint* joint_by_name(const std::string& name) const;
    Joint* joint_by_name(const std::string& name);
    // the contact pairs that can touch at the current poses
    std::vector<int> broad_phase() const;
    // one contact evaluation of several states: per state its kept pairs, and the
    // batch outputs of all of them (12, 144, 144 per pair)
    struct ContactEval {
        std::vector<std::vector<int>> pairs;
        std::vector<size_t> first;   // per state, its first pair's position in the outputs
        std::vector<double> f, K, D;
        std::vector<int32_t> na;
        bool derivs = false;
    };
    // the instances of the kept pairs at the current state, appended
    void collect_instances(std::vector<int>& pairs, std::vector<atac_instance>& inst) const;
    // evaluate instances of the given pair-list key with a cached batch
    void evaluate_instances(const std::vector<int>& key, std::vector<atac_instance>& inst, bool derivs, ContactEval& out);
    // the forces of state s of `ce` added at the bodies' maximal indices (and the history)
    void scatter_contacts(const ContactEval& ce, size_t s, VectorX& fm, MatrixX* Km, MatrixX* Dm);
    // when set, contact_forces takes state _override_state of this evaluation instead
    // of evaluating the current state
    const ContactEval* _override = nullptr;
    size_t _override_state = 0;
    // all contact forces at the current state: fm (6 per body), and with derivatives
    // Km, Dm (maximal coordinates)
    void contact_forces(bool derivs, VectorX& fm, MatrixX* Km, MatrixX* Dm);
    void body_forces(bool derivs, VectorX& fm, MatrixX* Km, MatrixX* Dm) const;
    // M, f, K, D with dM/dq as per-joint blocks (dMj[joint][k] n x n, the rest is zero)
    void assemble(MatrixX& M, VectorX& f, MatrixX& K, MatrixX& D, std::vector<std::vector<MatrixX>>& dMj);
    void evaluate_g(const VectorX& q1, VectorX& g);
    void evaluate_g_with_derivatives(const VectorX& q1, VectorX& g, MatrixX& H);
    void newton(VectorX& x);
    void integration_BDF1(const VectorX& q0, const VectorX& qdot0, dtype h, VectorX& q1, VectorX& qdot1);
};

}  // namespace redmax
