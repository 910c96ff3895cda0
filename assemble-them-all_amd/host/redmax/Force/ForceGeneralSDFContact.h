// Penalty + damping + friction contact between the sampled contact points of one
// body and the SDF of another, with the reference's interface
// (simulation/redmax/Force/ForceGeneralSDFContact.h:12-54). Detection, the wrench,
// Km / Dm and the contact-scale design column run on the GPU through the atac C ABI.
#pragma once
#include <memory>
#include <vector>
#include "CollisionDetection/Contact.h"
#include "Common.h"
#include "Force/Force.h"
#include "Utils.h"

namespace redmax {

class Body;
class BodySDFObj;

class ForceGeneralSDFContact : public Force {
public:
    Body* _contact_body;    // body whose contact points are tested
    BodySDFObj* _SDF_body;  // body whose distance field is queried
    dtype _kn;              // normal stiffness
    dtype _kt;              // tangential stiffness
    dtype _mu;              // coefficient of friction
    dtype _damping;         // damping of the contact force
    dtype _scale;           // stiffness scale for continuation (set_contact_scale)

    ForceGeneralSDFContact(Simulation* sim, Body* contact_body, Body* SDF_body, dtype kn = 1., dtype kt = 0.,
                           dtype mu = 0., dtype damping = 0.);
    ~ForceGeneralSDFContact() override;

    void set_stiffness(dtype kn, dtype kt);
    void set_friction(dtype mu);
    void set_damping(dtype damping);
    void set_scale(dtype scale);

    void computeForce(VectorX& fm, VectorX& fr, bool verbose = false) override;
    void computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr, MatrixX& Dr,
                                    bool verbose = false) override;
    void computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr, MatrixX& Dr,
                                    MatrixX& dfm_dp, MatrixX& dfr_dp, bool verbose = false) override;

    // finite-difference self-checks (ForceGeneralSDFContact.cpp:1027-1243); they print
    // with print_error and return the largest relative error found
    dtype test_derivatives();
    void test_derivatives_runtime() override;
    void test_design_derivatives_runtime() override;
    dtype test_derivatives_runtime_error();

    // number of active contacts in the last evaluation
    int last_active_count() const { return _last_active; }

protected:
    // over a fixed contact list (every contact counts, no detection), as the
    // reference's overloads at ForceGeneralSDFContact.cpp:147, :221, :449
    void computeForce(std::vector<Contact>& contacts, VectorX& fm, bool verbose = false);
    void computeForceWithDerivative(std::vector<Contact>& contacts, VectorX& fm, MatrixX& Km, MatrixX& Dm,
                                    bool verbose = false);
    void computeForceWithDerivative(std::vector<Contact>& contacts, VectorX& fm, MatrixX& Km, MatrixX& Dm,
                                    MatrixX& dfm_dp, bool verbose = false);

private:
    struct Device;
    std::unique_ptr<Device> _dev;
    int _last_active = 0;
    long long _last_detect_us = 0;   // GPU time of the last evaluation's detection (HIP events)
    void evaluate(uint32_t flags, double* fm, double* Km, double* Dm, double* p6);
    void evaluate_list(const std::vector<Contact>& contacts, uint32_t flags, double* fm, double* Km, double* Dm,
                       double* p6);
    void fill_instance(void* inst) const;
    std::vector<Contact> active_contacts();
};

}  // namespace redmax
