// Penalty + damping + friction contact between the sampled contact points of one
// body and the triangle mesh of a BodyBVHObj, with the reference's interface
// (simulation/redmax/Force/ForceGeneralBVHContact.h): detection
// (collision_detection_general_BVH, CollisionDetection.cpp:150-185), the wrench,
// Km / Dm and the contact-scale design column, on the GPU through the atac C ABI.
// The force math is ForceGeneralSDFContact's with the mesh's signed distance and
// pseudonormal (ForceGeneralBVHContact.cpp:147-219 and its derivative overloads).
#pragma once
#include "Common.h"
#include "Force/Force.h"
#include "Utils.h"

namespace redmax {

class Body;
class BodyBVHObj;

class ForceGeneralBVHContact : public Force {
public:
    Body* _contact_body;    // body whose contact points are tested
    BodyBVHObj* _BVH_body;  // body whose mesh is queried
    dtype _kn;              // normal stiffness
    dtype _kt;              // tangential stiffness
    dtype _mu;              // coefficient of friction
    dtype _damping;         // damping of the contact force
    dtype _scale;           // stiffness scale for continuation (set_contact_scale)

    ForceGeneralBVHContact(Simulation* sim, Body* contact_body, Body* BVH_body, dtype kn = 1., dtype kt = 0.,
                           dtype mu = 0., dtype damping = 0.);

    void set_stiffness(dtype kn, dtype kt);
    void set_friction(dtype mu);
    void set_damping(dtype damping);
    void set_scale(dtype scale);

    void computeForce(VectorX& fm, VectorX& fr, bool verbose = false) override;
    void computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr, MatrixX& Dr,
                                    bool verbose = false) override;
    void computeForceWithDerivative(VectorX& fm, VectorX& fr, MatrixX& Km, MatrixX& Dm, MatrixX& Kr, MatrixX& Dr,
                                    MatrixX& dfm_dp, MatrixX& dfr_dp, bool verbose = false) override;

    // number of active contacts in the last evaluation
    int last_active_count() const { return _last_active; }

private:
    int _last_active = 0;
    void evaluate(uint32_t flags, double* fm, double* Km, double* Dm, double* p6);
};

}  // namespace redmax
