// Joints of the standalone simulator: the reduced coordinates of one body attached to
// the world (simulation/redmax/Joint/Joint.{h,cpp}; the fixed, translational and
// free3d-exp types, Joint/JointFixed.cpp, JointTranslational.cpp, JointFree3DExp.cpp).
// A joint with pose E_pj_0 at q = 0 moves its frame to E_0j = E_pj_0 Q(q); its twist
// in its own frame is phi = S_j(q) qdot, [w; v] order. update() fills Q, S_j, its time
// derivative S_j_dot and their partial derivatives in q (holding qdot), which the
// reduced-coordinate dynamics needs for M, K and D.
#pragma once
#include <climits>
#include <string>
#include <vector>

#include "Common.h"

namespace redmax {

class Joint {
public:
    enum class Type { Fixed, Translational, Free3DExp };

    std::string _name;
    int _id = 0;
    int _ndof = 0;
    Type _type = Type::Fixed;
    std::vector<int> _index;   // positions of its dofs in the reduced coordinates
    VectorX _q, _qdot;
    dtype _Dr = 0.0;           // damping (Joint.cpp:226-231)
    dtype _joint_limit_lower = (dtype)INT_MIN, _joint_limit_upper = (dtype)INT_MAX, _joint_limit_stiffness = 0.0;

    Matrix4 _E_pj_0, _Q, _E_0j;
    Vector6 _phi;
    MatrixX _S_j, _S_j_dot;                     // 6 x ndof
    std::vector<MatrixX> _dSj_dq, _dSjdot_dq;   // per q_k: 6 x ndof

    Joint(std::string name, int id, Type type, const Matrix3& R, const Vector3& p);

    // Q, S_j, S_j_dot, E_0j, phi from q and qdot, and (derivs) the q-derivatives of
    // S_j and S_j_dot (only those need second derivatives of the rotation)
    void update(bool derivs = true);
    // free3d-exp: r -> (1 - 2 pi / |r|) r while |r| > 1.5 pi, the rotation unchanged and
    // rdot mapped so that the angular velocity is unchanged (JointSphericalExp.cpp:237-262)
    bool reparam();

    // fr += -Dr qdot plus the joint-limit springs (Joint.cpp:226-261)
    void computeJointForce(VectorX& fr) const;
    void computeJointForceWithDerivative(VectorX& fr, MatrixX& Kr, MatrixX& Dr) const;
};

// SO(3) of exponential coordinates r: R = exp([r]x) and the right Jacobian Jr (the
// body angular velocity is Jr(r) rdot), with their first and second partial derivatives
// in r (dR[k] = dR/dr_k, d2R[j][k] = d2R/dr_j dr_k, the same for Jr). Row-major 3 x 3.
// Without `second` the second derivatives are left unset.
struct So3Derivs {
    double R[9], Jr[9];
    double dR[3][9], dJr[3][9];
    double d2R[3][3][9], d2Jr[3][3][9];
};
void so3_derivatives(const double r[3], So3Derivs& out, bool second = true);

}  // namespace redmax
