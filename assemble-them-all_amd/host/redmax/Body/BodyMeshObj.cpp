#include "Body/BodyMeshObj.h"

#include <cmath>

#include "Body/ContactPoints.h"
#include "Joint/Joint.h"

namespace redmax {

namespace {

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

}  // namespace

BodyMeshObj::BodyMeshObj(Simulation* sim, Joint* joint, std::string filename, Matrix3 R, Vector3 p,
                         TransformType transform_type, dtype density, Vector3 scale, bool adaptive_sample)
    : Body(sim, "", {}, 0), _filename(std::move(filename)), _joint(joint), _density(density) {
    TriMesh m = load_obj(_filename, scale);
    const MassFrame mf = process_mesh(m, density);
    _V = std::move(m.V);
    _F = std::move(m.F);
    _mass = mf.mass;
    _Inertia = mf.inertia;
    _E_oi = mf.E_oi;
    _E_io = einv(_E_oi);
    // the pose relative to the joint by transform type (BodyMeshObj.cpp:22-38)
    Matrix4 Erp = Matrix4::Identity();
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) Erp(i, j) = R(i, j);
        Erp(i, 3) = p(i);
    }
    if (transform_type == BODY_TO_JOINT) {
        _E_ji = Erp;
    } else if (transform_type == OBJ_TO_WOLRD) {
        if (!_joint) throw_error("BodyMeshObj: OBJ_TO_WOLRD needs the joint");
        _E_ji = mul4(mul4(einv(_joint->_E_pj_0), Erp), _E_oi);
    } else {
        // R_ji = R R_oi, p_ji = R p_oi + p
        _E_ji = Matrix4::Identity();
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += R(i, k) * _E_oi(k, j);
                _E_ji(i, j) = s;
            }
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += R(i, k) * _E_oi(k, 3);
            _E_ji(i, 3) = s + p(i);
        }
    }
    _E_ij = einv(_E_ji);
    for (int a = 0; a < 3; ++a) {
        _bounding_box.first(a) = INFINITY;
        _bounding_box.second(a) = -INFINITY;
    }
    for (const auto& v : _V)
        for (int a = 0; a < 3; ++a) {
            _bounding_box.first(a) = std::min(_bounding_box.first(a), v(a));
            _bounding_box.second(a) = std::max(_bounding_box.second(a), v(a));
        }
    _contact_points = precompute_contact_points(_V, _F, adaptive_sample);
}

BodyMeshObj::BodyMeshObj(Simulation* sim, std::string name, std::vector<Vector3> contact_points, int index0)
    : Body(sim, std::move(name), std::move(contact_points), index0) {
    _E_oi = _E_io = _E_ji = _E_ij = Matrix4::Identity();
}

void BodyMeshObj::set_transform(const Matrix3& R_ji, const Vector3& p_ji) {
    _E_ji = Matrix4::Identity();
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) _E_ji(i, j) = R_ji(i, j);
        _E_ji(i, 3) = p_ji(i);
    }
    _E_ij = einv(_E_ji);
}

std::pair<Vector3, Vector3> BodyMeshObj::get_AABB() const {
    std::pair<Vector3, Vector3> box;
    for (int a = 0; a < 3; ++a) {
        box.first(a) = INFINITY;
        box.second(a) = -INFINITY;
    }
    for (int c = 0; c < 8; ++c) {
        const double x[3] = {(c & 1) ? _bounding_box.second(0) : _bounding_box.first(0),
                             (c & 2) ? _bounding_box.second(1) : _bounding_box.first(1),
                             (c & 4) ? _bounding_box.second(2) : _bounding_box.first(2)};
        for (int a = 0; a < 3; ++a) {
            const double y = _E_0i(a, 0) * x[0] + _E_0i(a, 1) * x[1] + _E_0i(a, 2) * x[2] + _E_0i(a, 3);
            box.first(a) = std::min(box.first(a), y);
            box.second(a) = std::max(box.second(a), y);
        }
    }
    return box;
}

}  // namespace redmax
