// Mesh body (simulation/redmax/Body/BodyMeshObj.{h,cpp}): a rigid body read from an OBJ
// file, moved into its principal frame, with its mass properties, its pose relative to
// its joint, its bounding box and its contact points (the vertices, or the adaptive
// sample of BodyMeshObj.cpp:184-238). The base of BodySDFObj.
#pragma once
#include <array>
#include <string>
#include <utility>
#include <vector>

#include "Body/Body.h"
#include "Body/MeshObj.h"

namespace redmax {

class Joint;

class BodyMeshObj : public Body {
public:
    // BodyMeshObj.h: the reference spells the second one OBJ_TO_WOLRD
    enum TransformType { BODY_TO_JOINT, OBJ_TO_WOLRD, OBJ_TO_JOINT };

    std::string _filename;
    Joint* _joint = nullptr;
    std::vector<Vector3> _V;                 // vertices in the body frame
    std::vector<std::array<int, 3>> _F;
    dtype _density = 1.0, _mass = 0.0;
    Vector6 _Inertia;                        // principal moments, then the mass three times
    Matrix4 _E_oi, _E_io;                    // body frame in the file's frame and back
    Matrix4 _E_ji, _E_ij;                    // body frame in its joint's frame and back
    std::pair<Vector3, Vector3> _bounding_box;
    Vector6 _f_external;                     // [torque; force], world frame

    // BodyMeshObj(sim, joint, filename, R, p, transform_type, density, scale, adaptive_sample)
    // (BodyMeshObj.cpp:9-45). The name and the maximal-coordinate index are given later
    // (the scene parser and Robot::init do that in the reference).
    BodyMeshObj(Simulation* sim, Joint* joint, std::string filename, Matrix3 R, Vector3 p,
                TransformType transform_type = BODY_TO_JOINT, dtype density = 1.0, Vector3 scale = Vector3::Ones(),
                bool adaptive_sample = false);
    // a body given by its contact points alone (no mesh, no mass), e.g. an SDF body built
    // from a grid
    BodyMeshObj(Simulation* sim, std::string name, std::vector<Vector3> contact_points, int index0);

    // Body::set_transform (Body.cpp:87-92)
    void set_transform(const Matrix3& R_ji, const Vector3& p_ji);
    std::vector<Vector3> get_vertices() const { return _V; }
    std::vector<std::array<int, 3>> get_faces() const { return _F; }
    // world AABB of the bounding box's corners at the current pose (Body.cpp:579-601)
    std::pair<Vector3, Vector3> get_AABB() const;
    Vector6 get_external_force() const { return _f_external; }
    void set_external_force(const Vector6& f) { _f_external = f; }
    void reset_external_force() { _f_external.setZero(); }
};

}  // namespace redmax
