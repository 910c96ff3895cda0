// The triangle mesh of a mesh body, read from a Wavefront OBJ file, with its mass
// properties and body frame: BodyMeshObj's load_mesh / process_mesh /
// compute_mass_property (simulation/redmax/Body/BodyMeshObj.cpp:47-177).
#pragma once
#include <array>
#include <string>
#include <vector>

#include "Common.h"

namespace redmax {

struct TriMesh {
    std::vector<Vector3> V;
    std::vector<std::array<int, 3>> F;
};

// load_mesh (BodyMeshObj.cpp:47-68): every vertex, scaled per axis, and the faces of the
// file's first shape (a shape ends where a 'g' or 'o' line follows faces), polygons split
// into triangle fans around their first corner; 1-based, negative (relative) indices and
// v/vt/vn corner tokens accepted. Throws when the file cannot be read or an index is out
// of range.
TriMesh load_obj(const std::string& filename, const Vector3& scale);

// compute_mass_property (BodyMeshObj.cpp:128-177): volume, centre of mass and inertia
// tensor of unit mass of a closed, outward-oriented mesh, summed over the signed
// tetrahedra the faces span with the origin (Melax's volume integration).
void compute_mass_property(const TriMesh& m, dtype& volume, Vector3& COM, Matrix3& I);

// Eigenvalues of a symmetric 3 x 3 matrix in ascending order and unit eigenvectors as
// the columns of `vectors` (cyclic Jacobi rotations). Each eigenvector's sign makes its
// component of largest magnitude positive (the first of equal ones).
void symmetric_eigen3(const Matrix3& A, Vector3& values, Matrix3& vectors);

struct MassFrame {
    dtype volume = 0, mass = 0;
    Vector6 inertia;   // principal moments (ascending), then the mass three times
    Matrix4 E_oi;      // body frame in the mesh file's frame: principal axes, origin at the COM
};

// process_mesh (BodyMeshObj.cpp:70-126): mass = volume * density, principal moments of
// mass * I, the frame of the principal axes (the third flipped if the three are
// left-handed) at the centre of mass, and the mesh moved into that frame. The
// reference takes the axes from Eigen's SelfAdjointEigenSolver, whose eigenvector signs
// are not specified; here they follow symmetric_eigen3's rule.
MassFrame process_mesh(TriMesh& m, dtype density);

}  // namespace redmax
