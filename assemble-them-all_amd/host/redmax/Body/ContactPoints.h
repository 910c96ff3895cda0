// The contact points of a mesh body (BodyMeshObj::precompute_contact_points,
// simulation/redmax/Body/BodyMeshObj.cpp:184-238).
#pragma once
#include <array>
#include <vector>
#include "Common.h"

namespace redmax {

// Every vertex (adaptive_sample false); with adaptive_sample, every
// sample_rate-th vertex on average, sample_rate = max(floor(nV / 1000), 1),
// picked by the C library's rand() after srand(1000), then on each face that the
// same draw keeps and whose area exceeds 0.5, floor(area / 0.5) points of the R2
// low-discrepancy sequence folded into the triangle. Vertices are V[i], faces
// index into them. Returns the points in that order; sample_rate (may be null)
// receives the rate used.
std::vector<Vector3> precompute_contact_points(const std::vector<Vector3>& V, const std::vector<std::array<int, 3>>& F,
                                               bool adaptive_sample, int* sample_rate = nullptr);

}  // namespace redmax
