// Shared by the two drop-in translation units of this directory, which replace
// simulation/redmax/Body/BodySDFObj.cpp and
// simulation/redmax/Force/ForceGeneralSDFContact.cpp inside the reference's own
// build and are compiled against its unchanged headers.
#pragma once
#include <string>

#include "atac.h"

namespace redmax {

class BodySDFObj;

// The GPU copy of an SDF body's grid (_SDF, _min_box, _max_box, _dx/_dy/_dz),
// created on first use and again whenever the body's grid storage, size, box or
// spacing changed. Throws (throw_error) when the library refuses the grid.
atac_grid atac_route_a_grid(const BodySDFObj* body);

// throw_error with the library's message when rc is not ATAC_OK.
void atac_route_a_check(int rc, const std::string& what);

}  // namespace redmax
