// Triangle-mesh body of the BVH contact variant with the reference's query methods
// (simulation/redmax/Body/BodyBVHObj.h:13-80): distance, distance_normal and their
// parallel forms over an exact signed distance (closest point on the mesh, sign
// from the angle-weighted pseudonormal of the closest feature; libigl's
// signed_distance_pseudonormal in the reference), evaluated on the GPU through the
// atac C ABI. The mesh is given in the body frame; its vertices' bounding box
// filters the queries as BodyMeshObj::filter does (BodyMeshObj.cpp:278-296).
#pragma once
#include <array>
#include <string>
#include <utility>
#include <vector>
#include "Body/Body.h"
#include "atac.h"

namespace redmax {

class BodyBVHObj : public Body {
public:
    std::vector<Vector3> _BVH_V;                // BVH mesh vertices (body frame)
    std::vector<std::array<int, 3>> _BVH_F;     // BVH mesh triangles
    dtype _col_th = 0;                          // collision threshold

    BodyBVHObj(Simulation* sim, std::string name, std::vector<Vector3> V, std::vector<std::array<int, 3>> F,
               dtype col_th, std::vector<Vector3> contact_points, int index0);
    ~BodyBVHObj() override;

    // the vertices' bounding box (BodyMeshObj::_bounding_box)
    std::pair<Vector3, Vector3> get_AABB() const;

    // signed distance + col_th of the world point xw (10 outside the bounding box;
    // BodyBVHObj.cpp:249-260, :322-356)
    void distance(Vector3 xw, dtype& d);
    // the same and its gradient, the pseudonormal of the closest feature (0 outside)
    void distance_normal(Vector3 xw, dtype& d, Vector3& dd_dx);
    void distance_parallel(const std::vector<Vector3>& xw, std::vector<dtype>& d);
    void distance_normal_parallel(const std::vector<Vector3>& xw, std::vector<dtype>& d, std::vector<Vector3>& dd_dx);

    // the mesh with its BVH in HBM, created on first use; call after changing the mesh
    atac_mesh device_mesh();
    void release_device_mesh();

private:
    atac_mesh mesh_ = nullptr;
};

}  // namespace redmax
