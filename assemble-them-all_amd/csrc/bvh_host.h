// Host side of the BVH contact variant: the mesh preprocessing (BVH and the
// pseudonormals libigl's per_face / per_edge (uniform) / per_vertex (angle)
// normals give, Body/BVH/BVHEngine.h:33-47) and the launches of bvh.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "bvh.h"

namespace atac {

struct MeshHost {
    std::vector<float4> nodes;
    std::vector<int32_t> tri, F, FE;
    std::vector<double> V, FN, EN, VN;
    double bmin[3], bmax[3];
    int64_t ne = 0;
};

// Preprocesses a triangle mesh (V: nv x 3, F: nf x 3 vertex indices); returns ""
// or the reason the mesh is refused.
std::string mesh_build(const double* V, int64_t nv, const int64_t* F, int64_t nf, MeshHost& out);

// BodyBVHObj::distance_normal_parallel (BodyBVHObj.cpp:322-356) for n world points.
struct MeshQueryArgs {
    MeshDev M;
    double R2[9], p2[3];
    double col_th;
    const double* xw;     // n x 3
    int64_t n;
    double* d;            // n
    double* normal;       // n x 3 (nullptr: not written)
    int32_t* face;        // n (nullptr: not written)
};

// One ForceGeneralBVHContact evaluation: detection and the wrench (+ p6, + Km / Dm).
struct BvhForceArgs {
    MeshDev M;
    InstDev in;
    const double* xi;     // n x 3 contact points of the contact body (body frame)
    int64_t n;
    uint32_t flags;       // ATAC_EVAL_* (DERIV, P6)
    double* out;          // 12 fm, 12 p6, 144 Km, 144 Dm
    int64_t* n_active;
    double* part;         // per workgroup: kBvhOut partial sums
    int64_t* part_cnt;    // per workgroup: contacts
};
constexpr int kBvhOut = 12 + 12 + 288;
constexpr int kBvhBlockPoints = 2048;   // points per force workgroup
inline int64_t bvh_force_blocks(int64_t n) { return n > 0 ? (n + kBvhBlockPoints - 1) / kBvhBlockPoints : 1; }

hipError_t launch_mesh_distance(const MeshQueryArgs& a, hipStream_t st);
hipError_t launch_bvh_force(const BvhForceArgs& a, hipStream_t st);

// Fixed scratch of atac_bvh_force_noalloc: room for the partial rows and counts of
// kBvhScratchBlocks force workgroups (kBvhNoAllocMaxPoints points), per device.
constexpr int64_t kBvhNoAllocMaxPoints = (int64_t)1 << 20;
constexpr int64_t kBvhScratchBlocks = kBvhNoAllocMaxPoints / kBvhBlockPoints;
// Looks up the device's copy of the scratch array; `device` must be current.
// Idempotent per device.
hipError_t bvh_scratch_setup(int device);
// The device address from bvh_scratch_setup, or nullptr before it.
void* bvh_scratch_address(int device);

}  // namespace atac
