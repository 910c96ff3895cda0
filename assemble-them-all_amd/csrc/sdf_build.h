// Mesh -> signed distance grid on the GPU (the role of sdfgen::make_level_set3).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace atac {

struct LsArgs {
    const float* x;           // 3 * nv vertex coordinates
    const uint32_t* tri;      // 3 * nf vertex indices
    int64_t nf;
    float o[3];               // grid origin (node (0, 0, 0))
    float dx, dy, dz;
    int32_t ni, nj, nk;
    int32_t band;             // exact band, in cells
    float phi0;               // initial (upper-bound) distance
    unsigned long long* best; // per node: (distance bits << 32) | triangle, min over the band
    float* phi;               // per node: distance, then signed distance
    int32_t* closest;         // per node: closest triangle so far (-1: none)
    int32_t* count;           // per node: x-ray crossings in (i - 1, i]
};

// Launches the whole construction on `st`: band distances and crossing counts,
// 2 x 8 fast-sweeping passes (one launch per diagonal plane), signs.
hipError_t ls_build(const LsArgs& a, hipStream_t st);

}  // namespace atac
