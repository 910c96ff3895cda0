// Device-side data layout shared by the kernels and the C-ABI implementation.
#pragma once
#include <cstdint>

namespace atac {

// One SDF grid in HBM: FP32 values, x fastest: v[i + ni*(j + nj*k)]
// (the order of sdfgen's Array3, simulation/externals/sdfgen/array3.h:70-79).
struct GridDev {
    const float* v;
    int32_t n[3];
    int32_t pad;
    double minb[3];
    double maxb[3];
    double d[3];
    double inv_d[3];
};

// One force object = (contact body points, SDF body grid, poses, parameters).
// Everything the kernels read per instance, in FP64 as the reference keeps it.
struct InstDev {
    // detection map straight into grid index space: c = A * xi + b
    // (A = diag(1/d) R2^T R1, b = diag(1/d) (R2^T (p1 - p2) - min_box))
    double A[9];
    double b[3];
    double cmax[3];     // (max_box - min_box) / d: the inside test in index space
    // poses (row-major R) and body-frame twists [w; v]
    double R1[9], p1[3], w1[3], v1[3];
    double R2[9], p2[3], w2[3], v2[3];
    // force parameters, effective values kn = _kn*_scale*contact_scale, kt likewise
    // (ForceGeneralSDFContact.cpp:155-156)
    double col_th, kn, kt, mu, damping, contact_scale;
    double kn_chain;    // d kn_eff / d kn = _scale * contact_scale (same for kt)
    const float* px;    // this instance's contact points, Morton-sorted SoA FP32
    const float* py;
    const float* pz;
    int64_t pt_count;
    int32_t grid;
    int32_t pad;
    int32_t tile_begin; // first tile of this instance (tiles are contiguous per instance)
    int32_t tile_count;
};

struct TileDev {
    int32_t inst;
    int32_t count;      // points in the tile
    int64_t begin;      // first point, index into the instance's arrays
    int64_t slot;       // first slot of the tile in the batch-wide active list
};

// Active-contact checkpoint, written by the forward pass at slot
// (tile.slot + queue position): (d, grad phi) of the active point; d is kept as
// an unevaluated FP32 pair (d = d_hi + d_lo) so the FP64 depth survives.
struct CkptDev {
    float d_hi, d_lo, g0, g1, g2;
};

constexpr int kFmOut = 12;          // wrench of both bodies
constexpr int kP6Out = 12;          // d fm / d contact_scale (reference p6 column)
constexpr int kKmOut = 144;
constexpr int kDmOut = 144;
constexpr int kVjpOut = 26;         // q1(6) q2(6) phi1(6) phi2(6) kn(1) kt(1)

}  // namespace atac
