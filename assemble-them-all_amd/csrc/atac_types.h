// Device-side data layout shared by the kernels and the C-ABI implementation.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace atac {

// One SDF grid in HBM: FP32 values, x fastest: v[i + ni*(j + nj*k)]
// (the order of sdfgen's Array3, simulation/externals/sdfgen/array3.h:70-79).
struct GridDev {
    const float* v;
    int32_t n[3];
    int32_t fast;       // 1: cells and blk below exist (every grid of the library has them)
    double minb[3];
    double maxb[3];
    double d[3];
    double inv_d[3];
    // corner-packed copy: cell (i, j, k) < n holds the 8 nodes a point in it
    // interpolates (v000 v100 v010 v110 | v001 v101 v011 v111, offsets clamped at
    // the last node) in 2 x float4, node-index order
    const float4* cells;

    // the FD stencil's reach in cell units, 1e-5 * inv_d + 1e-9 (fd_m), and 1 - fd_m
    double fd_m[3];
    double fd_1m[3];
};

// One force object = (contact body points, SDF body grid, poses, parameters).
// Everything the kernels read per instance, in FP64 as the reference keeps it.
struct InstDev {
    // detection map straight into grid index space, shifted by one cell:
    // c' = A * xi + bs, c = c' - 1 (A = diag(1/d) R2^T R1,
    // bs = diag(1/d) (R2^T (p1 - p2) - min_box) + 1)
    double A[9];
    double bs[3];
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
    // the same poses, twists and parameters rounded to FP32 for the FP32 contact
    // arithmetic (read as wave-uniform scalars)
    float fR1[9], fR2[9], fw1[3], fv1[3], fw2[3], fv2[3];
    float fkn, fkt, fmu, fdamping, fcontact_scale;
    float fpad0;
    // per-cell classification of the instance's grid for its col_th (cls_fill;
    // bricked, cls_offset): 0 no active point, 1 only active points, 2 open
    const uint8_t* cls;
    // the summed-area table of that classification's blocks (sat_*), and the
    // boxes of the instance's groups of kGroup points (8 floats each: lo, 0, hi, 0)
    const int32_t* csat;
    const float* gbox;
    double dpad;
    float fpad[4];
};
static_assert(sizeof(InstDev) == 704, "InstDev layout");

struct TileDev {
    int32_t inst;
    int32_t count;      // points in the tile
    int64_t begin;      // first point, index into the instance's arrays
    int64_t slot;       // first slot of the tile in the batch-wide active list
    int32_t grid;       // the instance's grid (read with the instance, not after it)
    int32_t pad;
};

// Active-contact record, one per point with d < 0, at the slot detection gave
// it in the batch's record pool (a tile's contacts in ascending point order in
// one range): written by the contact pass (the slow pass for contacts whose FD
// stencil leaves the cell) and read by the contact and adjoint passes (it is
// also the adjoint's checkpoint: the same d, grad phi and Hessian are used by both).
//   a = (xi1.x, xi1.y, xi1.z, bits(point offset within the instance))
//   b = (d, g.x, g.y, g.z)            d = query_SDF(x) + col_th, g = grad phi
//   c = (Hxy, Hxz, Hyz, bits(slow))   slow = -1, or the index of the contact in
//                                     the slow list (its FD stencil leaves the
//                                     cell; Hxx Hyy Hzz then sit in slow_diag)
// FP32 record (48 B), for the FP32 contact arithmetic.
struct Rec32 {
    float a[4];
    float b[4];
    float c[4];
};

// FP64 record (96 B), for the FP64 contact arithmetic: (xi1, point offset) as in
// Rec32, then d, grad phi and the six Hessian entries (xx yy zz xy xz yz).
struct Rec64 {
    float a[4];
    double d;
    double g[3];
    double H[6];
};

// A work item of the record, contact and adjoint passes: records
// [rec0, rec0 + n) (n <= kChunk) of one tile of instance `inst`.
struct ChunkDev {
    int32_t inst;
    int32_t n;
    int64_t rec0;
};

constexpr int kChunk = 256;         // records per work item

// Active list (the record pool's slots): the point's offset within its
// instance, with kActSlow set when its FD stencil leaves the cell (its record
// comes from the slow pass).
constexpr uint32_t kActSlow = 0x80000000u;

// Slow-list entry: the record slot and the instance of a contact whose FD
// stencil leaves its cell (its record is written by the slow pass).
struct SlowDev {
    uint32_t slot;
    int32_t inst;
};

// Device counters of one evaluation (reset before detection).
enum {
    // records reserved from the pool and work items pushed, the low and high
    // word of one 64-bit counter (a tile takes both with one atomic); both may
    // run past their capacity
    CNT_RECS = 0,
    CNT_CHUNKS = 1,
    CNT_SLOW = 2,                   // contacts in the slow list (also past its capacity)
    CNT_OVERFLOW = 3,               // non-zero: the record pool or the slow list was too small
    CNT_WORDS = 4
};

constexpr int kFmOut = 12;          // wrench of both bodies
constexpr int kP6Out = 12;          // d fm / d contact_scale (reference p6 column)
constexpr int kKmOut = 144;
constexpr int kDmOut = 144;
constexpr int kVjpOut = 26;         // q1(6) q2(6) phi1(6) phi2(6) kn(1) kt(1)

}  // namespace atac
