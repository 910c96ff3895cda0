// Kernel argument blocks and launchers (host side of atac_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "atac_types.h"

namespace atac {

// Tile queues: tiles order[begin[q] .. begin[q+1]) belong to queue q.
struct QueueDev {
    const int32_t* order;
    const int32_t* begin;   // nq + 1 entries
    int32_t nq;
    int32_t pad;
};

// Candidate buffer of a detection wavefront (entries x, y, z, tile offset |
// open << 15): room for what one step adds (<= 64 PPL) on top of < 64 waiting.
__host__ __device__ constexpr int detect_ring_cap(int ppl) { return 64 + 64 * ppl; }

// LDS of the detection kernel per wavefront: the output segment (span u16
// offsets) and the candidate buffer (16 B entries).
__host__ __device__ constexpr size_t detect_lds_per_wave(int span, int ppl) {
    return (size_t)span * 2 + (size_t)detect_ring_cap(ppl) * 16;
}

struct DetectArgs {
    QueueDev q;                 // tiles
    const TileDev* tiles;
    int32_t wave_span;          // output capacity per wavefront: its share of a tile's groups x 256 points
    int32_t all_active;         // every point is a contact (fixed contact lists)
    const InstDev* inst;
    const GridDev* grids;
    uint32_t* act;              // pool (rec_cap entries): offset of an active point within its instance (| kActSlow)
    int32_t rec_cap;
    int32_t slow_cap;           // entries of the slow list
    int32_t chunk_cap;          // entries of the work-item list
    int32_t pad;
    const uint32_t* tile_live;  // n_tiles: bit g set when group g of the tile can hold a candidate (null: all)
    int32_t* tile_active;       // n_tiles
    int32_t* tile_chunk0;       // n_tiles: first work item of the tile
    ChunkDev* chunks;           // work items, pushed per tile
    SlowDev* slow;              // slow list
    int32_t* counters;          // CNT_*
};

struct CullArgs {
    const TileDev* tiles;
    const InstDev* inst;
    const GridDev* grids;
    uint32_t* tile_live;        // n_tiles
    int32_t* counters;          // CNT_WORDS, cleared by the pass (the evaluation's reset)
    int32_t n_tiles;
    int32_t slots;              // lanes per tile: a power of two from 4 to 32, at least the groups of a tile
};

struct SlowArgs {
    const InstDev* inst;
    const GridDev* grids;
    const uint32_t* act;
    void* rec;                  // Rec32 or Rec64: the slow contacts' records are written here
    float* slow_diag;           // 4 per slow contact (FP32 records)
    const SlowDev* slow;
    const int32_t* counters;
    int32_t slow_cap;
    int32_t pad;
};

struct ContactArgs {
    const ChunkDev* chunks;
    const InstDev* inst;
    const GridDev* grids;
    const uint32_t* act;
    void* rec;                  // fast-path records are written here, slow ones read
    const float* slow_diag;
    int32_t* counters;
    int32_t chunk_cap;
    int32_t pad;
    double* partial;            // per work item: fm 12 | p6 12 | Km 144 | Dm 144 (as the mode has them);
                                // the fused pass: fm 12 | kVjpOut
    const double* g_fm;         // the fused pass: d(loss)/d(wrench) per instance (n_inst x 12)
};

struct VjpArgs {
    const ChunkDev* chunks;
    const InstDev* inst;
    const void* rec;
    const float* slow_diag;
    int32_t* counters;
    int32_t chunk_cap;
    int32_t pad;
    const double* g_fm;         // n_inst x 12
    double* partial;            // per work item: kVjpOut
};

// Per-instance sum of the work items' partials, in tile order then item order.
// Component c of the partial row goes to dst[s][i * width[s] + c - begin[s]]
// for the segment s holding it (dst[s] may be nullptr).
struct ReduceArgs {
    const InstDev* inst;
    int32_t n_inst;
    int32_t nacc;
    const int32_t* tile_active;
    const int32_t* tile_chunk0;
    const double* partial;
    int32_t nseg;
    int32_t host_dst;           // the outputs are in host memory: each wave ends with a system-scope fence
    double* dst[8];
    int32_t begin[8];
    int32_t width[8];
    int32_t* n_active;          // optional: active contacts per instance
    const int32_t* counters;    // CNT_OVERFLOW set: every output is NaN (the records were incomplete)
    int32_t* overflow_out;      // optional: receives counters[CNT_OVERFLOW] (host-visible memory allowed)
};

struct PlainArgs {
    const InstDev* inst;
    int32_t i;
    int32_t pad;
    const GridDev* grids;
    double* fm;
    double* p6;
    double* Km;
    double* Dm;
    int32_t* n_active;
    const double* g_fm;         // optional: d(loss)/d fm of the instance (12) ...
    double* grad;               // ... and the adjoint it gives (kVjpOut, as atac_batch_vjp's)
};

struct PointArgs {
    GridDev grid;
    InstDev in;                 // only R2, p2, w2, v2, col_th (and R1 = I, p1 = 0) are used
    const double* xw;           // n x 3
    const double* xwdot;        // n x 3 (collision only)
    int64_t n;
    double* out;                // distance: n; collision: n x 11
    double* out2;               // collision derivatives: n x 135 (nullptr: values only)
};

struct TouchedArgs {
    const InstDev* inst;
    const GridDev* grids;
    uint32_t* bits;
    const int64_t* bit_offset;  // per grid, in 32-bit words
};

hipError_t launch_detect(const DetectArgs& a, int block, int ppl, int grid, size_t lds, hipStream_t st);
const void* detect_kernel_ptr(int block, int ppl);
hipError_t launch_slow(const SlowArgs& a, bool precise, int grid, hipStream_t st);
hipError_t launch_contact(const ContactArgs& a, int mode, bool precise, int grid, hipStream_t st);
const void* contact_kernel_ptr(int mode, bool precise);
hipError_t launch_vjp(const VjpArgs& a, bool precise, int grid, hipStream_t st);
// the wrench and the adjoint of a known cotangent in one pass (no records kept)
hipError_t launch_fused(const ContactArgs& a, bool precise, int grid, hipStream_t st);
const void* vjp_kernel_ptr(bool precise);
hipError_t launch_reduce(const ReduceArgs& a, hipStream_t st);
hipError_t launch_grid_setup(const GridDev& G, float4* cells, hipStream_t st);
// the cell classification of a grid for one col_th (cls_bytes(G) bytes)
hipError_t launch_cls_setup(const GridDev& G, double col_th, uint8_t* cls, hipStream_t st);
// the summed-area table (sat_count(G) entries) of a classification's 2 x 2 x 2 blocks
hipError_t launch_sat_setup(const GridDev& G, const uint8_t* cls, int32_t* sat, hipStream_t st);
hipError_t launch_cull(const CullArgs& a, hipStream_t st);
hipError_t launch_reset(int32_t* c, int n, hipStream_t st);
hipError_t launch_zero(void* p, int64_t n_words, hipStream_t st);
hipError_t launch_plain(const PlainArgs& a, int64_t n_points, bool precise, hipStream_t st);
hipError_t launch_distance(const PointArgs& a, hipStream_t st);
hipError_t launch_collision(const PointArgs& a, bool precise, hipStream_t st);
hipError_t launch_touched(const TouchedArgs& a, int n_inst, hipStream_t st);
hipError_t launch_popcount(const uint32_t* bits, int64_t nwords, unsigned long long* out, hipStream_t st);

}  // namespace atac
