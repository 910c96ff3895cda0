// Kernel argument blocks and launchers (host side of atac_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "atac_types.h"

namespace atac {

// Work queues: items order[begin[q] .. begin[q+1]) belong to queue q.
struct QueueDev {
    const int32_t* order;
    const int32_t* begin;   // nq + 1 entries
    int32_t* counter;       // nq heads (reset before each launch)
    int32_t nq;
    int32_t pad;
};

struct DetectArgs {
    QueueDev q;                 // tiles
    const TileDev* tiles;
    int32_t wave_seg;           // LDS queue entries per wavefront
    int32_t all_active;         // every point is a contact (fixed contact lists)
    const InstDev* inst;
    const GridDev* grids;
    uint32_t* act_off;          // active offsets, tile slice starts at tile.slot
    int32_t* tile_active;       // n_tiles
};

struct ContactArgs {
    QueueDev q;                 // instances
    const TileDev* tiles;
    const InstDev* inst;
    const GridDev* grids;
    const uint32_t* act_off;
    const int32_t* tile_active;
    CkptDev* ck_dg;             // per active point (d, grad phi); nullptr: not kept
    double* fm;                 // outputs per instance (any may be nullptr)
    double* p6;
    double* Km;
    double* Dm;
    int32_t* n_active;
};

struct VjpArgs {
    QueueDev q;                 // instances
    const TileDev* tiles;
    const InstDev* inst;
    const GridDev* grids;
    const int32_t* tile_active;
    const uint32_t* act_off;
    const CkptDev* ck_dg;
    const double* g_fm;         // n_inst x 12
    double* g_q1;
    double* g_q2;
    double* g_phi1;
    double* g_phi2;
    double* g_kn;
    double* g_kt;
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

hipError_t launch_detect(const DetectArgs& a, int block, int grid, size_t lds, hipStream_t st);
const void* detect_kernel_ptr(int block);
hipError_t launch_contact(const ContactArgs& a, int mode, bool precise, int grid, hipStream_t st);
const void* contact_kernel_ptr(int mode, bool precise);
hipError_t launch_vjp(const VjpArgs& a, bool precise, int grid, hipStream_t st);
const void* vjp_kernel_ptr(bool precise);
hipError_t launch_reset(int32_t* c, int n, hipStream_t st);
hipError_t launch_zero(void* p, int64_t n_words, hipStream_t st);
hipError_t launch_plain(const PlainArgs& a, int64_t n_points, bool precise, hipStream_t st);
hipError_t launch_distance(const PointArgs& a, hipStream_t st);
hipError_t launch_collision(const PointArgs& a, bool precise, hipStream_t st);
hipError_t launch_touched(const TouchedArgs& a, int n_inst, hipStream_t st);
hipError_t launch_popcount(const uint32_t* bits, int64_t nwords, unsigned long long* out, hipStream_t st);

}  // namespace atac
