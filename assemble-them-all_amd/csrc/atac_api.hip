// C ABI of libatac_contact.so (declared in include/atac.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "atac.h"
#include "atac_compose.h"
#include "sdf_build.h"
#include "bvh_host.h"
#include "atac_kernels.h"
#include "atac_types.h"

using namespace atac;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define ATAC_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return fail(ATAC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return ATAC_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(ATAC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return ATAC_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

}  // namespace

struct atac_grid_s {
    GridDev dev;
    float* d_values = nullptr;
    float4* d_cells = nullptr;

    int64_t count = 0;
};

struct atac_points_s {
    float* d_x = nullptr;
    float* d_y = nullptr;
    float* d_z = nullptr;
    int64_t n = 0;
    std::vector<int64_t> seg;   // n_segments + 1 offsets
    std::vector<int64_t> dseg;  // device offset of each segment (multiples of 4: 16-byte aligned)
    int64_t dn = 0;             // device floats per coordinate (segments padded to multiples of 4)
    std::vector<int64_t> perm;  // sorted position -> original index within the segment
    float* d_gbox = nullptr;    // per group of kGroup sorted points of a segment: lo xyz, 0, hi xyz, 0
    std::vector<int64_t> gseg;  // first group of each segment
};

struct atac_batch_s {
    // the cell classification of one grid for one col_th (as the detection rounds it)
    struct ClsTable {
        int32_t grid;
        float cth;
        uint8_t* d;
        int32_t* sat;   // summed-area table of its 2 x 2 x 2 blocks (group culling)
    };
    std::vector<ClsTable> cls;
    int device = 0;
    int n_cu = 0;
    size_t lds_limit = 0;
    std::vector<atac_grid> grids;
    std::vector<atac_points> pts;
    std::vector<atac_instance> inst;
    std::vector<InstDev> inst_dev;
    InstDev* d_inst = nullptr;
    GridDev* d_grids = nullptr;
    atac_tuning tune{};
    std::vector<TileDev> tiles;
    TileDev* d_tiles = nullptr;
    int32_t* d_order = nullptr;       // tiles, queue-partitioned
    int32_t* d_tq_begin = nullptr;    // nq + 1
    int32_t nq = 1;
    int32_t n_tiles = 0;
    int64_t n_slots = 0;              // points of the batch (queries per evaluation)
    int64_t rec_cap = 0;              // record pool: the bound on active contacts per evaluation
    int64_t slow_cap = 0;             // slow list entries
    int64_t chunk_cap = 0;            // work items if the pool were full
    int32_t* d_tile_active = nullptr;
    uint32_t* d_tile_live = nullptr;  // per tile: the groups the culling pass keeps
    bool last_culled = false;         // the last evaluation enqueued ran the culling pass
    int32_t* d_tile_chunk0 = nullptr;
    ChunkDev* d_chunks = nullptr;
    int32_t* d_counters = nullptr;    // CNT_WORDS
    // per-evaluation buffers, allocated on first use
    uint32_t* d_act = nullptr;        // rec_cap: active point offsets, a tile's in one range
    Rec32* d_rec32 = nullptr;         // rec_cap
    Rec64* d_rec64 = nullptr;         // rec_cap
    SlowDev* d_slow = nullptr;        // slow_cap
    float* d_slow_diag = nullptr;     // 4 x slow_cap
    double* d_part = nullptr;         // chunk_cap x part_width
    int64_t part_width = 0;
    int grid_detect = 0, grid_contact[2][3] = {{0, 0, 0}, {0, 0, 0}}, grid_vjp[2] = {0, 0};
    int grid_slow = 0;
    size_t lds_detect = 0;
    int wave_span = 0;
    bool rec_valid = false;           // records of the last evaluation are in place
    bool graph_fused = false;         // the captured graph is a fused pass (it leaves no records)
    bool rec_precise = false;
    bool rec_captured = false;        // the captured graph's evaluation writes records when it is launched
    bool cap_precise = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t cap_stream = nullptr;
    // device staging for the synchronous host-buffer entry points
    double* h_out = nullptr;      // fm 12 | p6 12 | Km 144 | Dm 144 per instance
    int32_t* h_nact = nullptr;
    double* h_gfm = nullptr;      // 12 per instance
    double* h_grad = nullptr;     // 26 per instance
    // pinned host staging of the host-buffer entry points' results: the device-to-host
    // copies run as DMA without the runtime's pageable staging, then one wait
    char* pin = nullptr;          // 312 doubles per instance | n_active per instance | overflow flag
    InstDev* pin_inst = nullptr;  // pinned instance records for atac_batch_eval_instances_host
    // graphs of atac_batch_eval_instances_host per flag combination, valid while no device
    // buffer has been (re)allocated since their capture (buf_gen)
    struct HostGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
        uint64_t gen = 0;
    };
    std::map<uint32_t, HostGraph> host_graphs;
    uint64_t buf_gen = 1;
    // optional per-stage timing: event pairs recorded around each kernel launch
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;  // stage, (start, stop)
};

namespace {
enum { ST_DETECT = 0, ST_SLOW = 1, ST_CONTACT = 2, ST_REDUCE = 3, ST_VJP = 4, ST_VJP_REDUCE = 5, ST_COUNT = 6 };

hipEvent_t take_event(atac_batch b) {
    if (!b->ev_pool.empty()) {
        hipEvent_t e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records a start event on st when timing is enabled; returns it (or nullptr).
hipEvent_t stage_begin(atac_batch b, hipStream_t st) {
    if (!b->timing) return nullptr;
    hipEvent_t e = take_event(b);
    if (e) (void)hipEventRecord(e, st);
    return e;
}

void stage_end(atac_batch b, int stage, hipEvent_t start, hipStream_t st) {
    if (!b->timing || !start) return;
    hipEvent_t e = take_event(b);
    if (!e) return;
    (void)hipEventRecord(e, st);
    b->ev_used.push_back({stage, {start, e}});
}
}  // namespace

extern "C" {

const char* atac_last_error(void) { return g_err.c_str(); }

int atac_version(void) { return 100; }

int atac_device_init(int32_t device, int32_t* n_cu) {
    int count = 0;
    ATAC_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(ATAC_ERR_INVALID, "atac_device_init: no such device");
    ATAC_HIP(hipSetDevice(device));
    ATAC_HIP(bvh_scratch_setup(device));   // the device's fixed scratch (atac_bvh_force_noalloc)
    if (n_cu) {
        hipDeviceProp_t prop;
        ATAC_HIP(hipGetDeviceProperties(&prop, device));
        *n_cu = prop.multiProcessorCount;
    }
    return ATAC_OK;
}

// ---------------------------------------------------------------- grids
int atac_grid_create(const float* values, int32_t ni, int32_t nj, int32_t nk, const double min_box[3],
                     const double max_box[3], const double dxyz[3], atac_grid* out) {
    if (!values || !out || !min_box || !max_box || !dxyz) return fail(ATAC_ERR_INVALID, "atac_grid_create: null argument");
    if (ni < 1 || nj < 1 || nk < 1) return fail(ATAC_ERR_INVALID, "atac_grid_create: empty grid");
    for (int a = 0; a < 3; ++a)
        if (!(dxyz[a] > 0.0) || !(max_box[a] >= min_box[a]))
            return fail(ATAC_ERR_INVALID, "atac_grid_create: spacing must be > 0 and max_box >= min_box");
    // size limits, checked before anything is allocated or read: the detection
    // addresses the corner-packed cells (32 B per node) with 32-bit byte offsets
    // and forms the classification's brick index with 24-bit products
    if ((int64_t)ni * nj * nk * 32 >= ((int64_t)1 << 32))
        return fail(ATAC_ERR_INVALID, "atac_grid_create: grids hold fewer than 2^27 nodes");
    if ((int64_t)((nj + 5) / 4) * ((nk + 5) / 4) >= (1 << 24) || ni >= (1 << 24) || nj >= (1 << 24) || nk >= (1 << 24))
        return fail(ATAC_ERR_INVALID, "atac_grid_create: each n and (n_y + 2)(n_z + 2) / 16 stay below 2^24");
    auto* g = new atac_grid_s();
    g->count = (int64_t)ni * nj * nk;
    int rc = dalloc(&g->d_values, (size_t)g->count);
    if (rc) { delete g; return rc; }
    hipError_t e = hipMemcpy(g->d_values, values, sizeof(float) * g->count, hipMemcpyHostToDevice);
    if (e != hipSuccess) { dfree(g->d_values); delete g; return fail(ATAC_ERR_HIP, hipGetErrorString(e)); }
    GridDev& G = g->dev;
    G.v = g->d_values;
    G.n[0] = ni; G.n[1] = nj; G.n[2] = nk;
    G.fast = 0; G.cells = nullptr;
    for (int a = 0; a < 3; ++a) {
        G.minb[a] = min_box[a];
        G.maxb[a] = max_box[a];
        G.d[a] = dxyz[a];
        G.inv_d[a] = 1.0 / dxyz[a];
    }
    set_fd_reach(G);
    {
        // corner-packed cells, one per node (the last layer clamped), addressed
        // by 32-bit byte offsets in the detection kernel
        const int64_t nc = (int64_t)ni * nj * nk;
        rc = dalloc(&g->d_cells, (size_t)(2 * nc));
        if (!rc) {
            e = launch_grid_setup(G, g->d_cells, nullptr);
            if (e == hipSuccess) e = hipDeviceSynchronize();
            if (e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
        }
        if (rc) { dfree(g->d_values); dfree(g->d_cells); delete g; return rc; }
        G.cells = g->d_cells;
        G.fast = 1;
    }
    *out = g;
    return ATAC_OK;
}

int atac_grid_destroy(atac_grid g) {
    if (!g) return ATAC_OK;
    dfree(g->d_values); dfree(g->d_cells);
    delete g;
    return ATAC_OK;
}

// ---------------------------------------------------------------- points
static uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

int atac_points_create(const double* xyz, int64_t n, const int64_t* seg_offsets, int32_t n_segments,
                       double sort_cell, atac_points* out) {
    if (!out || n < 0 || (n > 0 && !xyz)) return fail(ATAC_ERR_INVALID, "atac_points_create: bad arguments");
    std::vector<int64_t> seg;
    if (seg_offsets) {
        if (n_segments < 1) return fail(ATAC_ERR_INVALID, "atac_points_create: n_segments < 1");
        seg.assign(seg_offsets, seg_offsets + n_segments + 1);
        if (seg.front() != 0 || seg.back() != n) return fail(ATAC_ERR_INVALID, "atac_points_create: offsets must span [0, n]");
        for (int s = 0; s < n_segments; ++s)
            if (seg[s + 1] < seg[s]) return fail(ATAC_ERR_INVALID, "atac_points_create: offsets not ascending");
    } else {
        seg = {0, n};
    }
    auto* P = new atac_points_s();
    P->n = n;
    P->seg = seg;
    P->perm.resize(n);
    P->dseg.resize(seg.size());
    int64_t dn = 0;
    for (size_t s = 0; s + 1 < seg.size(); ++s) {
        P->dseg[s] = dn;
        dn += (seg[s + 1] - seg[s] + 3) / 4 * 4;
    }
    P->dseg.back() = dn;
    P->dn = dn;
    std::vector<float> hx(dn, 0.0f), hy(dn, 0.0f), hz(dn, 0.0f);
    std::vector<std::pair<uint64_t, int64_t>> keys;
    for (size_t s = 0; s + 1 < seg.size(); ++s) {
        const int64_t b = seg[s], e = seg[s + 1], m = e - b, db = P->dseg[s];
        if (m == 0) continue;
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = b; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], xyz[3 * i + a]);
                hi[a] = std::max(hi[a], xyz[3 * i + a]);
            }
        double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        double cell = sort_cell > 0.0 ? sort_cell : (ext > 0.0 ? ext / 1024.0 : 1.0);
        keys.resize(m);
        for (int64_t i = 0; i < m; ++i) {
            uint64_t q[3];
            for (int a = 0; a < 3; ++a) {
                double c = std::floor((xyz[3 * (b + i) + a] - lo[a]) / cell);
                c = std::min(std::max(c, 0.0), 2097151.0);
                q[a] = (uint64_t)c;
            }
            keys[i] = {spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2), i};
        }
        std::sort(keys.begin(), keys.end());
        for (int64_t i = 0; i < m; ++i) {
            const int64_t src = b + keys[i].second;
            hx[db + i] = (float)xyz[3 * src];
            hy[db + i] = (float)xyz[3 * src + 1];
            hz[db + i] = (float)xyz[3 * src + 2];
            P->perm[b + i] = keys[i].second;
        }
    }
    // the groups' boxes (NaN for a group with a non-finite coordinate: never dropped)
    P->gseg.resize(seg.size());
    int64_t ngr = 0;
    for (size_t s = 0; s + 1 < seg.size(); ++s) {
        P->gseg[s] = ngr;
        ngr += (seg[s + 1] - seg[s] + kGroup - 1) / kGroup;
    }
    P->gseg.back() = ngr;
    std::vector<float> gbox((size_t)ngr * 8, 0.0f);
    for (size_t s = 0; s + 1 < seg.size(); ++s) {
        const int64_t m = seg[s + 1] - seg[s], db = P->dseg[s];
        for (int64_t g = 0; g * kGroup < m; ++g) {
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            bool finite = true;
            for (int64_t i = g * kGroup; i < std::min<int64_t>(m, (g + 1) * kGroup); ++i) {
                const float c[3] = {hx[db + i], hy[db + i], hz[db + i]};
                for (int a = 0; a < 3; ++a) {
                    finite = finite && std::isfinite(c[a]);
                    lo[a] = std::min(lo[a], c[a]);
                    hi[a] = std::max(hi[a], c[a]);
                }
            }
            float* o = &gbox[(size_t)(P->gseg[s] + g) * 8];
            for (int a = 0; a < 3; ++a) {
                o[a] = finite ? lo[a] : NAN;
                o[4 + a] = finite ? hi[a] : NAN;
            }
        }
    }
    int rc = dalloc(&P->d_x, dn);
    if (!rc) rc = dalloc(&P->d_y, dn);
    if (!rc) rc = dalloc(&P->d_z, dn);
    if (!rc) rc = dalloc(&P->d_gbox, gbox.size());
    if (!rc && !gbox.empty()) {
        hipError_t e = hipMemcpy(P->d_gbox, gbox.data(), sizeof(float) * gbox.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
    }
    if (!rc && dn > 0) {
        hipError_t e = hipMemcpy(P->d_x, hx.data(), sizeof(float) * dn, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->d_y, hy.data(), sizeof(float) * dn, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(P->d_z, hz.data(), sizeof(float) * dn, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
    }
    if (rc) {
        dfree(P->d_x); dfree(P->d_y); dfree(P->d_z); dfree(P->d_gbox);
        delete P;
        return rc;
    }
    *out = P;
    return ATAC_OK;
}

int atac_points_destroy(atac_points p) {
    if (!p) return ATAC_OK;
    dfree(p->d_x); dfree(p->d_y); dfree(p->d_z); dfree(p->d_gbox);
    delete p;
    return ATAC_OK;
}

int atac_points_permutation(atac_points p, int32_t segment, int64_t* sorted_to_original, int64_t capacity) {
    if (!p || !sorted_to_original) return fail(ATAC_ERR_INVALID, "atac_points_permutation: null argument");
    if (segment < 0 || segment + 1 >= (int32_t)p->seg.size()) return fail(ATAC_ERR_INVALID, "atac_points_permutation: bad segment");
    const int64_t b = p->seg[segment], m = p->seg[segment + 1] - b;
    if (capacity < m) return fail(ATAC_ERR_INVALID, "atac_points_permutation: capacity too small");
    std::copy(p->perm.begin() + b, p->perm.begin() + b + m, sorted_to_original);
    return ATAC_OK;
}

// ---------------------------------------------------------------- batches
static void compose_for(atac_batch b, const atac_instance& s, InstDev& d) {
    const atac_points_s& P = *b->pts[s.points];
    const int64_t beg = P.dseg[s.segment];
    compose_instance(s, b->grids[s.grid]->dev, P.d_x + beg, P.d_y + beg, P.d_z + beg,
                     P.seg[s.segment + 1] - P.seg[s.segment], d);
    d.gbox = P.d_gbox ? P.d_gbox + 8 * P.gseg[s.segment] : nullptr;
}

// A classification table for every (grid, col_th) the instances use, built on
// first use (synchronously) and set in their device records; tables no longer
// used are freed once the device is idle.
static int ensure_cls(atac_batch b) {
    std::vector<char> used(b->cls.size(), 0);
    for (size_t i = 0; i < b->inst_dev.size(); ++i) {
        InstDev& d = b->inst_dev[i];
        const float cth = (float)d.col_th;
        size_t t = 0;
        while (t < b->cls.size() && !(b->cls[t].grid == d.grid && __builtin_memcmp(&b->cls[t].cth, &cth, sizeof cth) == 0)) ++t;
        if (t == b->cls.size()) {
            const GridDev& G = b->grids[d.grid]->dev;
            atac_batch_s::ClsTable ct{d.grid, cth, nullptr, nullptr};
            int rc = dalloc(&ct.d, (size_t)cls_bytes(G));
            if (!rc) rc = dalloc(&ct.sat, sat_count(G));
            if (rc) { dfree(ct.d); return rc; }
            hipError_t e = launch_cls_setup(G, d.col_th, ct.d, nullptr);
            if (e == hipSuccess) e = launch_sat_setup(G, ct.d, ct.sat, nullptr);
            if (e == hipSuccess) e = hipDeviceSynchronize();
            if (e != hipSuccess) { dfree(ct.d); dfree(ct.sat); return fail(ATAC_ERR_HIP, hipGetErrorString(e)); }
            b->cls.push_back(ct);
            used.push_back(0);
        }
        used[t] = 1;
        d.cls = b->cls[t].d;
        d.csat = b->cls[t].sat;
    }
    bool idle = false;
    for (size_t t = b->cls.size(); t-- > 0;) {
        if (used[t]) continue;
        if (!idle) { ATAC_HIP(hipDeviceSynchronize()); idle = true; }   // an earlier evaluation may still read it
        dfree(b->cls[t].d);
        dfree(b->cls[t].sat);
        b->cls.erase(b->cls.begin() + t);
    }
    return ATAC_OK;
}

static int check_instances(atac_batch b, const atac_instance* inst, int32_t n) {
    for (int32_t i = 0; i < n; ++i) {
        const atac_instance& s = inst[i];
        if (s.grid < 0 || s.grid >= (int32_t)b->grids.size()) return fail(ATAC_ERR_INVALID, "instance " + std::to_string(i) + ": bad grid index");
        if (s.points < 0 || s.points >= (int32_t)b->pts.size()) return fail(ATAC_ERR_INVALID, "instance " + std::to_string(i) + ": bad point-set index");
        if (s.segment < 0 || s.segment + 1 >= (int32_t)b->pts[s.points]->seg.size())
            return fail(ATAC_ERR_INVALID, "instance " + std::to_string(i) + ": bad segment");
        if (!(s.contact_scale != 0.0)) return fail(ATAC_ERR_INVALID, "instance " + std::to_string(i) + ": contact_scale must be non-zero");
    }
    return ATAC_OK;
}

constexpr int kXcds = 8;   // MI355X: 8 XCDs, each with its own L2

static void release_plan(atac_batch b) {
    dfree(b->d_tiles); dfree(b->d_order); dfree(b->d_tq_begin); dfree(b->d_counters);
    dfree(b->d_tile_active); dfree(b->d_tile_chunk0); dfree(b->d_chunks); dfree(b->d_tile_live);
    dfree(b->d_act); dfree(b->d_rec32); dfree(b->d_rec64); dfree(b->d_slow); dfree(b->d_slow_diag);
    dfree(b->d_part);
    b->part_width = 0;
    b->rec_valid = false;
    b->rec_captured = false;
    if (b->graph_exec) { (void)hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

// Width of a work item's partial row for an evaluation mode (0 wrench, 1 + p6,
// 2 + Km/Dm) and for the adjoint (mode 3).
// partial row per work item: mode 0-2 the evaluation's, 3 the adjoint's, 4 the
// fused pass's (wrench and adjoint)
static int64_t part_width_of(int mode) {
    return mode == 4 ? kFmOut + kVjpOut : mode == 3 ? kVjpOut : (mode == 2 ? 24 + 288 : (mode == 1 ? 24 : 12));
}

// Allocates what an evaluation of this mode and precision (or the adjoint,
// mode 3) needs; called before any launch or graph capture.
static int ensure_buffers(atac_batch b, int mode, bool precise) {
    int rc = ATAC_OK;
    if (b->n_slots > 0) {
        const void* before[5] = {b->d_act, b->d_rec64, b->d_rec32, b->d_slow, b->d_slow_diag};
        if (!b->d_act) rc = dalloc(&b->d_act, (size_t)b->rec_cap);
        if (!rc && precise && !b->d_rec64) rc = dalloc(&b->d_rec64, (size_t)b->rec_cap);
        if (!rc && !precise && !b->d_rec32) rc = dalloc(&b->d_rec32, (size_t)b->rec_cap);
        if (!rc && !b->d_slow) rc = dalloc(&b->d_slow, (size_t)b->slow_cap);
        if (!rc && !precise && !b->d_slow_diag) rc = dalloc(&b->d_slow_diag, (size_t)b->slow_cap * 4);
        const void* after[5] = {b->d_act, b->d_rec64, b->d_rec32, b->d_slow, b->d_slow_diag};
        for (int k = 0; k < 5; ++k) b->buf_gen += before[k] != after[k];
    }
    // wide enough for the adjoint too, so a VJP after a captured evaluation does
    // not reallocate the rows (which would drop the captured graph)
    const int64_t w = std::max<int64_t>(part_width_of(mode), part_width_of(3));
    if (!rc && b->chunk_cap > 0 && w > b->part_width) {
        // a captured graph holds the old buffer: it is dropped with it
        if (b->graph_exec) { (void)hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        (void)hipDeviceSynchronize();
        dfree(b->d_part);
        rc = dalloc(&b->d_part, (size_t)(b->chunk_cap * w));
        b->part_width = rc ? 0 : w;
        ++b->buf_gen;
    }
    return rc;
}

// Output capacity per wavefront of the detection kernel: the wavefronts share
// a tile's groups, at most ceil(groups / wavefronts) each.
static int wave_span_of(int tile_points, int block) {
    const int nw = block / 64, ng = (tile_points + kGroup - 1) / kGroup;
    return (ng + nw - 1) / nw * kGroup;
}

static int plan(atac_batch b, const atac_tuning* t) {
    ++b->buf_gen;   // the tile plan and its buffers are rebuilt: cached graphs are stale
    atac_tuning tu = t ? *t : atac_tuning{};
    if (tu.tile_points <= 0) tu.tile_points = 4096;
    if (tu.detect_block <= 0) tu.detect_block = 256;
    if (tu.order == 0) tu.order = 1;
    if (tu.points_per_lane <= 0) tu.points_per_lane = 4;
    if (tu.tile_points > 32 * kGroup || tu.tile_points % kGroup != 0)
        return fail(ATAC_ERR_INVALID, "tile_points must be a multiple of " + std::to_string(kGroup) + " and <= " +
                                          std::to_string(32 * kGroup));
    if (tu.detect_block != 64 && tu.detect_block != 128 && tu.detect_block != 256 && tu.detect_block != 512 &&
        tu.detect_block != 1024)
        return fail(ATAC_ERR_INVALID, "detect_block must be 64, 128, 256, 512 or 1024");
    if (tu.points_per_lane != 1 && tu.points_per_lane != 2 && tu.points_per_lane != 4)
        return fail(ATAC_ERR_INVALID, "points_per_lane must be 1, 2 or 4");
    if (tu.blocks_per_cu < 0 || tu.order < 0 || tu.order > 4 || tu.group_cull < -1 || tu.group_cull > 1)
        return fail(ATAC_ERR_INVALID, "bad tuning");
    if (tu.group_cull == 0) tu.group_cull = 1;
    const int B = tu.detect_block;
    const int span = wave_span_of(tu.tile_points, B);
    const size_t lds = detect_lds_per_wave(span, tu.points_per_lane) * (B / 64);
    if (lds + 1024 > b->lds_limit)
        return fail(ATAC_ERR_INVALID, "tile_points " + std::to_string(tu.tile_points) + " with detect_block " +
                                          std::to_string(B) + " needs " + std::to_string(lds) +
                                          " B of LDS; the device allows " + std::to_string(b->lds_limit) +
                                          " B per workgroup");
    int64_t total_points = 0;
    for (const InstDev& d : b->inst_dev) total_points += d.pt_count;
    if (total_points >= ((int64_t)1 << 31)) return fail(ATAC_ERR_INVALID, "a batch holds fewer than 2^31 points");
    release_plan(b);
    b->tune = tu;
    const int T = tu.tile_points;
    b->tiles.clear();
    int64_t slot = 0, chunks = 0;
    for (size_t i = 0; i < b->inst_dev.size(); ++i) {
        InstDev& d = b->inst_dev[i];
        d.tile_begin = (int32_t)b->tiles.size();
        for (int64_t o = 0; o < d.pt_count; o += T) {
            TileDev td;
            td.inst = (int32_t)i;
            td.count = (int32_t)std::min<int64_t>(T, d.pt_count - o);
            td.begin = o;
            td.slot = slot + o;
            td.grid = d.grid;
            td.pad = 0;
            b->tiles.push_back(td);
            chunks += (td.count + kChunk - 1) / kChunk;
        }
        d.tile_count = (int32_t)b->tiles.size() - d.tile_begin;
        slot += d.pt_count;
    }
    b->n_slots = slot;
    // the record pool bounds the active contacts of one evaluation: a record per
    // point up to 4M points, a quarter of the points beyond (max_active sets it:
    // > 0 that many, < 0 one per point); the slow list holds 1/16 of it
    {
        const int64_t dflt = std::min<int64_t>(slot, std::max<int64_t>((int64_t)1 << 22, slot / 4));
        b->rec_cap = tu.max_active > 0 ? std::min<int64_t>(slot, tu.max_active) : (tu.max_active < 0 ? slot : dflt);
        b->slow_cap = std::min<int64_t>(b->rec_cap, std::max<int64_t>(4096, b->rec_cap / 16));
    }
    // work items: at most one partial item per tile beyond the full ones of the pool
    b->chunk_cap = std::min<int64_t>(chunks, b->rec_cap / kChunk + b->tiles.size());
    b->n_tiles = (int32_t)b->tiles.size();
    // tile queues: 1 = one queue in instance order; 2 = one queue grouped by grid;
    // 3 = grouped by grid and cut into one queue per XCD with balanced point counts;
    // 4 = one queue per XCD by the tile's place in its instance's point order: a
    // point set is Morton-sorted, so the tiles at the same fraction of the instances
    // of one grid cover about the same part of that grid, and that part's
    // classification lines and cells stay in one XCD's L2; the eighths are rotated
    // by the grid index so that every XCD takes a share of every grid
    b->nq = (tu.order == 3 || tu.order == 4) ? kXcds : 1;
    std::vector<int32_t> order(b->n_tiles);
    std::iota(order.begin(), order.end(), 0);
    std::vector<int32_t> tq(b->nq + 1, b->n_tiles);
    if (tu.order == 2 || tu.order == 3) {
        auto grid_of_tile = [&](int32_t x) { return b->inst_dev[b->tiles[x].inst].grid; };
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return grid_of_tile(x) < grid_of_tile(y); });
    }
    if (tu.order == 4) {
        std::vector<int32_t> xcd(b->n_tiles);
        for (int32_t x = 0; x < b->n_tiles; ++x) {
            const TileDev& td = b->tiles[x];
            const InstDev& d = b->inst_dev[td.inst];
            const double mid = ((double)td.begin + 0.5 * td.count) / (double)d.pt_count;
            const int eighth = std::min(kXcds - 1, (int)(mid * kXcds));
            xcd[x] = (eighth + d.grid) % kXcds;
        }
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return xcd[x] < xcd[y]; });
        tq[0] = 0;
        for (int q = 1; q < kXcds; ++q)
            tq[q] = (int32_t)(std::lower_bound(order.begin(), order.end(), q,
                                               [&](int32_t x, int q_) { return xcd[x] < q_; }) - order.begin());
    } else {
        tq[0] = 0;
        double total = 0.0, acc = 0.0;
        for (int32_t it : order) total += (double)b->tiles[it].count;
        int q = 1;
        for (size_t k = 0; k < order.size() && q < b->nq; ++k) {
            acc += (double)b->tiles[order[k]].count;
            while (q < b->nq && acc >= total * q / b->nq) tq[q++] = (int32_t)(k + 1);
        }
    }
    int rc = dalloc(&b->d_tiles, b->n_tiles);
    if (!rc) rc = dalloc(&b->d_order, b->n_tiles);
    if (!rc) rc = dalloc(&b->d_tq_begin, b->nq + 1);
    if (!rc) rc = dalloc(&b->d_counters, CNT_WORDS);
    if (!rc) rc = dalloc(&b->d_tile_active, b->n_tiles);
    if (!rc) rc = dalloc(&b->d_tile_live, b->n_tiles);
    if (!rc) rc = dalloc(&b->d_tile_chunk0, b->n_tiles);
    if (!rc) rc = dalloc(&b->d_chunks, (size_t)b->chunk_cap);
    if (rc) return rc;
    ATAC_HIP(hipMemset(b->d_counters, 0, sizeof(int32_t) * CNT_WORDS));   // no stale work-item count before the first evaluation
    ATAC_HIP(hipMemcpy(b->d_tq_begin, tq.data(), sizeof(int32_t) * tq.size(), hipMemcpyHostToDevice));
    if (b->n_tiles > 0) {
        ATAC_HIP(hipMemcpy(b->d_tiles, b->tiles.data(), sizeof(TileDev) * b->n_tiles, hipMemcpyHostToDevice));
        ATAC_HIP(hipMemcpy(b->d_order, order.data(), sizeof(int32_t) * b->n_tiles, hipMemcpyHostToDevice));
    }
    // inst_dev carries the tile ranges: re-upload
    if (!b->inst_dev.empty())
        ATAC_HIP(hipMemcpy(b->d_inst, b->inst_dev.data(), sizeof(InstDev) * b->inst_dev.size(), hipMemcpyHostToDevice));
    // persistent launch shapes (the FP32 and FP64 detection kernels share LDS and block size)
    b->wave_span = span;
    b->lds_detect = lds;
    int nb = 0;
    ATAC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, detect_kernel_ptr(B, tu.points_per_lane), B, lds));
    if (tu.blocks_per_cu > 0) nb = std::min(nb, tu.blocks_per_cu);
    nb = std::max(nb, 1);
    // workgroup g serves queue g % nq. Default: one workgroup per tile (the
    // dispatcher balances the uneven tiles); with blocks_per_cu > 0, that many
    // persistent workgroups per CU walk their queue's tiles
    int max_q = 0;
    for (int q = 0; q < b->nq; ++q) max_q = std::max(max_q, tq[q + 1] - tq[q]);
    if (tu.blocks_per_cu > 0)
        b->grid_detect = std::max(b->nq, std::min(b->n_tiles, nb * b->n_cu));
    else
        b->grid_detect = std::max(b->nq, b->nq * max_q);
    const int64_t items = std::max<int64_t>(1, (b->chunk_cap + 3) / 4);
    for (int pr = 0; pr < 2; ++pr) {
        for (int m = 0; m < 3; ++m) {
            int nc = 0;
            ATAC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nc, contact_kernel_ptr(m, pr == 1), 256, 0));
            // the Km/Dm mode: a work item per workgroup, as many workgroups as are
            // resident. The others: a wavefront per work item (of the most the
            // evaluation can push; those past its count return at once), so that the
            // dispatcher hands the uneven items out as wavefronts finish (8% faster
            // than a resident grid walking them in a fixed order, measured; a shared
            // atomic counter instead was 2.6x slower, its one address serialising
            // the chip's L2s)
            b->grid_contact[pr][m] = (int)std::max<int64_t>(
                1, m == 2 ? std::min<int64_t>(std::max<int64_t>(1, b->chunk_cap), (int64_t)std::max(nc, 1) * b->n_cu) : items);
        }
        b->grid_vjp[pr] = (int)std::max<int64_t>(1, items);   // a wavefront per work item, as above
    }
    b->grid_slow = std::max(1, 8 * b->n_cu);
    return ATAC_OK;
}

int atac_batch_create(const atac_grid* grids, int32_t n_grids, const atac_points* points, int32_t n_point_sets,
                      const atac_instance* inst, int32_t n_inst, const atac_tuning* tuning, atac_batch* out) {
    if (!out || n_grids < 0 || n_point_sets < 0 || n_inst < 0) return fail(ATAC_ERR_INVALID, "atac_batch_create: bad arguments");
    if ((n_grids && !grids) || (n_point_sets && !points) || (n_inst && !inst)) return fail(ATAC_ERR_INVALID, "atac_batch_create: null array");
    for (int i = 0; i < n_grids; ++i) if (!grids[i]) return fail(ATAC_ERR_INVALID, "atac_batch_create: null grid");
    for (int i = 0; i < n_point_sets; ++i) if (!points[i]) return fail(ATAC_ERR_INVALID, "atac_batch_create: null point set");
    auto* b = new atac_batch_s();
    b->grids.assign(grids, grids + n_grids);
    b->pts.assign(points, points + n_point_sets);
    int rc = check_instances(b, inst, n_inst);
    if (rc) { delete b; return rc; }
    hipError_t e = hipGetDevice(&b->device);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, b->device);
    if (e != hipSuccess) { delete b; return fail(ATAC_ERR_HIP, hipGetErrorString(e)); }
    b->n_cu = prop.multiProcessorCount;
    b->lds_limit = prop.sharedMemPerBlock;
    b->inst.assign(inst, inst + n_inst);
    b->inst_dev.resize(n_inst);
    for (int i = 0; i < n_inst; ++i)
        compose_for(b, inst[i], b->inst_dev[i]);
    rc = ensure_cls(b);
    if (rc) { atac_batch_destroy(b); return rc; }
    std::vector<GridDev> gd(n_grids);
    for (int i = 0; i < n_grids; ++i) gd[i] = b->grids[i]->dev;
    rc = dalloc(&b->d_grids, n_grids);
    if (!rc) rc = dalloc(&b->d_inst, n_inst);
    if (!rc && n_grids) {
        e = hipMemcpy(b->d_grids, gd.data(), sizeof(GridDev) * n_grids, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
    }
    if (!rc) rc = plan(b, tuning);
    if (!rc) {
        e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
        if (e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
    }
    if (rc) { atac_batch_destroy(b); return rc; }
    *out = b;
    return ATAC_OK;
}

int atac_batch_destroy(atac_batch b) {
    if (!b) return ATAC_OK;
    release_plan(b);
    for (auto& u : b->ev_used) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
    for (auto e : b->ev_pool) (void)hipEventDestroy(e);
    dfree(b->d_inst); dfree(b->d_grids);
    for (auto& ct : b->cls) { dfree(ct.d); dfree(ct.sat); }
    dfree(b->h_out); dfree(b->h_nact); dfree(b->h_gfm); dfree(b->h_grad);
    if (b->pin) (void)hipHostFree(b->pin);
    if (b->pin_inst) (void)hipHostFree(b->pin_inst);
    for (auto& kv : b->host_graphs) {
        if (kv.second.e) (void)hipGraphExecDestroy(kv.second.e);
        if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
    }
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
    return ATAC_OK;
}

int atac_batch_set_instances(atac_batch b, const atac_instance* inst, int32_t n_inst, void* stream) {
    if (!b || (n_inst && !inst)) return fail(ATAC_ERR_INVALID, "atac_batch_set_instances: null argument");
    if (n_inst != (int32_t)b->inst.size()) return fail(ATAC_ERR_INVALID, "atac_batch_set_instances: instance count differs");
    for (int i = 0; i < n_inst; ++i)
        if (inst[i].grid != b->inst[i].grid || inst[i].points != b->inst[i].points || inst[i].segment != b->inst[i].segment)
            return fail(ATAC_ERR_INVALID, "atac_batch_set_instances: instance " + std::to_string(i) + " changes its grid/points/segment");
    int rc = check_instances(b, inst, n_inst);
    if (rc) return rc;
    for (int i = 0; i < n_inst; ++i) {
        const int32_t tb = b->inst_dev[i].tile_begin, tc = b->inst_dev[i].tile_count;
        compose_for(b, inst[i], b->inst_dev[i]);
        b->inst_dev[i].tile_begin = tb;
        b->inst_dev[i].tile_count = tc;
    }
    rc = ensure_cls(b);   // a new col_th needs its classification table
    if (rc) return rc;
    b->inst.assign(inst, inst + n_inst);
    // the records of the last evaluation belong to the old poses: the adjoint
    // must not combine them with the new instance records
    b->rec_valid = false;
    if (n_inst) {
        if (stream) {
            ATAC_HIP(hipMemcpyAsync(b->d_inst, b->inst_dev.data(), sizeof(InstDev) * n_inst, hipMemcpyHostToDevice,
                                    (hipStream_t)stream));
        } else {
            // no stream: a synchronous upload after everything already queued on
            // the batch's own stream (the host-buffer entry points run there)
            ATAC_HIP(hipStreamSynchronize(b->cap_stream));
            ATAC_HIP(hipMemcpy(b->d_inst, b->inst_dev.data(), sizeof(InstDev) * n_inst, hipMemcpyHostToDevice));
        }
    }
    return ATAC_OK;
}

int atac_batch_retune(atac_batch b, const atac_tuning* tuning) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_retune: null batch");
    ATAC_HIP(hipDeviceSynchronize());
    return plan(b, tuning);
}

int atac_batch_info(atac_batch b, int64_t* n_points, int32_t* n_tiles, int32_t* persistent_blocks, atac_tuning* tuning) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_info: null batch");
    if (n_points) *n_points = b->n_slots;
    if (n_tiles) *n_tiles = b->n_tiles;
    if (persistent_blocks) *persistent_blocks = b->grid_detect;
    if (tuning) *tuning = b->tune;
    return ATAC_OK;
}

static int mode_of(uint32_t flags) { return (flags & ATAC_EVAL_DERIV) ? 2 : ((flags & ATAC_EVAL_P6) ? 1 : 0); }

static QueueDev tile_queues(atac_batch b) {
    QueueDev q{};
    q.order = b->d_order;
    q.begin = b->d_tq_begin;
    q.nq = b->nq;
    return q;
}

static int32_t* counters(atac_batch b) { return b->d_counters; }

// overflow_out (optional) receives the evaluation's overflow flag from the reduction.
// g_fused != nullptr: the fused pass (ATAC_EVAL_FUSED_ADJOINT), the wrench and the
// adjoint of g_fused into out and gg, no records kept.
static int enqueue_eval(atac_batch b, uint32_t flags, const atac_outputs* out, hipStream_t st,
                        int32_t* overflow_out = nullptr, const double* g_fused = nullptr,
                        const atac_gradients* gg = nullptr) {
    if (flags & ~(ATAC_EVAL_DERIV | ATAC_EVAL_P6 | ATAC_EVAL_CHECKPOINT | ATAC_EVAL_PRECISE | ATAC_EVAL_ALL_ACTIVE |
                  (g_fused ? ATAC_EVAL_FUSED_ADJOINT : 0u)))
        return fail(ATAC_ERR_INVALID, "atac_batch_eval: unknown flag");
    const bool fused = g_fused != nullptr;
    if (fused && (flags & (ATAC_EVAL_DERIV | ATAC_EVAL_P6)))
        return fail(ATAC_ERR_INVALID, "ATAC_EVAL_FUSED_ADJOINT computes the wrench only (no ATAC_EVAL_DERIV or ATAC_EVAL_P6)");
    const int mode = fused ? 4 : mode_of(flags);
    const bool precise = (flags & ATAC_EVAL_PRECISE) != 0;
    if (out && out->dfm_dp6 && mode < 1) return fail(ATAC_ERR_INVALID, "dfm_dp6 requested without ATAC_EVAL_P6 or ATAC_EVAL_DERIV");
    if (out && (out->Km || out->Dm) && mode < 2) return fail(ATAC_ERR_INVALID, "Km/Dm requested without ATAC_EVAL_DERIV");
    if ((flags & ATAC_EVAL_ALL_ACTIVE) && b->rec_cap < b->n_slots)
        return fail(ATAC_ERR_INVALID, "ATAC_EVAL_ALL_ACTIVE needs a record per point: plan the batch with max_active < 0");
    int rc = ensure_buffers(b, mode, precise);
    if (rc) return rc;
    if (b->inst.empty()) return ATAC_OK;
    void* rec = precise ? (void*)b->d_rec64 : (void*)b->d_rec32;
    // groups whose boxes reach only cells where no point can be active are dropped
    // first; that pass also clears the counters (all points are contacts with
    // ATAC_EVAL_ALL_ACTIVE: nothing is dropped)
    const bool cull = b->tune.group_cull > 0 && !(flags & ATAC_EVAL_ALL_ACTIVE) && b->n_tiles > 0;
    if (!cull) ATAC_HIP(launch_reset(b->d_counters, CNT_WORDS, st));
    if (b->n_tiles > 0) {
        DetectArgs da{};
        da.q = tile_queues(b);
        da.tiles = b->d_tiles; da.wave_span = b->wave_span; da.inst = b->d_inst; da.grids = b->d_grids;
        da.all_active = (flags & ATAC_EVAL_ALL_ACTIVE) ? 1 : 0;
        da.act = b->d_act; da.rec_cap = (int32_t)b->rec_cap; da.slow_cap = (int32_t)b->slow_cap;
        da.chunk_cap = (int32_t)b->chunk_cap;
        da.tile_active = b->d_tile_active; da.tile_chunk0 = b->d_tile_chunk0; da.chunks = b->d_chunks;
        da.slow = b->d_slow; da.counters = counters(b);
        da.tile_live = cull ? b->d_tile_live : nullptr;
        b->last_culled = cull;
        hipEvent_t t0 = stage_begin(b, st);
        if (cull) {
            CullArgs cu{};
            cu.tiles = b->d_tiles; cu.inst = b->d_inst; cu.grids = b->d_grids; cu.tile_live = b->d_tile_live;
            cu.counters = b->d_counters; cu.n_tiles = b->n_tiles;
            const int ng = (b->tune.tile_points + kGroup - 1) / kGroup;
            cu.slots = 4;
            while (cu.slots < ng) cu.slots *= 2;
            ATAC_HIP(launch_cull(cu, st));
        }
        ATAC_HIP(launch_detect(da, b->tune.detect_block, b->tune.points_per_lane, b->grid_detect, b->lds_detect, st));
        stage_end(b, ST_DETECT, t0, st);
        SlowArgs sa{};
        sa.inst = b->d_inst; sa.grids = b->d_grids; sa.act = b->d_act; sa.rec = rec; sa.slow_diag = b->d_slow_diag;
        sa.slow = b->d_slow; sa.counters = counters(b); sa.slow_cap = (int32_t)b->slow_cap;
        hipEvent_t t2 = stage_begin(b, st);
        ATAC_HIP(launch_slow(sa, precise, b->grid_slow, st));
        stage_end(b, ST_SLOW, t2, st);
        ContactArgs ca{};
        ca.chunks = b->d_chunks; ca.inst = b->d_inst; ca.grids = b->d_grids; ca.act = b->d_act; ca.rec = rec;
        ca.slow_diag = b->d_slow_diag;
        ca.counters = counters(b); ca.chunk_cap = (int32_t)b->chunk_cap; ca.partial = b->d_part;
        ca.g_fm = g_fused;
        hipEvent_t t3 = stage_begin(b, st);
        if (fused) ATAC_HIP(launch_fused(ca, precise, b->grid_vjp[precise ? 1 : 0], st));
        else ATAC_HIP(launch_contact(ca, mode, precise, b->grid_contact[precise ? 1 : 0][mode], st));
        stage_end(b, ST_CONTACT, t3, st);
    }
    ReduceArgs ra{};
    ra.inst = b->d_inst; ra.n_inst = (int32_t)b->inst.size(); ra.nacc = (int32_t)part_width_of(mode);
    ra.tile_active = b->d_tile_active; ra.tile_chunk0 = b->d_tile_chunk0; ra.partial = b->d_part;
    ra.nseg = 4;
    ra.dst[0] = out ? out->fm : nullptr;      ra.begin[0] = 0;   ra.width[0] = 12;
    ra.dst[1] = out ? out->dfm_dp6 : nullptr; ra.begin[1] = 12;  ra.width[1] = 12;
    ra.dst[2] = out ? out->Km : nullptr;      ra.begin[2] = 24;  ra.width[2] = 144;
    ra.dst[3] = out ? out->Dm : nullptr;      ra.begin[3] = 168; ra.width[3] = 144;
    if (fused) {   // fm 12 | q1 6 | q2 6 | phi1 6 | phi2 6 | kn | kt
        double* gd[6] = {gg ? gg->g_q1 : nullptr, gg ? gg->g_q2 : nullptr, gg ? gg->g_phi1 : nullptr,
                         gg ? gg->g_phi2 : nullptr, gg ? gg->g_kn : nullptr, gg ? gg->g_kt : nullptr};
        const int gw[6] = {6, 6, 6, 6, 1, 1};
        ra.nseg = 7;
        ra.dst[1] = nullptr; ra.dst[2] = nullptr; ra.dst[3] = nullptr;
        for (int k = 0, off = 12; k < 6; off += gw[k], ++k) {
            ra.dst[1 + k] = gd[k];
            ra.begin[1 + k] = off;
            ra.width[1 + k] = gw[k];
        }
    }
    ra.n_active = out ? out->n_active : nullptr;
    ra.counters = counters(b);
    ra.overflow_out = overflow_out;
    ra.host_dst = overflow_out ? 1 : 0;   // the small-batch graph's outputs are its pinned staging
    hipEvent_t t3 = stage_begin(b, st);
    ATAC_HIP(launch_reduce(ra, st));
    stage_end(b, ST_REDUCE, t3, st);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ATAC_HIP(hipStreamIsCapturing(st, &cs));
    if (fused) {
        // the fast-path records were not written: nothing for atac_batch_vjp
        b->rec_valid = false;
        if (cs != hipStreamCaptureStatusNone) b->rec_captured = false;
    } else if (cs == hipStreamCaptureStatusNone) {
        b->rec_valid = true;
        b->rec_precise = precise;
    } else {
        // recorded only: the records exist once the graph has run (atac_batch_graph_launch)
        b->rec_valid = false;
        b->rec_captured = true;
        b->cap_precise = precise;
    }
    return ATAC_OK;
}

// The adjoint of the last evaluation that has run (or, while capturing, of the
// evaluation recorded just before it in the same graph).
static int enqueue_vjp(atac_batch b, const double* g_fm, const atac_gradients* g, hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ATAC_HIP(hipStreamIsCapturing(st, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (capturing ? !b->rec_captured : !b->rec_valid)
        return fail(ATAC_ERR_STATE, "atac_batch_vjp: no evaluation of this batch to differentiate");
    if (!g_fm) return fail(ATAC_ERR_INVALID, "atac_batch_vjp: g_fm is null");
    const bool precise = capturing ? b->cap_precise : b->rec_precise;
    int rc = ensure_buffers(b, 3, precise);
    if (rc) return rc;
    if (b->inst.empty()) return ATAC_OK;
    if (b->n_tiles > 0) {
        VjpArgs va{};
        va.chunks = b->d_chunks; va.inst = b->d_inst;
        va.rec = precise ? (const void*)b->d_rec64 : (const void*)b->d_rec32; va.slow_diag = b->d_slow_diag;
        va.counters = counters(b); va.chunk_cap = (int32_t)b->chunk_cap; va.g_fm = g_fm; va.partial = b->d_part;
        hipEvent_t t0 = stage_begin(b, st);
        ATAC_HIP(launch_vjp(va, precise, b->grid_vjp[precise ? 1 : 0], st));
        stage_end(b, ST_VJP, t0, st);
    }
    ReduceArgs ra{};
    ra.inst = b->d_inst; ra.n_inst = (int32_t)b->inst.size(); ra.nacc = kVjpOut;
    ra.tile_active = b->d_tile_active; ra.tile_chunk0 = b->d_tile_chunk0; ra.partial = b->d_part;
    ra.nseg = 6;
    ra.dst[0] = g ? g->g_q1 : nullptr;   ra.begin[0] = 0;  ra.width[0] = 6;
    ra.dst[1] = g ? g->g_q2 : nullptr;   ra.begin[1] = 6;  ra.width[1] = 6;
    ra.dst[2] = g ? g->g_phi1 : nullptr; ra.begin[2] = 12; ra.width[2] = 6;
    ra.dst[3] = g ? g->g_phi2 : nullptr; ra.begin[3] = 18; ra.width[3] = 6;
    ra.dst[4] = g ? g->g_kn : nullptr;   ra.begin[4] = 24; ra.width[4] = 1;
    ra.dst[5] = g ? g->g_kt : nullptr;   ra.begin[5] = 25; ra.width[5] = 1;
    ra.counters = counters(b);
    hipEvent_t t1 = stage_begin(b, st);
    ATAC_HIP(launch_reduce(ra, st));
    stage_end(b, ST_VJP_REDUCE, t1, st);
    return ATAC_OK;
}

int atac_batch_eval(atac_batch b, uint32_t flags, const atac_outputs* out, void* stream) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_eval: null batch");
    return enqueue_eval(b, flags, out, (hipStream_t)stream);
}

int atac_batch_eval_vjp(atac_batch b, uint32_t flags, const atac_outputs* out, const double* g_fm,
                        const atac_gradients* grads, void* stream) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_eval_vjp: null batch");
    if (!g_fm) return fail(ATAC_ERR_INVALID, "atac_batch_eval_vjp: g_fm is null");
    return enqueue_eval(b, flags | ATAC_EVAL_FUSED_ADJOINT, out, (hipStream_t)stream, nullptr, g_fm, grads);
}

int atac_batch_vjp(atac_batch b, const double* g_fm, const atac_gradients* out, void* stream) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_vjp: null batch");
    return enqueue_vjp(b, g_fm, out, (hipStream_t)stream);
}

int atac_batch_graph_capture(atac_batch b, uint32_t flags, const atac_outputs* out, const double* g_fm,
                             const atac_gradients* grads) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_graph_capture: null batch");
    const bool fused = (flags & ATAC_EVAL_FUSED_ADJOINT) != 0;
    if (fused && !g_fm) return fail(ATAC_ERR_INVALID, "atac_batch_graph_capture: ATAC_EVAL_FUSED_ADJOINT needs g_fm");
    // every buffer the captured launches use exists before capture begins
    int rc = ensure_buffers(b, fused ? 4 : mode_of(flags), (flags & ATAC_EVAL_PRECISE) != 0);
    if (!rc && g_fm) rc = ensure_buffers(b, 3, (flags & ATAC_EVAL_PRECISE) != 0);
    if (rc) return rc;
    if (b->graph_exec) { (void)hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    ATAC_HIP(hipStreamSynchronize(b->cap_stream));
    const bool timing = b->timing;
    b->timing = false;
    ATAC_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    b->rec_captured = false;
    b->graph_fused = fused;
    if (fused) {
        rc = enqueue_eval(b, flags, out, b->cap_stream, nullptr, g_fm, grads);
    } else {
        rc = enqueue_eval(b, flags, out, b->cap_stream);
        if (!rc && g_fm) rc = enqueue_vjp(b, g_fm, grads, b->cap_stream);
    }
    b->timing = timing;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(ATAC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    ATAC_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    return ATAC_OK;
}

int atac_batch_graph_launch(atac_batch b, void* stream) {
    if (!b || !b->graph_exec) return fail(ATAC_ERR_STATE, "atac_batch_graph_launch: no captured graph");
    ATAC_HIP(hipGraphLaunch(b->graph_exec, (hipStream_t)stream));
    if (b->graph_fused) {
        b->rec_valid = false;   // its pass overwrote the work lists and kept no records
    } else if (b->rec_captured) {   // the graph's evaluation leaves its records for atac_batch_vjp
        b->rec_valid = true;
        b->rec_precise = b->cap_precise;
    }
    return ATAC_OK;
}

int atac_batch_active_ids(atac_batch b, int32_t inst, int64_t* ids, int64_t capacity, int64_t* count) {
    if (!b || !count) return fail(ATAC_ERR_INVALID, "atac_batch_active_ids: null argument");
    if (inst < 0 || inst >= (int32_t)b->inst.size()) return fail(ATAC_ERR_INVALID, "atac_batch_active_ids: bad instance");
    if (!b->rec_valid) return fail(ATAC_ERR_STATE, "atac_batch_active_ids: the batch has not been evaluated");
    ATAC_HIP(hipDeviceSynchronize());
    int32_t ovf = 0;
    ATAC_HIP(hipMemcpy(&ovf, b->d_counters + CNT_OVERFLOW, sizeof ovf, hipMemcpyDeviceToHost));
    if (ovf) return fail(ATAC_ERR_STATE, "atac_batch_active_ids: the last evaluation overflowed its record pool");
    const InstDev& d = b->inst_dev[inst];
    std::vector<int32_t> ta(d.tile_count), tc(d.tile_count);
    if (d.tile_count) {
        ATAC_HIP(hipMemcpy(ta.data(), b->d_tile_active + d.tile_begin, sizeof(int32_t) * d.tile_count, hipMemcpyDeviceToHost));
        ATAC_HIP(hipMemcpy(tc.data(), b->d_tile_chunk0 + d.tile_begin, sizeof(int32_t) * d.tile_count, hipMemcpyDeviceToHost));
    }
    const atac_points_s& P = *b->pts[b->inst[inst].points];
    const int64_t seg_begin = P.seg[b->inst[inst].segment];
    std::vector<int64_t> out;
    for (int k = 0; k < d.tile_count; ++k) {
        if (!ta[k]) continue;
        // the tile's first work item gives the start of its range in the pool
        ChunkDev ch;
        ATAC_HIP(hipMemcpy(&ch, b->d_chunks + tc[k], sizeof ch, hipMemcpyDeviceToHost));
        std::vector<uint32_t> off(ta[k]);
        ATAC_HIP(hipMemcpy(off.data(), b->d_act + ch.rec0, sizeof(uint32_t) * ta[k], hipMemcpyDeviceToHost));
        for (uint32_t o : off) out.push_back(P.perm[seg_begin + (o & ~kActSlow)]);
    }
    std::sort(out.begin(), out.end());
    *count = (int64_t)out.size();
    if (ids) {
        if (capacity < (int64_t)out.size()) return fail(ATAC_ERR_INVALID, "atac_batch_active_ids: capacity too small");
        std::copy(out.begin(), out.end(), ids);
    }
    return ATAC_OK;
}

int atac_batch_touched_values(atac_batch b, int64_t* count) {
    if (!b || !count) return fail(ATAC_ERR_INVALID, "atac_batch_touched_values: null argument");
    std::vector<int64_t> off(b->grids.size() + 1, 0);
    for (size_t g = 0; g < b->grids.size(); ++g) off[g + 1] = off[g] + (b->grids[g]->count + 31) / 32;
    uint32_t* bits = nullptr;
    int64_t* d_off = nullptr;
    unsigned long long* d_cnt = nullptr;
    int rc = dalloc(&bits, (size_t)std::max<int64_t>(off.back(), 1));
    if (!rc) rc = dalloc(&d_off, off.size());
    if (!rc) rc = dalloc(&d_cnt, 1);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemset(bits, 0, sizeof(uint32_t) * std::max<int64_t>(off.back(), 1));
    if (!rc && e == hipSuccess) e = hipMemset(d_cnt, 0, sizeof(unsigned long long));
    if (!rc && e == hipSuccess) e = hipMemcpy(d_off, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) {
        TouchedArgs ta{};
        ta.inst = b->d_inst; ta.grids = b->d_grids; ta.bits = bits; ta.bit_offset = d_off;
        e = launch_touched(ta, (int)b->inst.size(), nullptr);
        if (e == hipSuccess) e = launch_popcount(bits, off.back(), d_cnt, nullptr);
        unsigned long long c = 0;
        if (e == hipSuccess) e = hipMemcpy(&c, d_cnt, sizeof c, hipMemcpyDeviceToHost);
        *count = (int64_t)c;
    }
    dfree(bits); dfree(d_off); dfree(d_cnt);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ATAC_ERR_HIP, hipGetErrorString(e));
    return ATAC_OK;
}

int atac_eval_plain(atac_batch b, uint32_t flags, const atac_outputs* out, void* stream) {
    if (!b || !out || !out->fm || !out->n_active) return fail(ATAC_ERR_INVALID, "atac_eval_plain: fm and n_active are required");
    if (flags & ~(ATAC_EVAL_DERIV | ATAC_EVAL_P6 | ATAC_EVAL_PRECISE)) return fail(ATAC_ERR_INVALID, "atac_eval_plain: unsupported flag");
    const bool deriv = (flags & ATAC_EVAL_DERIV) != 0;
    if (deriv && (!out->Km || !out->Dm)) return fail(ATAC_ERR_INVALID, "atac_eval_plain: Km and Dm are required with ATAC_EVAL_DERIV");
    hipStream_t st = (hipStream_t)stream;
    const int n = (int)b->inst.size();
    ATAC_HIP(launch_zero(out->fm, (int64_t)n * 12 * 2, st));
    ATAC_HIP(launch_zero(out->n_active, n, st));
    if (out->dfm_dp6) ATAC_HIP(launch_zero(out->dfm_dp6, (int64_t)n * 12 * 2, st));
    if (deriv) {
        ATAC_HIP(launch_zero(out->Km, (int64_t)n * 144 * 2, st));
        ATAC_HIP(launch_zero(out->Dm, (int64_t)n * 144 * 2, st));
    }
    for (int i = 0; i < n; ++i) {
        PlainArgs pa{};
        pa.inst = b->d_inst; pa.i = i; pa.grids = b->d_grids;
        pa.fm = out->fm + 12 * (int64_t)i;
        pa.p6 = out->dfm_dp6 ? out->dfm_dp6 + 12 * (int64_t)i : nullptr;
        pa.Km = deriv ? out->Km + 144 * (int64_t)i : nullptr;
        pa.Dm = deriv ? out->Dm + 144 * (int64_t)i : nullptr;
        pa.n_active = out->n_active + i;
        ATAC_HIP(launch_plain(pa, b->inst_dev[i].pt_count, (flags & ATAC_EVAL_PRECISE) != 0, st));
    }
    return ATAC_OK;
}

static int ensure_staging(atac_batch b);

int atac_eval_plain_vjp(atac_batch b, uint32_t flags, const double* g_fm, const atac_outputs* out,
                        const atac_gradients* grads, void* stream) {
    if (!b || !g_fm || !out || !out->fm || !out->n_active || !grads)
        return fail(ATAC_ERR_INVALID, "atac_eval_plain_vjp: g_fm, fm, n_active and the gradients are required");
    if (flags & ~ATAC_EVAL_PRECISE) return fail(ATAC_ERR_INVALID, "atac_eval_plain_vjp: only ATAC_EVAL_PRECISE is accepted");
    hipStream_t st = (hipStream_t)stream;
    const int n = (int)b->inst.size();
    if (n == 0) return ATAC_OK;
    // the per-instance adjoint rows (kVjpOut each) in library staging, then split
    int rc = ensure_staging(b);
    if (rc) return rc;
    double* G = b->h_grad;
    ATAC_HIP(launch_zero(out->fm, (int64_t)n * 12 * 2, st));
    ATAC_HIP(launch_zero(out->n_active, n, st));
    ATAC_HIP(launch_zero(G, (int64_t)n * kVjpOut * 2, st));
    for (int i = 0; i < n; ++i) {
        PlainArgs pa{};
        pa.inst = b->d_inst; pa.i = i; pa.grids = b->d_grids;
        pa.fm = out->fm + 12 * (int64_t)i;
        pa.n_active = out->n_active + i;
        pa.g_fm = g_fm + 12 * (int64_t)i;
        pa.grad = G + (int64_t)kVjpOut * i;
        ATAC_HIP(launch_plain(pa, b->inst_dev[i].pt_count, (flags & ATAC_EVAL_PRECISE) != 0, st));
    }
    double* dst[6] = {grads->g_q1, grads->g_q2, grads->g_phi1, grads->g_phi2, grads->g_kn, grads->g_kt};
    const int begin[6] = {0, 6, 12, 18, 24, 25}, width[6] = {6, 6, 6, 6, 1, 1};
    for (int s = 0; s < 6; ++s)
        if (dst[s])
            ATAC_HIP(hipMemcpy2DAsync(dst[s], sizeof(double) * width[s], G + begin[s], sizeof(double) * kVjpOut,
                                      sizeof(double) * width[s], n, hipMemcpyDeviceToDevice, st));
    return ATAC_OK;
}

// ---------------------------------------------------------------- per-point API
static PointArgs point_args(atac_grid g, const double E2[16], const double* phi2, double col_th) {
    PointArgs a{};
    a.grid = g->dev;
    InstDev& d = a.in;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) { d.R1[3 * r + c] = (r == c) ? 1.0 : 0.0; d.R2[3 * r + c] = E2[4 * r + c]; }
        d.p1[r] = 0.0;
        d.p2[r] = E2[4 * r + 3];
        d.w1[r] = d.v1[r] = 0.0;
        d.w2[r] = phi2 ? phi2[r] : 0.0;
        d.v2[r] = phi2 ? phi2[3 + r] : 0.0;
    }
    d.col_th = col_th;
    d.contact_scale = 1.0;
    d.kn_chain = 1.0;
    set_f32_copies(d);   // the FP32 collision arithmetic reads the rounded copies
    return a;
}

int atac_sdf_distance(atac_grid g, const double E2[16], double col_th, const double* xw, int64_t n, double* d, void* stream) {
    if (!g || !E2 || n < 0 || (n > 0 && (!xw || !d))) return fail(ATAC_ERR_INVALID, "atac_sdf_distance: bad arguments");
    PointArgs a = point_args(g, E2, nullptr, col_th);
    a.xw = xw; a.n = n; a.out = d;
    ATAC_HIP(launch_distance(a, (hipStream_t)stream));
    return ATAC_OK;
}

int atac_sdf_collision(atac_grid g, const double E2[16], const double phi2[6], double col_th, uint32_t flags,
                       const double* xw, const double* xw_dot, int64_t n, double* values, double* derivs, void* stream) {
    if (!g || !E2 || !phi2 || n < 0 || (n > 0 && (!xw || !xw_dot || !values)))
        return fail(ATAC_ERR_INVALID, "atac_sdf_collision: bad arguments");
    if (flags & ~ATAC_EVAL_PRECISE) return fail(ATAC_ERR_INVALID, "atac_sdf_collision: only ATAC_EVAL_PRECISE is accepted");
    PointArgs a = point_args(g, E2, phi2, col_th);
    a.xw = xw; a.xwdot = xw_dot; a.n = n; a.out = values; a.out2 = derivs;
    ATAC_HIP(launch_collision(a, (flags & ATAC_EVAL_PRECISE) != 0, (hipStream_t)stream));
    return ATAC_OK;
}

int atac_batch_workspace(atac_batch b, int64_t* capacity, int64_t* used, int32_t* overflowed) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_workspace: null batch");
    int32_t c[CNT_WORDS] = {0, 0, 0, 0};
    ATAC_HIP(hipDeviceSynchronize());
    ATAC_HIP(hipMemcpy(c, b->d_counters, sizeof c, hipMemcpyDeviceToHost));
    if (capacity) *capacity = b->rec_cap;
    if (used) *used = c[CNT_RECS];
    if (overflowed) *overflowed = c[CNT_OVERFLOW] != 0 ? 1 : 0;
    return ATAC_OK;
}

int atac_batch_live_points(atac_batch b, int64_t* live_points, int64_t* points) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_live_points: null batch");
    int64_t live = b->n_slots;
    if (b->last_culled && b->n_tiles > 0) {
        std::vector<uint32_t> m(b->n_tiles);
        ATAC_HIP(hipDeviceSynchronize());
        ATAC_HIP(hipMemcpy(m.data(), b->d_tile_live, sizeof(uint32_t) * m.size(), hipMemcpyDeviceToHost));
        live = 0;
        for (int32_t t = 0; t < b->n_tiles; ++t) {
            const int64_t c = b->tiles[t].count;
            for (int g = 0; g * (int64_t)kGroup < c; ++g)
                if ((m[t] >> g) & 1u) live += std::min<int64_t>(kGroup, c - (int64_t)g * kGroup);
        }
    }
    if (live_points) *live_points = live;
    if (points) *points = b->n_slots;
    return ATAC_OK;
}

int atac_batch_timing(atac_batch b, int32_t enable, double* ms_sum, int32_t* launches) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_timing: null batch");
    if (ms_sum || launches) {
        for (int s = 0; s < ST_COUNT; ++s) {
            if (ms_sum) ms_sum[s] = 0.0;
            if (launches) launches[s] = 0;
        }
        for (auto& u : b->ev_used) {
            ATAC_HIP(hipEventSynchronize(u.second.second));
            float ms = 0.0f;
            ATAC_HIP(hipEventElapsedTime(&ms, u.second.first, u.second.second));
            if (ms_sum) ms_sum[u.first] += ms;
            if (launches) launches[u.first] += 1;
        }
    }
    for (auto& u : b->ev_used) { b->ev_pool.push_back(u.second.first); b->ev_pool.push_back(u.second.second); }
    b->ev_used.clear();
    b->timing = enable != 0;
    return ATAC_OK;
}

// ---------------------------------------------------------------- host-buffer variants
static int ensure_staging(atac_batch b) {
    const size_t n = b->inst.size();
    int rc = ATAC_OK;
    if (!b->h_out) rc = dalloc(&b->h_out, n * (12 + 12 + 144 + 144));
    if (!rc && !b->h_nact) rc = dalloc(&b->h_nact, n);
    if (!rc && !b->h_gfm) rc = dalloc(&b->h_gfm, n * 12);
    if (!rc && !b->h_grad) rc = dalloc(&b->h_grad, n * kVjpOut);
    if (!rc && !b->pin) {
        const hipError_t e = hipHostMalloc((void**)&b->pin, n * (312 * sizeof(double) + sizeof(int32_t)) + 16,
                                           hipHostMallocDefault);
        if (e != hipSuccess) {
            b->pin = nullptr;
            rc = fail(ATAC_ERR_NOMEM, std::string("pinned staging: ") + hipGetErrorString(e));
        }
    }
    return rc;
}

int atac_batch_eval_host(atac_batch b, uint32_t flags, double* fm, double* Km, double* Dm, double* dfm_dp6,
                         int32_t* n_active) {
    if (!b) return fail(ATAC_ERR_INVALID, "atac_batch_eval_host: null batch");
    if (b->inst.empty()) return ATAC_OK;
    int rc = ensure_staging(b);
    if (rc) return rc;
    const size_t n = b->inst.size();
    double* base = b->h_out;
    atac_outputs o{};
    o.fm = base;
    o.dfm_dp6 = dfm_dp6 ? base + 12 * n : nullptr;
    o.Km = Km ? base + 24 * n : nullptr;
    o.Dm = Dm ? base + 168 * n : nullptr;
    o.n_active = b->h_nact;
    hipStream_t st = b->cap_stream;
    rc = enqueue_eval(b, flags, &o, st);
    if (rc) return rc;
    // the device staging's layout (fm | p6 | Km | Dm) copied in at most three pieces
    double* P = reinterpret_cast<double*>(b->pin);
    int32_t* Pn = reinterpret_cast<int32_t*>(b->pin + n * 312 * sizeof(double));
    ATAC_HIP(hipMemcpyAsync(P, base, sizeof(double) * (dfm_dp6 ? 24 : 12) * n, hipMemcpyDeviceToHost, st));
    if (Km || Dm)
        ATAC_HIP(hipMemcpyAsync(P + 24 * n, base + 24 * n, sizeof(double) * 288 * n, hipMemcpyDeviceToHost, st));
    ATAC_HIP(hipMemcpyAsync(Pn, o.n_active, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    ATAC_HIP(hipMemcpyAsync(Pn + n, b->d_counters + CNT_OVERFLOW, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    ATAC_HIP(hipStreamSynchronize(st));
    if (fm) std::memcpy(fm, P, sizeof(double) * 12 * n);
    if (dfm_dp6) std::memcpy(dfm_dp6, P + 12 * n, sizeof(double) * 12 * n);
    if (Km) std::memcpy(Km, P + 24 * n, sizeof(double) * 144 * n);
    if (Dm) std::memcpy(Dm, P + 168 * n, sizeof(double) * 144 * n);
    if (n_active) std::memcpy(n_active, Pn, sizeof(int32_t) * n);
    const int32_t ovf = Pn[n];
    if (ovf)
        return fail(ATAC_ERR_NOMEM, "atac_batch_eval_host: more active contacts than the record pool holds (" +
                                        std::to_string(b->rec_cap) + "); plan the batch with a larger max_active");
    return ATAC_OK;
}

int atac_batch_eval_instances_host(atac_batch b, const atac_instance* inst, int32_t n_inst, uint32_t flags, double* fm,
                                   double* Km, double* Dm, double* dfm_dp6, int32_t* n_active) {
    if (!b || (n_inst && !inst)) return fail(ATAC_ERR_INVALID, "atac_batch_eval_instances_host: null argument");
    if (n_inst != (int32_t)b->inst.size())
        return fail(ATAC_ERR_INVALID, "atac_batch_eval_instances_host: instance count differs");
    for (int i = 0; i < n_inst; ++i)
        if (inst[i].grid != b->inst[i].grid || inst[i].points != b->inst[i].points || inst[i].segment != b->inst[i].segment)
            return fail(ATAC_ERR_INVALID, "atac_batch_eval_instances_host: instance " + std::to_string(i) +
                                              " changes its grid/points/segment");
    if (flags & ~(ATAC_EVAL_DERIV | ATAC_EVAL_P6 | ATAC_EVAL_CHECKPOINT | ATAC_EVAL_PRECISE | ATAC_EVAL_ALL_ACTIVE))
        return fail(ATAC_ERR_INVALID, "atac_batch_eval_instances_host: unknown flag");
    const int mode = mode_of(flags);
    if (dfm_dp6 && mode < 1) return fail(ATAC_ERR_INVALID, "dfm_dp6 requested without ATAC_EVAL_P6 or ATAC_EVAL_DERIV");
    if ((Km || Dm) && mode < 2) return fail(ATAC_ERR_INVALID, "Km/Dm requested without ATAC_EVAL_DERIV");
    int rc = check_instances(b, inst, n_inst);
    if (rc) return rc;
    for (int i = 0; i < n_inst; ++i) {
        const int32_t tb = b->inst_dev[i].tile_begin, tc = b->inst_dev[i].tile_count;
        compose_for(b, inst[i], b->inst_dev[i]);
        b->inst_dev[i].tile_begin = tb;
        b->inst_dev[i].tile_count = tc;
    }
    rc = ensure_cls(b);
    if (rc) return rc;
    b->inst.assign(inst, inst + n_inst);
    b->rec_valid = false;
    if (n_inst == 0) return ATAC_OK;
    rc = ensure_staging(b);
    if (!rc) rc = ensure_buffers(b, mode, (flags & ATAC_EVAL_PRECISE) != 0);
    if (rc) return rc;
    const size_t n = (size_t)n_inst;
    if (!b->pin_inst) {
        const hipError_t e = hipHostMalloc((void**)&b->pin_inst, sizeof(InstDev) * n, hipHostMallocDefault);
        if (e != hipSuccess) {
            b->pin_inst = nullptr;
            return fail(ATAC_ERR_NOMEM, std::string("pinned staging: ") + hipGetErrorString(e));
        }
    }
    double* P = reinterpret_cast<double*>(b->pin);
    int32_t* Pn = reinterpret_cast<int32_t*>(b->pin + n * 312 * sizeof(double));
    const uint32_t key = flags | ((Km || Dm) ? 1u << 8 : 0u) | (dfm_dp6 ? 1u << 9 : 0u);
    atac_batch_s::HostGraph& hg = b->host_graphs[key];
    hipStream_t st = b->cap_stream;
    if (!hg.e || hg.gen != b->buf_gen) {
        // capture: instance upload and evaluation, whose reduction writes the results
        // and the overflow flag straight into the pinned staging (no stage timing inside)
        if (hg.e) (void)hipGraphExecDestroy(hg.e);
        if (hg.g) (void)hipGraphDestroy(hg.g);
        hg = atac_batch_s::HostGraph{};
        ATAC_HIP(hipStreamSynchronize(st));
        const bool timing = b->timing, rc_cap = b->rec_captured, cp = b->cap_precise;
        b->timing = false;
        atac_outputs o{};
        o.fm = P;
        o.dfm_dp6 = dfm_dp6 ? P + 12 * n : nullptr;
        o.Km = (Km || Dm) ? P + 24 * n : nullptr;
        o.Dm = (Km || Dm) ? P + 168 * n : nullptr;
        o.n_active = Pn;
        ATAC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipMemcpyAsync(b->d_inst, b->pin_inst, sizeof(InstDev) * n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) rc = enqueue_eval(b, flags, &o, st, Pn + n);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(st, &g);
        b->timing = timing;
        b->rec_captured = rc_cap;   // this graph is the library's own: a user capture's state stays as it was
        b->cap_precise = cp;
        if (rc || e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            return fail(ATAC_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        hg.g = g;
        ATAC_HIP(hipGraphInstantiate(&hg.e, hg.g, nullptr, nullptr, 0));
        hg.gen = b->buf_gen;
    }
    std::memcpy(b->pin_inst, b->inst_dev.data(), sizeof(InstDev) * n);   // the previous call has finished with it
    ATAC_HIP(hipGraphLaunch(hg.e, st));
    ATAC_HIP(hipStreamSynchronize(st));
    b->rec_valid = true;   // the replay left its records for atac_batch_vjp
    b->rec_precise = (flags & ATAC_EVAL_PRECISE) != 0;
    if (fm) std::memcpy(fm, P, sizeof(double) * 12 * n);
    if (dfm_dp6) std::memcpy(dfm_dp6, P + 12 * n, sizeof(double) * 12 * n);
    if (Km) std::memcpy(Km, P + 24 * n, sizeof(double) * 144 * n);
    if (Dm) std::memcpy(Dm, P + 168 * n, sizeof(double) * 144 * n);
    if (n_active) std::memcpy(n_active, Pn, sizeof(int32_t) * n);
    if (Pn[n])
        return fail(ATAC_ERR_NOMEM, "atac_batch_eval_instances_host: more active contacts than the record pool holds (" +
                                        std::to_string(b->rec_cap) + "); plan the batch with a larger max_active");
    return ATAC_OK;
}

int atac_batch_vjp_host(atac_batch b, const double* g_fm, double* g_q1, double* g_q2, double* g_phi1, double* g_phi2,
                        double* g_kn, double* g_kt) {
    if (!b || !g_fm) return fail(ATAC_ERR_INVALID, "atac_batch_vjp_host: null argument");
    if (b->inst.empty()) return ATAC_OK;
    int rc = ensure_staging(b);
    if (rc) return rc;
    const size_t n = b->inst.size();
    hipStream_t st = b->cap_stream;
    ATAC_HIP(hipMemcpyAsync(b->h_gfm, g_fm, sizeof(double) * 12 * n, hipMemcpyHostToDevice, st));
    double* G = b->h_grad;
    atac_gradients g{G, G + 6 * n, G + 12 * n, G + 18 * n, G + 24 * n, G + 25 * n};
    rc = enqueue_vjp(b, b->h_gfm, &g, st);
    if (rc) return rc;
    // the gradients are contiguous in the staging (q1 | q2 | phi1 | phi2 | kn | kt): one copy
    double* P = reinterpret_cast<double*>(b->pin);
    ATAC_HIP(hipMemcpyAsync(P, G, sizeof(double) * 26 * n, hipMemcpyDeviceToHost, st));
    ATAC_HIP(hipStreamSynchronize(st));
    double* outs[6] = {g_q1, g_q2, g_phi1, g_phi2, g_kn, g_kt};
    const size_t begin[6] = {0, 6, 12, 18, 24, 25}, w[6] = {6, 6, 6, 6, 1, 1};
    for (int k = 0; k < 6; ++k)
        if (outs[k]) std::memcpy(outs[k], P + begin[k] * n, sizeof(double) * w[k] * n);
    return ATAC_OK;
}

int atac_sdf_distance_host(atac_grid g, const double E2[16], double col_th, const double* xw, int64_t n, double* d) {
    if (!g || !E2 || n < 0 || (n > 0 && (!xw || !d))) return fail(ATAC_ERR_INVALID, "atac_sdf_distance_host: bad arguments");
    if (n == 0) return ATAC_OK;
    double *dx = nullptr, *dd = nullptr;
    int rc = dalloc(&dx, (size_t)n * 3);
    if (!rc) rc = dalloc(&dd, (size_t)n);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpy(dx, xw, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) rc = atac_sdf_distance(g, E2, col_th, dx, n, dd, nullptr);
    if (!rc && e == hipSuccess) e = hipMemcpy(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost);
    dfree(dx); dfree(dd);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ATAC_ERR_HIP, hipGetErrorString(e));
    return ATAC_OK;
}

int atac_sdf_collision_host(atac_grid g, const double E2[16], const double phi2[6], double col_th, uint32_t flags,
                            const double* xw, const double* xw_dot, int64_t n, double* values, double* derivs) {
    if (!g || n < 0 || (n > 0 && (!xw || !xw_dot || !values))) return fail(ATAC_ERR_INVALID, "atac_sdf_collision_host: bad arguments");
    if (n == 0) return ATAC_OK;
    double *dx = nullptr, *dv = nullptr, *dval = nullptr, *dder = nullptr;
    int rc = dalloc(&dx, (size_t)n * 3);
    if (!rc) rc = dalloc(&dv, (size_t)n * 3);
    if (!rc) rc = dalloc(&dval, (size_t)n * 11);
    if (!rc && derivs) rc = dalloc(&dder, (size_t)n * 135);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpy(dx, xw, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemcpy(dv, xw_dot, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) rc = atac_sdf_collision(g, E2, phi2, col_th, flags, dx, dv, n, dval, dder, nullptr);
    if (!rc && e == hipSuccess) e = hipMemcpy(values, dval, sizeof(double) * 11 * n, hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess && derivs) e = hipMemcpy(derivs, dder, sizeof(double) * 135 * n, hipMemcpyDeviceToHost);
    dfree(dx); dfree(dv); dfree(dval); dfree(dder);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ATAC_ERR_HIP, hipGetErrorString(e));
    return ATAC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ mesh -> SDF grid
int atac_sdf_from_mesh(const float* x, int64_t nv, const uint32_t* tri, int64_t nf, const float origin[3], float dx,
                       float dy, float dz, int32_t ni, int32_t nj, int32_t nk, int32_t exact_band, float* phi) {
    if (!origin || !phi || nv < 0 || nf < 0 || (nv > 0 && !x) || (nf > 0 && !tri))
        return fail(ATAC_ERR_INVALID, "atac_sdf_from_mesh: null argument");
    if (ni < 1 || nj < 1 || nk < 1 || exact_band < 0 || !(dx > 0.0f) || !(dy > 0.0f) || !(dz > 0.0f))
        return fail(ATAC_ERR_INVALID, "atac_sdf_from_mesh: bad grid");
    const int64_t n = (int64_t)ni * nj * nk;
    if (n >= ((int64_t)1 << 31) || nf >= ((int64_t)1 << 32))
        return fail(ATAC_ERR_INVALID, "atac_sdf_from_mesh: grids hold fewer than 2^31 nodes and meshes fewer than 2^32 triangles");
    for (int64_t e = 0; e < 3 * nf; ++e)
        if ((int64_t)tri[e] >= nv) return fail(ATAC_ERR_INVALID, "atac_sdf_from_mesh: triangle index out of range");
    LsArgs a{};
    a.nf = nf;
    a.o[0] = origin[0]; a.o[1] = origin[1]; a.o[2] = origin[2];
    a.dx = dx; a.dy = dy; a.dz = dz;
    a.ni = ni; a.nj = nj; a.nk = nk;
    a.band = exact_band;
    a.phi0 = (float)(ni + nj + nk) * (dx + dy + dz);   // upper bound on any distance
    float* d_x = nullptr;
    uint32_t* d_tri = nullptr;
    unsigned long long* d_best = nullptr;
    float* d_phi = nullptr;
    int32_t *d_closest = nullptr, *d_count = nullptr;
    int rc = dalloc(&d_x, (size_t)std::max<int64_t>(3 * nv, 1));
    if (!rc) rc = dalloc(&d_tri, (size_t)std::max<int64_t>(3 * nf, 1));
    if (!rc) rc = dalloc(&d_best, (size_t)n);
    if (!rc) rc = dalloc(&d_phi, (size_t)n);
    if (!rc) rc = dalloc(&d_closest, (size_t)n);
    if (!rc) rc = dalloc(&d_count, (size_t)n);
    hipError_t e = hipSuccess;
    if (!rc && nv > 0) e = hipMemcpy(d_x, x, sizeof(float) * 3 * nv, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess && nf > 0) e = hipMemcpy(d_tri, tri, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) e = hipMemset(d_best, 0xff, sizeof(unsigned long long) * n);
    if (!rc && e == hipSuccess) e = hipMemset(d_count, 0, sizeof(int32_t) * n);
    if (!rc && e == hipSuccess) {
        a.x = d_x; a.tri = d_tri; a.best = d_best; a.phi = d_phi; a.closest = d_closest; a.count = d_count;
        e = ls_build(a, nullptr);
        if (e == hipSuccess) e = hipMemcpy(phi, d_phi, sizeof(float) * n, hipMemcpyDeviceToHost);
    }
    dfree(d_x); dfree(d_tri); dfree(d_best); dfree(d_phi); dfree(d_closest); dfree(d_count);
    if (rc) return rc;
    if (e != hipSuccess) return fail(ATAC_ERR_HIP, hipGetErrorString(e));
    return ATAC_OK;
}

int atac_sdf_mesh_grid(const double* V, int64_t nv, double dx, int32_t min_size, float min_box[3], float max_box[3],
                       double d[3], int32_t n[3]) {
    if (!V || nv < 1 || min_size < 1 || !(dx > 0.0) || !min_box || !max_box || !d || !n)
        return fail(ATAC_ERR_INVALID, "atac_sdf_mesh_grid: bad arguments");
    // FP32 bounds of the FP32 vertices (sdfgen::update_minmax), FP64 bounds for the spacing
    float lo[3], hi[3];
    double blo[3], bhi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = 3.402823466e+38f;
        hi[a] = -3.402823466e+38f;
        blo[a] = bhi[a] = V[a];
    }
    for (int64_t i = 0; i < nv; ++i)
        for (int a = 0; a < 3; ++a) {
            const double v = V[3 * i + a];
            const float f = (float)v;
            if (f < lo[a]) lo[a] = f;
            if (f > hi[a]) hi[a] = f;
            if (v < blo[a]) blo[a] = v;
            if (v > bhi[a]) bhi[a] = v;
        }
    for (int a = 0; a < 3; ++a) {
        const double s = (bhi[a] - blo[a]) / min_size;
        d[a] = s < dx ? s : dx;
    }
    for (int a = 0; a < 3; ++a) {
        const float unit = (float)d[a];
        lo[a] -= 1.0f * unit;   // one cell of padding
        hi[a] += 1.0f * unit;
        const float ext = hi[a] - lo[a] + (float)(0.5 * d[a]);
        const float cells = ext / (float)d[a];
        if (!(cells >= 0.0f && cells < 2147483000.0f)) return fail(ATAC_ERR_INVALID, "atac_sdf_mesh_grid: bad extent");
        n[a] = (int32_t)(uint32_t)cells + 1;
        min_box[a] = lo[a];
        max_box[a] = hi[a];
    }
    return ATAC_OK;
}

// ------------------------------------------------------------------ BVH contact variant
struct atac_mesh_s {
    MeshDev dev;
    float4* d_nodes = nullptr;
    int32_t* d_tri = nullptr;
    int32_t* d_F = nullptr;
    int32_t* d_FE = nullptr;
    double* d_V = nullptr;
    double* d_FN = nullptr;
    double* d_EN = nullptr;
    double* d_VN = nullptr;
    void release() {
        dfree(d_nodes); dfree(d_tri); dfree(d_F); dfree(d_FE);
        dfree(d_V); dfree(d_FN); dfree(d_EN); dfree(d_VN);
    }
};

namespace {
template <typename T>
int upload(T** d, const std::vector<T>& h) {
    int rc = dalloc(d, h.size());
    if (rc) return rc;
    if (!h.empty()) ATAC_HIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return ATAC_OK;
}
}  // namespace

int atac_mesh_create(const double* V, int64_t nv, const int64_t* F, int64_t nf, atac_mesh* out) {
    if (!out) return fail(ATAC_ERR_INVALID, "atac_mesh_create: out is NULL");
    *out = nullptr;
    MeshHost h;
    const std::string why = mesh_build(V, nv, F, nf, h);
    if (!why.empty()) return fail(ATAC_ERR_INVALID, "atac_mesh_create: " + why);
    auto* m = new atac_mesh_s();
    int rc = upload(&m->d_nodes, h.nodes);
    if (!rc) rc = upload(&m->d_tri, h.tri);
    if (!rc) rc = upload(&m->d_F, h.F);
    if (!rc) rc = upload(&m->d_FE, h.FE);
    if (!rc) rc = upload(&m->d_V, h.V);
    if (!rc) rc = upload(&m->d_FN, h.FN);
    if (!rc) rc = upload(&m->d_EN, h.EN);
    if (!rc) rc = upload(&m->d_VN, h.VN);
    if (rc) {
        m->release();
        delete m;
        return rc;
    }
    MeshDev& D = m->dev;
    D.nodes = m->d_nodes; D.tri = m->d_tri; D.V = m->d_V; D.F = m->d_F;
    D.FN = m->d_FN; D.FE = m->d_FE; D.EN = m->d_EN; D.VN = m->d_VN;
    for (int a = 0; a < 3; ++a) { D.bmin[a] = h.bmin[a]; D.bmax[a] = h.bmax[a]; }
    D.n_nodes = (int32_t)(h.nodes.size() / 2);
    D.nf = (int32_t)nf;
    *out = m;
    return ATAC_OK;
}

int atac_mesh_destroy(atac_mesh m) {
    if (!m) return ATAC_OK;
    m->release();
    delete m;
    return ATAC_OK;
}

int atac_mesh_distance_normal(atac_mesh m, const double E2[16], double col_th, const double* xw, int64_t n, double* d,
                              double* normal, int32_t* face, void* stream) {
    if (!m || !E2 || n < 0 || (n > 0 && (!xw || !d)))
        return fail(ATAC_ERR_INVALID, "atac_mesh_distance_normal: bad arguments");
    MeshQueryArgs a{};
    a.M = m->dev;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) a.R2[3 * r + c] = E2[4 * r + c];
        a.p2[r] = E2[4 * r + 3];
    }
    a.col_th = col_th;
    a.xw = xw; a.n = n; a.d = d; a.normal = normal; a.face = face;
    ATAC_HIP(launch_mesh_distance(a, (hipStream_t)stream));
    return ATAC_OK;
}

int atac_mesh_distance_normal_host(atac_mesh m, const double E2[16], double col_th, const double* xw, int64_t n,
                                   double* d, double* normal, int32_t* face) {
    if (!m || !E2 || n < 0 || (n > 0 && (!xw || !d)))
        return fail(ATAC_ERR_INVALID, "atac_mesh_distance_normal_host: bad arguments");
    if (n == 0) return ATAC_OK;
    double *dx = nullptr, *dd = nullptr, *dn = nullptr;
    int32_t* df = nullptr;
    int rc = dalloc(&dx, (size_t)(3 * n));
    if (!rc) rc = dalloc(&dd, (size_t)n);
    if (!rc && normal) rc = dalloc(&dn, (size_t)(3 * n));
    if (!rc && face) rc = dalloc(&df, (size_t)n);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpy(dx, xw, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) rc = atac_mesh_distance_normal(m, E2, col_th, dx, n, dd, dn, df, nullptr);
    if (!rc && e == hipSuccess) e = hipMemcpy(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess && normal) e = hipMemcpy(normal, dn, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess && face) e = hipMemcpy(face, df, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
    if (!rc && e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
    dfree(dx); dfree(dd); dfree(dn); dfree(df);
    return rc;
}

int atac_bvh_force_host(atac_mesh m, const atac_instance* inst, const double* xi, int64_t n, uint32_t flags,
                        double* fm, double* Km, double* Dm, double* dfm_dp6, int64_t* n_active) {
    if (!m || !inst || !fm || n < 0 || (n > 0 && !xi)) return fail(ATAC_ERR_INVALID, "atac_bvh_force_host: bad arguments");
    if ((flags & ATAC_EVAL_DERIV) && (!Km || !Dm)) return fail(ATAC_ERR_INVALID, "atac_bvh_force_host: Km / Dm needed");
    if ((flags & ATAC_EVAL_P6) && !dfm_dp6) return fail(ATAC_ERR_INVALID, "atac_bvh_force_host: dfm_dp6 needed");
    // the detection map of compose_instance over a unit grid at the origin: body 2's frame
    GridDev unit{};
    for (int a = 0; a < 3; ++a) { unit.d[a] = 1.0; unit.inv_d[a] = 1.0; }
    BvhForceArgs a{};
    compose_instance(*inst, unit, nullptr, nullptr, nullptr, n, a.in);
    a.M = m->dev;
    a.n = n;
    a.flags = flags;
    double *dx = nullptr, *dout = nullptr, *dpart = nullptr;
    int64_t *dna = nullptr, *dcnt = nullptr;
    const int64_t nb = bvh_force_blocks(n);
    if (nb >= ((int64_t)1 << 31)) return fail(ATAC_ERR_INVALID, "atac_bvh_force_host: too many points");
    int rc = dalloc(&dx, (size_t)std::max<int64_t>(3 * n, 1));
    if (!rc) rc = dalloc(&dout, (size_t)kBvhOut);
    if (!rc) rc = dalloc(&dna, 1);
    if (!rc) rc = dalloc(&dpart, (size_t)(nb * kBvhOut));
    if (!rc) rc = dalloc(&dcnt, (size_t)nb);
    hipError_t e = hipSuccess;
    std::vector<double> h(kBvhOut);
    int64_t na = 0;
    if (!rc && n > 0) e = hipMemcpy(dx, xi, sizeof(double) * 3 * n, hipMemcpyHostToDevice);
    if (!rc && e == hipSuccess) {
        a.xi = dx; a.out = dout; a.n_active = dna; a.part = dpart; a.part_cnt = dcnt;
        e = launch_bvh_force(a, nullptr);
    }
    if (!rc && e == hipSuccess) e = hipMemcpy(h.data(), dout, sizeof(double) * kBvhOut, hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess) e = hipMemcpy(&na, dna, sizeof(int64_t), hipMemcpyDeviceToHost);
    if (!rc && e != hipSuccess) rc = fail(ATAC_ERR_HIP, hipGetErrorString(e));
    dfree(dx); dfree(dout); dfree(dna); dfree(dpart); dfree(dcnt);
    if (rc) return rc;
    for (int k = 0; k < 12; ++k) fm[k] = h[k];
    if (flags & ATAC_EVAL_P6)
        for (int k = 0; k < 12; ++k) dfm_dp6[k] = h[12 + k];
    if (flags & ATAC_EVAL_DERIV)
        for (int k = 0; k < 144; ++k) { Km[k] = h[24 + k]; Dm[k] = h[168 + k]; }
    if (n_active) *n_active = na;
    return ATAC_OK;
}

int atac_bvh_force_noalloc(atac_mesh m, const atac_instance* inst, const double* xi, int64_t n, uint32_t flags,
                           double* out, int64_t* n_active, void* stream) {
    if (n > kBvhNoAllocMaxPoints)
        return fail(ATAC_ERR_INVALID, "atac_bvh_force_noalloc: " + std::to_string(n) + " points; at most " +
                                          std::to_string(kBvhNoAllocMaxPoints) + " fit its fixed scratch");
    if (!m || !inst || !out || n < 0 || (n > 0 && !xi))
        return fail(ATAC_ERR_INVALID, "atac_bvh_force_noalloc: bad arguments");
    if (flags & ~(ATAC_EVAL_DERIV | ATAC_EVAL_P6 | ATAC_EVAL_PRECISE))
        return fail(ATAC_ERR_INVALID, "atac_bvh_force_noalloc: unsupported flag");
    int dev = 0;
    ATAC_HIP(hipGetDevice(&dev));
    ATAC_HIP(bvh_scratch_setup(dev));   // no-op after the device's first call (or atac_device_init)
    double* part = static_cast<double*>(bvh_scratch_address(dev));
    if (!part) return fail(ATAC_ERR_HIP, "atac_bvh_force_noalloc: no scratch address for the device");
    GridDev unit{};
    for (int a = 0; a < 3; ++a) { unit.d[a] = 1.0; unit.inv_d[a] = 1.0; }
    BvhForceArgs a{};
    compose_instance(*inst, unit, nullptr, nullptr, nullptr, n, a.in);
    a.M = m->dev;
    a.n = n;
    a.flags = flags;
    a.xi = xi;
    a.out = out;
    a.n_active = n_active;
    // every workgroup writes its whole row and count, and the reduction reads
    // only the rows of this launch: nothing needs clearing between calls
    a.part = part;
    a.part_cnt = reinterpret_cast<int64_t*>(part + kBvhScratchBlocks * kBvhOut);
    ATAC_HIP(launch_bvh_force(a, (hipStream_t)stream));
    return ATAC_OK;
}
