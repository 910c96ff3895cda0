// gfx950 kernels of the point-vs-SDF contact path.
//
// Batched path (one persistent launch over all instances):
//   detect_kernel  tiles of Morton-sorted FP32 contact points, one tile per
//                  workgroup iteration, tiles pulled from an atomic queue. It
//                  streams the tile's points (nontemporal loads), maps each into
//                  grid index space with an FP64 affine map, gathers the 8 FP32
//                  corners and compacts the points with d < 0 (per-wave LDS
//                  segments filled by ballot + popcount, deterministic order)
//                  into the tile's slice of a global active list.
//   contact_kernel one force object per work item: the full contact (and
//                  optionally the Km/Dm tangent columns) of its active points
//                  with every lane busy, compensated FP32 sums, DPP wave
//                  reductions, the FP64 result written in a fixed order.
//   vjp_kernel     the adjoint over the checkpointed active contacts only.
// Work items come from per-XCD queues (next_item).
// Plain path (the one-launch-per-instance baseline): plain_kernel, one lane per
// point, inline contact math, FP64 atomics.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "atac_kernels.h"
#include "contact_math.h"

namespace atac {

// ------------------------------------------------------------------ wave primitives
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(dpp_i<CTRL>(hi), dpp_i<CTRL>(lo));
}
__device__ __forceinline__ float read_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l), __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sum over the 64 lanes of a wavefront; every lane must be active. Within each
// row of 16 lanes: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror; the four row totals are then read as scalars. The result is
// wave-uniform and its summation order is fixed.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

template <typename R>
struct Acc;
template <>
struct Acc<float> {
    Comp32 c;
    __device__ void zero() { c.zero(); }
    __device__ void add(float v) { c.add(v); }
    __device__ double value() const { return c.value(); }
};
template <>
struct Acc<double> {
    double s;
    __device__ void zero() { s = 0.0; }
    __device__ void add(double v) { s += v; }
    __device__ double value() const { return s; }
};

__device__ __forceinline__ float stream_load(const float* p) { return __builtin_nontemporal_load(p); }

// ------------------------------------------------------------------ work queues
// Items (tiles or instances) are split into nq contiguous queues; workgroup b
// starts on queue b % nq (the XCD it was dispatched to, with round-robin
// dispatch over the XCDs) and moves to the next queue when its own is drained.
// Called by one thread; qid/tried persist in that thread across calls. Returns
// -1 when every queue is empty, so every workgroup's loop ends.
__device__ __forceinline__ int next_item(const QueueDev& Q, int& qid, int& tried) {
    while (tried < Q.nq) {
        const int s = atomicAdd(Q.counter + qid, 1);
        const int b = Q.begin[qid], len = Q.begin[qid + 1] - b;
        if (s < len) return Q.order[b + s];
        qid = (qid + 1 == Q.nq) ? 0 : qid + 1;
        ++tried;
    }
    return -1;
}

// ------------------------------------------------------------------ batched forward
// Phase 1, streaming: every point of the tile is tested; active offsets are
// compacted into per-wave LDS segments (ballot + popcount, deterministic order)
// and copied to the tile's slice of the global active list.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void detect_kernel(DetectArgs a) {
    constexpr int NW = BLOCK / 64;
    extern __shared__ uint16_t s_queue[];   // tile offsets (tiles hold at most 65536 points)
    __shared__ int s_item;
    __shared__ int s_wcount[NW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int seg = a.wave_seg;
    int qid = blockIdx.x % a.q.nq, tried = 0;
    for (;;) {
        if (tid == 0) s_item = next_item(a.q, qid, tried);
        __syncthreads();
        const int tile = s_item;
        if (tile < 0) break;
        const TileDev t = a.tiles[tile];
        const InstDev& in = a.inst[t.inst];
        const GridDev G = a.grids[in.grid];
        uint16_t* q = s_queue + w * seg;
        int wcount = 0;
        for (int o0 = w * 64; o0 < t.count; o0 += BLOCK) {
            const int o = o0 + lane;
            bool act = false;
            if (o < t.count) {
                const int64_t p = t.begin + o;
                act = a.all_active ||
                      detect_value(G, in, stream_load(in.px + p), stream_load(in.py + p), stream_load(in.pz + p)) < 0.0f;
            }
            const uint64_t m = __ballot(act);
            if (act) q[wcount + __popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)o;
            wcount += __popcll(m);
        }
        if (lane == 0) s_wcount[w] = wcount;
        __syncthreads();
        int pre = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            pre += (k < w) ? s_wcount[k] : 0;
            cnt += s_wcount[k];
        }
        uint32_t* dst = a.act_off + t.slot + pre;
        for (int k = lane; k < wcount; k += 64) dst[k] = q[k];
        if (tid == 0) a.tile_active[tile] = cnt;
        __syncthreads();
    }
}

// The active contacts of one instance, flattened over its tiles: an exclusive
// prefix of the tiles' active counts in LDS maps a flat index to (tile, entry).
// Instances with more than kMaxTilesPerPass tiles are walked in passes.
constexpr int kMaxTilesPerPass = 256;

template <int BLOCK>
struct ActiveList {
    int32_t* pre;     // LDS, kMaxTilesPerPass + 1 entries
    int t0, nt;       // tiles [t0, t0 + nt) of this pass
    int total;
    // fills pre[] for tiles t0 .. t0+nt-1 of the instance (all threads call)
    __device__ void load(const int32_t* tile_active, int tile_first, int pass_t0, int pass_nt) {
        t0 = pass_t0;
        nt = pass_nt;
        const int tid = threadIdx.x;
        for (int k = tid; k < kMaxTilesPerPass + 1; k += BLOCK) pre[k] = 0;
        __syncthreads();
        for (int k = tid; k < nt; k += BLOCK) pre[k + 1] = tile_active[tile_first + t0 + k];
        __syncthreads();
        if (tid == 0)
            for (int k = 0; k < nt; ++k) pre[k + 1] += pre[k];
        __syncthreads();
        total = pre[nt];
    }
    // tile index (relative to t0) holding flat entry q
    __device__ int tile_of(int q) const {
        int lo = 0, hi = nt - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= q) lo = mid; else hi = mid - 1;
        }
        return lo;
    }
};

// Phase 2: full contact (wrench, p6 column, optionally Km/Dm tangent columns)
// of the active points of one instance per work item; final FP64 results are
// written directly (fixed summation order: tile order, then queue order).
template <int BLOCK, int MODE, typename R>
__global__ __launch_bounds__(BLOCK) void contact_kernel(ContactArgs a) {
    constexpr int NW = BLOCK / 64;
    constexpr bool P6 = MODE >= 1;
    constexpr bool DERIV = MODE >= 2;
    constexpr int NACC = P6 ? 24 : 12;
    __shared__ int s_item;
    __shared__ int32_t s_pre[kMaxTilesPerPass + 1];
    __shared__ double s_red[NW][NACC];
    __shared__ double s_km[DERIV ? NW : 1][DERIV ? 288 : 1];
    __shared__ int s_cnt;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int qid = blockIdx.x % a.q.nq, tried = 0;
    ActiveList<BLOCK> L;
    L.pre = s_pre;
    for (;;) {
        if (tid == 0) s_item = next_item(a.q, qid, tried);
        __syncthreads();
        const int i = s_item;
        if (i < 0) break;
        const InstDev& in = a.inst[i];
        const GridDev G = a.grids[in.grid];
        if (DERIV)
            for (int k = lane; k < 288; k += 64) s_km[DERIV ? w : 0][DERIV ? k : 0] = 0.0;
        Acc<R> acc[NACC];
#pragma unroll
        for (int k = 0; k < NACC; ++k) acc[k].zero();
        const InstR<R> I = make_inst<R>(in);
        int n_active = 0;
        for (int pt0 = 0; pt0 < in.tile_count; pt0 += kMaxTilesPerPass) {
            L.load(a.tile_active, in.tile_begin, pt0, min(kMaxTilesPerPass, in.tile_count - pt0));
            n_active += L.total;
            for (int base = 0; base < L.total; base += BLOCK) {
                const int qi = base + tid;
                const bool have = qi < L.total;
                ContactFwd<R> s;
                if (have) {
                    const int tt = L.tile_of(qi);
                    const TileDev t = a.tiles[in.tile_begin + L.t0 + tt];
                    const int64_t slot = t.slot + (qi - L.pre[tt]);
                    const int64_t p = t.begin + a.act_off[slot];
                    contact_sample<R>(G, in, (double)in.px[p], (double)in.py[p], (double)in.pz[p], s);
                    R p6[12];
                    contact_forward<R>(I, s, P6 ? p6 : nullptr);
#pragma unroll
                    for (int k = 0; k < 12; ++k) acc[k].add(s.f[k]);
                    if (P6) {
#pragma unroll
                        for (int k = 0; k < 12; ++k) acc[P6 ? 12 + k : 0].add(p6[k]);
                    }
                    if (a.ck_dg) {
                        const float hi = (float)s.d;
                        a.ck_dg[slot] = CkptDev{hi, (float)((double)s.d - (double)hi), (float)s.g.x, (float)s.g.y,
                                                (float)s.g.z};
                    }
                }
                if (DERIV) {
                    // columns of Km (j < 12) and Dm (j >= 12); every lane takes part in the sums
                    for (int j = 0; j < 24; ++j) {
                        R col[12];
#pragma unroll
                        for (int r = 0; r < 12; ++r) col[r] = (R)0;
                        if (have) contact_jvp<R>(I, s, unit_tangent<R>(j), col);
                        const int obase = (j < 12) ? j : 144 + (j - 12);
#pragma unroll
                        for (int r = 0; r < 12; ++r) {
                            const R v = wave_sum(col[r]);
                            if (lane == 0) s_km[DERIV ? w : 0][DERIV ? obase + 12 * r : 0] += (double)v;
                        }
                    }
                }
            }
            __syncthreads();   // s_pre is reloaded by the next pass
        }
#pragma unroll
        for (int k = 0; k < NACC; ++k) {
            const double v = wave_sum(acc[k].value());
            if (lane == 0) s_red[w][k] = v;
        }
        if (tid == 0) s_cnt = n_active;
        __syncthreads();
        if (tid < NACC) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < NW; ++k) sum += s_red[k][tid];
            double* dst = tid < 12 ? a.fm : a.p6;
            if (dst) dst[(int64_t)i * 12 + (tid % 12)] = sum;
        }
        if (DERIV) {
            for (int k = tid; k < 288; k += BLOCK) {
                double sum = 0.0;
#pragma unroll
                for (int ww = 0; ww < NW; ++ww) sum += s_km[DERIV ? ww : 0][DERIV ? k : 0];
                double* dst = k < 144 ? a.Km : a.Dm;
                if (dst) dst[(int64_t)i * 144 + (k % 144)] = sum;
            }
        }
        if (tid == 0 && a.n_active) a.n_active[i] = s_cnt;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ batched adjoint
// One instance per work item over its checkpointed active contacts.
template <int BLOCK, typename R>
__global__ __launch_bounds__(BLOCK) void vjp_kernel(VjpArgs a) {
    constexpr int NW = BLOCK / 64;
    __shared__ int s_item;
    __shared__ int32_t s_pre[kMaxTilesPerPass + 1];
    __shared__ double s_red[NW][kVjpOut];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int qid = blockIdx.x % a.q.nq, tried = 0;
    ActiveList<BLOCK> L;
    L.pre = s_pre;
    for (;;) {
        if (tid == 0) s_item = next_item(a.q, qid, tried);
        __syncthreads();
        const int i = s_item;
        if (i < 0) break;
        const InstDev& in = a.inst[i];
        const GridDev G = a.grids[in.grid];
        const InstR<R> I = make_inst<R>(in);
        R gb[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) gb[k] = (R)a.g_fm[(int64_t)i * 12 + k];
        R acc[kVjpOut];
#pragma unroll
        for (int k = 0; k < kVjpOut; ++k) acc[k] = (R)0;
        for (int pt0 = 0; pt0 < in.tile_count; pt0 += kMaxTilesPerPass) {
            L.load(a.tile_active, in.tile_begin, pt0, min(kMaxTilesPerPass, in.tile_count - pt0));
            for (int qi = tid; qi < L.total; qi += BLOCK) {
                const int tt = L.tile_of(qi);
                const TileDev t = a.tiles[in.tile_begin + L.t0 + tt];
                const int64_t slot = t.slot + (qi - L.pre[tt]);
                const int64_t p = t.begin + a.act_off[slot];
                const CkptDev ck = a.ck_dg[slot];
                ContactFwd<R> s;
                contact_sample<R>(G, in, (double)in.px[p], (double)in.py[p], (double)in.pz[p], s, &ck);
                contact_forward<R>(I, s, nullptr);
                R v[kVjpOut];
                contact_vjp<R>(I, s, gb, v);
                v[24] *= (R)in.kn_chain;   // d/d kn = kn_chain * d/d kn_eff (ForceGeneralSDFContact.cpp:155-156)
                v[25] *= (R)in.kn_chain;
#pragma unroll
                for (int k = 0; k < kVjpOut; ++k) acc[k] += v[k];
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kVjpOut; ++k) {
            const double v = wave_sum((double)acc[k]);
            if (lane == 0) s_red[w][k] = v;
        }
        __syncthreads();
        if (tid < kVjpOut) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < NW; ++k) sum += s_red[k][tid];
            double* dst = nullptr;
            int col = 0, width = 6;
            if (tid < 6) { dst = a.g_q1; col = tid; }
            else if (tid < 12) { dst = a.g_q2; col = tid - 6; }
            else if (tid < 18) { dst = a.g_phi1; col = tid - 12; }
            else if (tid < 24) { dst = a.g_phi2; col = tid - 18; }
            else if (tid == 24) { dst = a.g_kn; col = 0; width = 1; }
            else { dst = a.g_kt; col = 0; width = 1; }
            if (dst) dst[(int64_t)i * width + col] = sum;
        }
        __syncthreads();
    }
}

__global__ void reset_counters_kernel(int32_t* c, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

__global__ void zero_words_kernel(uint32_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// ------------------------------------------------------------------ plain per-instance path
template <typename R>
__global__ void plain_kernel(PlainArgs a) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const InstDev& in = a.inst[a.i];
    if (o >= in.pt_count) return;
    const GridDev G = a.grids[in.grid];
    const float x = in.px[o], y = in.py[o], z = in.pz[o];
    if (!(detect_value(G, in, x, y, z) < 0.0f)) return;
    const InstR<R> I = make_inst<R>(in);
    ContactFwd<R> s;
    contact_sample<R>(G, in, (double)x, (double)y, (double)z, s);
    R p6[12];
    contact_forward<R>(I, s, a.p6 ? p6 : nullptr);
    for (int k = 0; k < 12; ++k) atomicAdd(a.fm + k, (double)s.f[k]);
    if (a.p6)
        for (int k = 0; k < 12; ++k) atomicAdd(a.p6 + k, (double)p6[k]);
    if (a.Km) {
        for (int j = 0; j < 24; ++j) {
            R col[12];
            contact_jvp<R>(I, s, unit_tangent<R>(j), col);
            double* dst = (j < 12) ? a.Km + j : a.Dm + (j - 12);
            for (int r = 0; r < 12; ++r) atomicAdd(dst + 12 * r, (double)col[r]);
        }
    }
    atomicAdd(a.n_active, 1);
}

// ------------------------------------------------------------------ per-point API kernels
// BodySDFObj::distance for world points (BodySDFObj.cpp:207-213): x in FP64,
// cell and fraction in FP64, corners and interpolation in FP32.
__global__ void distance_kernel(PointArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double xw0 = a.xw[3 * i], xw1 = a.xw[3 * i + 1], xw2 = a.xw[3 * i + 2];
    const double q0 = xw0 - a.in.p2[0], q1 = xw1 - a.in.p2[1], q2 = xw2 - a.in.p2[2];
    double x[3];
    for (int r = 0; r < 3; ++r) x[r] = a.in.R2[r] * q0 + a.in.R2[3 + r] * q1 + a.in.R2[6 + r] * q2;
    const GridDev& G = a.grid;
    const bool inside = x[0] >= G.minb[0] && x[0] <= G.maxb[0] && x[1] >= G.minb[1] && x[1] <= G.maxb[1] &&
                        x[2] >= G.minb[2] && x[2] <= G.maxb[2];
    if (!inside) {
        a.out[i] = (double)FLT_MAX + a.in.col_th;
        return;
    }
    int idx[3];
    float f[3];
    int64_t st[3];
    for (int k = 0; k < 3; ++k) {
        const double c = (x[k] - G.minb[k]) / G.d[k];
        int ii = (int)floor(c);
        const double fr = c - ii;
        ii = ii < 0 ? 0 : (ii > G.n[k] - 1 ? G.n[k] - 1 : ii);
        idx[k] = ii;
        f[k] = (float)fr;
    }
    st[0] = idx[0] < G.n[0] - 1 ? 1 : 0;
    st[1] = idx[1] < G.n[1] - 1 ? G.n[0] : 0;
    st[2] = idx[2] < G.n[2] - 1 ? (int64_t)G.n[0] * G.n[1] : 0;
    const float* v = G.v + grid_index(G, idx[0], idx[1], idx[2]);
    const float p = lerpf(v[0], v[st[0]], f[0]), q = lerpf(v[st[1]], v[st[1] + st[0]], f[0]);
    const float r = lerpf(v[st[2]], v[st[2] + st[0]], f[0]), s = lerpf(v[st[2] + st[1]], v[st[2] + st[1] + st[0]], f[0]);
    a.out[i] = (double)lerpf(lerpf(p, q, f[1]), lerpf(r, s, f[1]), f[2]) + a.in.col_th;
}

// BodySDFObj::collision for world points and velocities (BodySDFObj.cpp:239-377):
// values (11 per point) and, optionally, the 14 derivative blocks (135 per point)
// in the reference's order, each block row-major.
template <typename R>
__global__ void collision_kernel(PointArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const InstR<R> I = make_inst<R>(a.in);
    ContactFwd<R> s;
    contact_sample<R>(a.grid, a.in, a.xw[3 * i], a.xw[3 * i + 1], a.xw[3 * i + 2], s);
    s.xw1dot = mk3<R>((R)a.xwdot[3 * i], (R)a.xwdot[3 * i + 1], (R)a.xwdot[3 * i + 2]);
    collision_forward<R>(I, s);
    double* v = a.out + 11 * i;
    v[0] = s.d; v[1] = s.n.x; v[2] = s.n.y; v[3] = s.n.z; v[4] = s.ddot;
    v[5] = s.tdot.x; v[6] = s.tdot.y; v[7] = s.tdot.z; v[8] = s.xi2.x; v[9] = s.xi2.y; v[10] = s.xi2.z;
    if (!a.out2) return;
    double* o = a.out2 + 135 * i;
    // offsets of the blocks in the 135-value record
    enum { DD_DXW = 0, DD_DQ2 = 3, DN_DXW = 9, DN_DQ2 = 18, DDD_DXW = 36, DDD_DXWDOT = 39, DDD_DQ2 = 42, DDD_DPHI2 = 48,
           DT_DXW = 54, DT_DXWDOT = 63, DT_DQ2 = 72, DT_DPHI2 = 90, DXI_DXW = 108, DXI_DQ2 = 117 };
    const v3<R> Z = zero3<R>();
    for (int c = 0; c < 18; ++c) {
        v3<R> dxw = Z, dxwdot = Z, om2 = Z, v2 = Z, w2p = Z, v2p = Z;
        v3<R>* slot[6] = {&dxw, &dxwdot, &om2, &v2, &w2p, &v2p};
        v3<R>& sl = *slot[c / 3];
        if (c % 3 == 0) sl.x = 1; else if (c % 3 == 1) sl.y = 1; else sl.z = 1;
        CollTangent<R> t;
        collision_jvp<R>(I, s, dxw, dxwdot, om2, v2, w2p, v2p, t);
        const int k = c % 3;
        const R dn[3] = {t.dn.x, t.dn.y, t.dn.z}, dt[3] = {t.dtdot.x, t.dtdot.y, t.dtdot.z};
        const R dx[3] = {t.dxi2.x, t.dxi2.y, t.dxi2.z};
        if (c < 3) {                       // d / d xw
            o[DD_DXW + k] = t.dd;
            o[DDD_DXW + k] = t.dddot;
            for (int r = 0; r < 3; ++r) { o[DN_DXW + 3 * r + k] = dn[r]; o[DT_DXW + 3 * r + k] = dt[r]; o[DXI_DXW + 3 * r + k] = dx[r]; }
        } else if (c < 6) {                // d / d xw_dot
            o[DDD_DXWDOT + k] = t.dddot;
            for (int r = 0; r < 3; ++r) o[DT_DXWDOT + 3 * r + k] = dt[r];
        } else if (c < 12) {               // d / d q2 = [w2; v2]
            const int col = c - 6;
            o[DD_DQ2 + col] = t.dd;
            o[DDD_DQ2 + col] = t.dddot;
            for (int r = 0; r < 3; ++r) { o[DN_DQ2 + 6 * r + col] = dn[r]; o[DT_DQ2 + 6 * r + col] = dt[r]; o[DXI_DQ2 + 6 * r + col] = dx[r]; }
        } else {                           // d / d phi2
            const int col = c - 12;
            o[DDD_DPHI2 + col] = t.dddot;
            for (int r = 0; r < 3; ++r) o[DT_DPHI2 + 6 * r + col] = dt[r];
        }
    }
}

// Marks every grid value read by the 8-corner stencil of an inside point.
__global__ void touched_kernel(TouchedArgs a) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int inst_i = blockIdx.y;
    const InstDev& in = a.inst[inst_i];
    for (int64_t o = gid; o < in.pt_count; o += (int64_t)gridDim.x * blockDim.x) {
        const GridDev G = a.grids[in.grid];
        const double x = in.px[o], y = in.py[o], z = in.pz[o];
        const double c0 = in.A[0] * x + in.A[1] * y + in.A[2] * z + in.b[0];
        const double c1 = in.A[3] * x + in.A[4] * y + in.A[5] * z + in.b[1];
        const double c2 = in.A[6] * x + in.A[7] * y + in.A[8] * z + in.b[2];
        if (!(c0 >= 0.0 && c0 <= in.cmax[0] && c1 >= 0.0 && c1 <= in.cmax[1] && c2 >= 0.0 && c2 <= in.cmax[2])) continue;
        int i = (int)c0, j = (int)c1, k = (int)c2;
        i = i < G.n[0] - 1 ? i : G.n[0] - 1;
        j = j < G.n[1] - 1 ? j : G.n[1] - 1;
        k = k < G.n[2] - 1 ? k : G.n[2] - 1;
        const int ip = i < G.n[0] - 1 ? i + 1 : i, jp = j < G.n[1] - 1 ? j + 1 : j, kp = k < G.n[2] - 1 ? k + 1 : k;
        uint32_t* bits = a.bits + a.bit_offset[in.grid];
        const int ii[2] = {i, ip}, jj[2] = {j, jp}, kk[2] = {k, kp};
        for (int c = 0; c < 8; ++c) {
            const int64_t idx = grid_index(G, ii[c & 1], jj[(c >> 1) & 1], kk[c >> 2]);
            atomicOr(bits + (idx >> 5), 1u << (idx & 31));
        }
    }
}

__global__ void popcount_kernel(const uint32_t* bits, int64_t nwords, unsigned long long* out) {
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * blockDim.x)
        c += __popc(bits[i]);
    atomicAdd(out, c);
}

// ------------------------------------------------------------------ launchers
hipError_t launch_detect(const DetectArgs& a, int block, int grid, size_t lds, hipStream_t st) {
    switch (block) {
        case 128: hipLaunchKernelGGL(detect_kernel<128>, dim3(grid), dim3(128), lds, st, a); break;
        case 256: hipLaunchKernelGGL(detect_kernel<256>, dim3(grid), dim3(256), lds, st, a); break;
        case 512: hipLaunchKernelGGL(detect_kernel<512>, dim3(grid), dim3(512), lds, st, a); break;
        case 1024: hipLaunchKernelGGL(detect_kernel<1024>, dim3(grid), dim3(1024), lds, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

const void* detect_kernel_ptr(int block) {
    switch (block) {
        case 128: return (const void*)&detect_kernel<128>;
        case 256: return (const void*)&detect_kernel<256>;
        case 512: return (const void*)&detect_kernel<512>;
        case 1024: return (const void*)&detect_kernel<1024>;
        default: return nullptr;
    }
}

template <int MODE, typename R>
static const void* contact_ptr_t() { return (const void*)&contact_kernel<256, MODE, R>; }

const void* contact_kernel_ptr(int mode, bool precise) {
    if (precise) return mode == 0 ? contact_ptr_t<0, double>() : mode == 1 ? contact_ptr_t<1, double>() : contact_ptr_t<2, double>();
    return mode == 0 ? contact_ptr_t<0, float>() : mode == 1 ? contact_ptr_t<1, float>() : contact_ptr_t<2, float>();
}

hipError_t launch_contact(const ContactArgs& a, int mode, bool precise, int grid, hipStream_t st) {
    if (precise) {
        if (mode == 0) hipLaunchKernelGGL((contact_kernel<256, 0, double>), dim3(grid), dim3(256), 0, st, a);
        else if (mode == 1) hipLaunchKernelGGL((contact_kernel<256, 1, double>), dim3(grid), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((contact_kernel<256, 2, double>), dim3(grid), dim3(256), 0, st, a);
    } else {
        if (mode == 0) hipLaunchKernelGGL((contact_kernel<256, 0, float>), dim3(grid), dim3(256), 0, st, a);
        else if (mode == 1) hipLaunchKernelGGL((contact_kernel<256, 1, float>), dim3(grid), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((contact_kernel<256, 2, float>), dim3(grid), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_vjp(const VjpArgs& a, bool precise, int grid, hipStream_t st) {
    if (precise) hipLaunchKernelGGL((vjp_kernel<256, double>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((vjp_kernel<256, float>), dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

const void* vjp_kernel_ptr(bool precise) {
    return precise ? (const void*)&vjp_kernel<256, double> : (const void*)&vjp_kernel<256, float>;
}

hipError_t launch_reset(int32_t* c, int n, hipStream_t st) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3((n + 63) / 64), dim3(64), 0, st, c, n);
    return hipGetLastError();
}

hipError_t launch_zero(void* p, int64_t n_words, hipStream_t st) {
    if (n_words <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n_words + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, n_words);
    return hipGetLastError();
}

hipError_t launch_plain(const PlainArgs& a, int64_t n_points, bool precise, hipStream_t st) {
    if (n_points <= 0) return hipSuccess;
    const int64_t blocks = (n_points + 255) / 256;
    if (precise) hipLaunchKernelGGL(plain_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(plain_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_distance(const PointArgs& a, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(distance_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_collision(const PointArgs& a, bool precise, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    if (precise) hipLaunchKernelGGL(collision_kernel<double>, dim3((unsigned)((a.n + 127) / 128)), dim3(128), 0, st, a);
    else hipLaunchKernelGGL(collision_kernel<float>, dim3((unsigned)((a.n + 127) / 128)), dim3(128), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_touched(const TouchedArgs& a, int n_inst, hipStream_t st) {
    if (n_inst <= 0) return hipSuccess;
    hipLaunchKernelGGL(touched_kernel, dim3(64, n_inst), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_popcount(const uint32_t* bits, int64_t nwords, unsigned long long* out, hipStream_t st) {
    hipLaunchKernelGGL(popcount_kernel, dim3(256), dim3(256), 0, st, bits, nwords, out);
    return hipGetLastError();
}

}  // namespace atac
