// gfx950 kernels of the point-vs-SDF contact path.
//
// Batched path:
//   cull_kernel    per group of 256 sorted points: can its box reach a cell where
//                  a point may be active? (the tiles' live-group masks); it also
//                  clears the evaluation's counters.
//   detect_kernel  tiles of Morton-sorted FP32 contact points, taken from one
//                  queue in instance order (or grid-grouped / per-XCD queues).
//                  Streams each tile's live groups (several points per lane, nontemporal
//                  loads), maps every point into the grid's padded cell space
//                  with an FP64 affine map and reads one byte of the grid's cell
//                  classification table (inactive / active / open). Open cells
//                  go through an LDS candidate ring and the exact test; the
//                  active points are appended to the tile's slice of the
//                  active list, with one 64-bit atomic per tile reserving
//                  record slots and work items of the bounded record pool.
//   slow_kernel    the FP64 FD-stencil samples of the few active points whose
//                  stencil leaves their cell.
//   contact_kernel work items of kChunk active points, one wavefront each (one
//                  workgroup each with Km/Dm): builds
//                  the records (sample of phi, grad phi, Hessian) and sums the
//                  wrench (optionally the p6 column and Km/Dm tangent columns)
//                  into the item's partial row.
//   vjp_kernel     the adjoint of the same records.
//   reduce_kernel  per-instance sums of the partial rows in a fixed order.
// Plain path (the one-launch-per-instance baseline): plain_kernel, one lane per
// point, inline contact math, FP64 atomics.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
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

// Per-lane wrench accumulator: FP64 for both arithmetics (an FP32 term is
// converted exactly and added in FP64: two full-rate instructions, and the
// lane's sum carries no FP32 rounding).
template <typename R>
struct LaneAcc {
    double s;
    __device__ void zero() { s = 0.0; }
    __device__ void add(R v) { s += (double)v; }
    __device__ double value() const { return s; }
};

// ------------------------------------------------------------------ tile assignment
// Tiles are split into nq contiguous queues of balanced point counts (nq is 1,
// or 8 for per-XCD queues). Workgroup b belongs to queue b % nq (with nq = 8 the
// XCD it is dispatched to, with round-robin dispatch over the XCDs, so a queue's
// tiles share that XCD's L2) and takes the
// queue's tiles r, r + m, r + 2m, ... where r = b / nq is its rank among the m
// workgroups of the queue. The assignment is static: no shared counter, and
// every workgroup's tile sequence is known from the start.
struct TileWalk {
    int k, end, step;
    __device__ TileWalk(const QueueDev& Q) {
        const int nq = Q.nq, q = (int)blockIdx.x % nq, r = (int)blockIdx.x / nq;
        step = ((int)gridDim.x - q + nq - 1) / nq;
        k = as_const(Q.begin)[q] + r;
        end = as_const(Q.begin)[q + 1];
    }
};

// ------------------------------------------------------------------ records
template <typename R> struct RecOf;
template <> struct RecOf<float> { using type = Rec32; };
template <> struct RecOf<double> { using type = Rec64; };

__device__ __forceinline__ void store_rec(Rec32* p, const Rec32& r) {
    float4* q = reinterpret_cast<float4*>(p);
    q[0] = make_float4(r.a[0], r.a[1], r.a[2], r.a[3]);
    q[1] = make_float4(r.b[0], r.b[1], r.b[2], r.b[3]);
    q[2] = make_float4(r.c[0], r.c[1], r.c[2], r.c[3]);
}
__device__ __forceinline__ Rec32 load_rec(const Rec32* p) {
    const float4* q = reinterpret_cast<const float4*>(p);
    const float4 a = q[0], b = q[1], c = q[2];
    Rec32 r;
    r.a[0] = a.x; r.a[1] = a.y; r.a[2] = a.z; r.a[3] = a.w;
    r.b[0] = b.x; r.b[1] = b.y; r.b[2] = b.z; r.b[3] = b.w;
    r.c[0] = c.x; r.c[1] = c.y; r.c[2] = c.z; r.c[3] = c.w;
    return r;
}
__device__ __forceinline__ void store_rec(Rec64* p, const Rec64& r) { *p = r; }
__device__ __forceinline__ Rec64 load_rec(const Rec64* p) { return *p; }

// A slow-path FP32 record points at its diagonal Hessian entries in slow_diag.
__device__ __forceinline__ void mark_slow_index(Rec32& r, int slow) { r.c[3] = __int_as_float(slow); }
__device__ __forceinline__ void mark_slow_index(Rec64&, int) {}

// ------------------------------------------------------------------ batched detection
// Points o0 + 64 k + lane, k < PPL, of a step into (x, y, z)[k]: each load
// instruction reads 64 consecutive points (a contiguous 256-byte span), so the
// points a wavefront handles together are neighbours in the Morton order and
// their block-bound reads share cache lines. Points at or past `end` read point 0
// (in bounds; their results are masked by the caller).
template <int PPL, typename P>
__device__ __forceinline__ void load_points(P px, P py, P pz, int o0, int lane, int end, float (&x)[PPL],
                                            float (&y)[PPL], float (&z)[PPL]) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int o = o0 + 64 * k + lane;
        const uint32_t oo = (uint32_t)(o < end ? o : 0) * 4u;   // byte offset
        x[k] = nt_load_at<float>(px, oo);
        y[k] = nt_load_at<float>(py, oo);
        z[k] = nt_load_at<float>(pz, oo);
    }
}

// Points of one tile per work item. The tile's live groups are shared among its
// wavefronts by rank (contiguous ranges, so their outputs follow one another in
// point order), and each wavefront streams its groups: a step takes 64 * PPL
// consecutive points of a group, lane l the points 64 k + l (each load and
// table-read instruction covers 64 neighbouring points; the next step's points
// are loaded behind the table reads of the current ones, so a step waits once),
// and only the classification table is read (detect_candidates). The candidates a step finds (settled active or open; a
// few per cent, near the surface) are appended, with their coordinates, to the
// wavefront's LDS ring in ascending order (ballots per point slot k); whenever 64
// are waiting, the next step finishes them one per lane (candidate_locate and
// candidate_finish: cell read and interpolation of the open ones, FD-stencil
// test), overlapped with its own work, instead of every stream step branching
// into that work for a few lanes. The active offsets (bit 15: slow path; tiles
// hold at most 8192 points) go to a per-wave LDS segment, then to the tile's
// slots of the active list, and the tile pushes its work items for the record,
// contact and adjoint passes.
template <int BLOCK, int PPL>
__global__ __launch_bounds__(BLOCK) void detect_kernel(DetectArgs a) {
    constexpr int NW = BLOCK / 64;
    constexpr int CAP = detect_ring_cap(PPL);
    extern __shared__ __align__(16) unsigned char s_lds[];
    __shared__ int s_wcount[NW];
    __shared__ int s_base;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int span = a.wave_span;
    uint16_t* s_queue = reinterpret_cast<uint16_t*>(s_lds);   // [NW][span] active offsets, bit 15 = slow path
    float4* cand = reinterpret_cast<float4*>(s_lds + (size_t)NW * span * 2) + w * CAP;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const bool all_active = a.all_active != 0;
    const auto order = as_const(a.q.order);
    const TileWalk walk(a.q);
    for (int kk = walk.k; kk < walk.end; kk += walk.step) {
        const int tile = order[kk];
        const TileDev t = as_const(a.tiles)[tile];
        const InstDev in = as_const(a.inst)[t.inst];
        const GridDev G = as_const(a.grids)[t.grid];
        const auto px = as_global(in.px) + t.begin;
        const auto py = as_global(in.py) + t.begin;
        const auto pz = as_global(in.pz) + t.begin;
        uint16_t* q = s_queue + w * span;
        // the tile's groups that can hold a candidate (all without the mask); this
        // wave takes those of ranks [w m / NW, (w + 1) m / NW), in ascending order,
        // so the waves' output segments follow one another in point order
        const int ng = (t.count + kGroup - 1) / kGroup;
        const uint32_t all_groups = ng >= 32 ? 0xffffffffu : ((1u << ng) - 1u);
        uint32_t mine = a.tile_live ? (as_const(a.tile_live)[tile] & all_groups) : all_groups;
        {
            const int wu = __builtin_amdgcn_readfirstlane(w);
            const int m = __popc(mine), r0 = (wu * m) / NW, r1 = ((wu + 1) * m) / NW;
            for (int r = 0; r < r0; ++r) mine &= mine - 1u;
            uint32_t later = mine;
            for (int r = r0; r < r1; ++r) later &= later - 1u;
            mine &= ~later;
        }
        // candidates wait in a ring of CAP entries: [rd, rd + occ) (mod CAP)
        int wcount = 0, rd = 0, occ = 0, npend = 0;   // wave-uniform
        auto ring = [](int i) { return i >= CAP ? i - CAP : i; };   // i < 2 CAP
        const float cth = (float)in.col_th;
        const ClsConsts kc = cls_consts_vgpr(G, in);
        // A batch: the n <= 64 oldest candidates, one per lane. start() reads
        // them and issues their cell reads; complete() decides them into the
        // output segment. A step starts one at its top, before its own work, and
        // completes it once its own reads are out, so the cell reads' latency
        // overlaps the step. Every step starts one, empty or not, and every lane
        // issues the two reads (lanes without an open candidate read the first
        // cell), so the reads in flight are the same on every path.
        CandRead pend;
        uint32_t ptag = 0;
        const auto cells = ATAC_GLOBAL(G.cells);
        auto start = [&](int n) {
            uint32_t off = 0;
            if (lane < n) {
                const float4 e = cand[ring(rd + lane)];
                ptag = __float_as_uint(e.w);
                off = candidate_locate(G, in, e.x, e.y, e.z, (ptag >> 15) != 0u, pend);
            }
            pend.lo = load_f4_at(cells, off);
            pend.hi = load_f4_at(cells, off + 16u);
            rd = ring(rd + n);
            occ -= n;
            npend = n;
        };
        auto complete = [&]() {
            bool act = false, slow = false;
            if (lane < npend) act = candidate_finish(pend, cth, slow);
            const uint64_t bal = __ballot(act);
            if (act) q[wcount + __popcll(bal & lt_mask)] = (uint16_t)((ptag & 0x7fffu) | (slow ? 0x8000u : 0u));
            wcount += __popcll(bal);
            npend = 0;
        };
        // one step: points (cx, cy, cz) at o0 (valid below we) are tested while
        // the next step's points (at on, valid below wn) load into (nx, ny, nz);
        // the loop alternates the two buffers, so no register copy waits for the
        // loads in flight
        auto step = [&](int o0, int we, int on, int wn, const float (&cx)[PPL], const float (&cy)[PPL],
                        const float (&cz)[PPL], float (&nx)[PPL], float (&ny)[PPL], float (&nz)[PPL]) {
            start(occ >= 64 ? 64 : 0);
            unsigned valid = 0;
#pragma unroll
            for (int k = 0; k < PPL; ++k)
                if (o0 + 64 * k + lane < we) valid |= 1u << k;
            unsigned op;
            const unsigned m = detect_candidates<PPL>(G, in, kc, cx, cy, cz, valid, all_active, op, [] {}, [&]() {
                __builtin_amdgcn_sched_barrier(0);
                load_points<PPL>(px, py, pz, on, lane, wn, nx, ny, nz);
                __builtin_amdgcn_sched_barrier(0);
                complete();   // its cell reads went out before this step's reads
            });
            // appended in point order: point k of every lane (lane order) before point k + 1
            int pos = rd + occ;   // stays < 2 CAP: occ + (this step's candidates) <= CAP
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
                const bool is_c = ((m >> k) & 1u) != 0u;
                const uint64_t bal = __ballot(is_c);
                if (is_c)
                    cand[ring(pos + __popcll(bal & lt_mask))] = make_float4(
                        cx[k], cy[k], cz[k], __uint_as_float((uint32_t)(o0 + 64 * k + lane) | (((op >> k) & 1u) << 15)));
                pos += __popcll(bal);
            }
            occ = pos - rd;
            // the next step takes at most 64 * PPL and starts by taking 64 if it
            // has them: room for it while occ <= 128 (always with one point per
            // lane); beyond (rare), batches finish here
            if constexpr (PPL > 1) {
                while (occ > 128) {
                    start(64);
                    complete();
                }
            }
        };
        // the wave's steps: kGroup / (64 PPL) per group of `mine`, lowest group
        // first; (g, s) is the current one, advance() moves to the next and says
        // whether there is one (a step without a next loads point 0, in bounds)
        if (mine) {
            constexpr int SP = 64 * PPL, NS = kGroup / SP;
            auto gend = [&](int gg) { return min(t.count, (gg + 1) * kGroup); };
            auto advance = [](uint32_t& lv, int& g, int& s) {
                if (++s < NS) return true;
                s = 0;
                lv &= lv - 1u;
                if (!lv) return false;
                g = __builtin_ctz(lv);
                return true;
            };
            uint32_t lv = mine;
            int g = __builtin_ctz(lv), s = 0;
            float x0[PPL], y0[PPL], z0[PPL], x1[PPL], y1[PPL], z1[PPL];
            load_points<PPL>(px, py, pz, g * kGroup, lane, gend(g), x0, y0, z0);
            for (;;) {
                uint32_t lv1 = lv;
                int g1 = g, s1 = s;
                const bool more = advance(lv1, g1, s1);
                step(g * kGroup + s * SP, gend(g), more ? g1 * kGroup + s1 * SP : 0, more ? gend(g1) : 0, x0, y0, z0,
                     x1, y1, z1);
                if (!more) break;
                lv = lv1; g = g1; s = s1;
                const bool more2 = advance(lv1, g1, s1);
                step(g * kGroup + s * SP, gend(g), more2 ? g1 * kGroup + s1 * SP : 0, more2 ? gend(g1) : 0, x1, y1,
                     z1, x0, y0, z0);
                if (!more2) break;
                lv = lv1; g = g1; s = s1;
            }
        }
        while (occ > 0) {
            start(occ < 64 ? occ : 64);
            complete();
        }
        if (lane == 0) s_wcount[w] = wcount;
        __syncthreads();
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) cnt += s_wcount[k];
        // the tile's records: one range of the pool, reserved by one atomic; a
        // tile that does not fit marks the evaluation as overflowed and writes
        // nothing (its outputs then come out NaN)
        if (tid == 0) {
            int base = -1;
            if (cnt > 0) {
                const int nch = (cnt + kChunk - 1) / kChunk;
                const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long*>(a.counters + CNT_RECS),
                                                         ((unsigned long long)nch << 32) | (unsigned long long)cnt);
                const int r0 = (int)(uint32_t)old, c0 = (int)(uint32_t)(old >> 32);
                const bool ok = r0 <= a.rec_cap - cnt && c0 <= a.chunk_cap - nch;
                if (ok) {
                    base = r0;
                    a.tile_chunk0[tile] = c0;
                } else {
                    atomicOr(a.counters + CNT_OVERFLOW, 1);
                }
                // the items this tile took: its ranges, or (overflowed) empty ones
                for (int k = 0; k < nch && c0 + k < a.chunk_cap; ++k)
                    a.chunks[c0 + k] = ok ? ChunkDev{t.inst, min(kChunk, cnt - k * kChunk), (int64_t)base + (int64_t)k * kChunk}
                                          : ChunkDev{t.inst, 0, 0};
            }
            s_base = base;
            a.tile_active[tile] = base >= 0 ? cnt : 0;
        }
        __syncthreads();
        const int base = s_base;
        if (base >= 0) {
            // the tile's active offsets, in ascending point order, at its pool range;
            // the contact pass builds their records (the slow pass those of the
            // contacts whose FD stencil leaves the cell)
            uint32_t* act = a.act + base;
            const uint32_t tb = (uint32_t)t.begin;
            for (int qi = tid; qi < cnt; qi += BLOCK) {
                int rem = qi, ww = 0;
                while (rem >= s_wcount[ww]) { rem -= s_wcount[ww]; ++ww; }
                const uint32_t e = s_queue[ww * span + rem];
                const bool slow = (e >> 15) != 0u;
                act[qi] = (tb + (e & 0x7fffu)) | (slow ? kActSlow : 0u);
                if (slow) {
                    const int si = atomicAdd(a.counters + CNT_SLOW, 1);
                    if (si < a.slow_cap) a.slow[si] = SlowDev{(uint32_t)(base + qi), t.inst};
                    else atomicOr(a.counters + CNT_OVERFLOW, 1);
                }
            }
        }
        __syncthreads();
    }
}

// Wavefront-granular passes take work items in a fixed stride: wavefront g of
// the launch takes items g, g + G, g + 2G, ... (G = wavefronts in the launch),
// the items being near-equal (at most kChunk records). No shared counter is
// touched, and the item index is wave-uniform (readfirstlane), so every branch
// on it is uniform.
__device__ __forceinline__ int wave_id() {
    return (int)blockIdx.x * (int)(blockDim.x >> 6) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}
__device__ __forceinline__ int wave_count() { return (int)gridDim.x * (int)(blockDim.x >> 6); }

// Records of the slow-path contacts (a few per thousand): the 20 FD-stencil
// evaluations of a point run on 20 lanes at once (three points per wavefront),
// and one lane per point combines them and writes the record.
template <bool PREC>
__global__ __launch_bounds__(64) void slow_kernel(SlowArgs a) {
    using Rec = typename RecOf<typename std::conditional<PREC, double, float>::type>::type;
    constexpr int GROUPS = 64 / kStencil;
    __shared__ double s_v[GROUPS][kStencil];
    const int lane = threadIdx.x, grp = lane / kStencil, st = lane % kStencil;
    const int n = min(a.counters[CNT_SLOW], a.slow_cap);
    for (int base = (int)blockIdx.x * GROUPS; base < n; base += (int)gridDim.x * GROUPS) {
        const int i = base + grp;
        const bool mine = grp < GROUPS && i < n;
        SlowDev sd{0u, 0};
        Rec r;
        if (mine) {
            sd = a.slow[i];
            const InstDev& in = a.inst[sd.inst];
            const uint32_t off = a.act[sd.slot] & ~kActSlow;
            r.a[0] = in.px[off]; r.a[1] = in.py[off]; r.a[2] = in.pz[off]; r.a[3] = __uint_as_float(off);
            double x[3], q[3];
            bool accurate;
            body2_point(in, (double)r.a[0], (double)r.a[1], (double)r.a[2], x);
            stencil_point(x, st, q, accurate);
            s_v[grp][st] = query_sdf64(a.grids[in.grid], q[0], q[1], q[2], accurate);
        }
        __syncthreads();
        if (mine && st == 0) {
            double v[kStencil], val, g[3], H[6];
#pragma unroll
            for (int k = 0; k < kStencil; ++k) v[k] = s_v[grp][k];
            stencil_combine(v, val, g, H);
            float* diag = a.slow_diag ? a.slow_diag + 4 * (int64_t)i : nullptr;
            record_slow_fill(a.inst[sd.inst], val, g, H, r, diag);
            mark_slow_index(r, i);
            store_rec(static_cast<Rec*>(a.rec) + sd.slot, r);
        }
        __syncthreads();
    }
}

// Column sums of a wavefront's 64 x N lane values through LDS (s: this wave's
// 64 * N doubles), in a fixed order: lane k < N adds column k as four interleaved
// sums of rows r = 4i + q, combined as (q0 + q1) + (q2 + q3). Returns lane k's
// column sum (0 on lanes >= N). One pass for all N values: no per-value wave
// reduction and no wave-uniform results to hold.
template <int N>
__device__ __forceinline__ double lds_column_sums(double* s, const double (&v)[N]) {
    static_assert(N <= 64, "one column per lane");
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < N; ++k) s[lane * N + k] = v[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double r = 0.0;
    if (lane < N) {
        double q[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int i = 0; i < 64; i += 4)
#pragma unroll
            for (int j = 0; j < 4; ++j) q[j] += s[(i + j) * N + lane];
        r = (q[0] + q[1]) + (q[2] + q[3]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();   // the buffer is free for the wave's next use
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return r;
}

// ------------------------------------------------------------------ batched contact pass
// Work items of kChunk records, one wavefront each: the wrench (and p6 column,
// and the Km/Dm tangent columns) of every record, summed per lane, then over
// the wave in a fixed order into the item's partial row.
template <int MODE, typename R>
__global__ __launch_bounds__(256) void contact_kernel(ContactArgs a) {
    using Rec = typename RecOf<R>::type;
    constexpr bool P6 = MODE >= 1;
    constexpr bool DERIV = MODE >= 2;
    constexpr int NF = P6 ? 24 : 12;
    constexpr int NACC = NF + (DERIV ? 288 : 0);
    // records per lane: a wavefront takes a whole work item (kChunk records, 4 per lane),
    // except with the Km/Dm columns, where the 24 tangent columns and their wave sums
    // dominate and a workgroup takes an item, its wave w records [64 w, 64 w + 64): four
    // times the wavefronts for the same work, which small batches need to fill the GPU
    constexpr int U = DERIV ? 1 : kChunk / 64;
    __shared__ double s_km[DERIV ? 4 : 1][DERIV ? 288 : 1];
    __shared__ double s_red[4][64 * NF];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int total = min(a.counters[CNT_CHUNKS], a.chunk_cap);
    const int j0 = DERIV ? 64 * w : 0;
    for (int c = DERIV ? (int)blockIdx.x : wave_id(); c < total; c += DERIV ? (int)gridDim.x : wave_count()) {
        const ChunkDev ch = as_const(a.chunks)[c];
        const InstDev in = as_const(a.inst)[ch.inst];
        const GridDev G = as_const(a.grids)[in.grid];
        const int n = ch.n;
        Rec* rec = static_cast<Rec*>(a.rec) + ch.rec0;
        const auto act = as_global(a.act) + ch.rec0;
        const InstR<R> I = make_inst<R>(in);
        LaneAcc<R> acc[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) acc[k].zero();
        if (DERIV)
            for (int k = lane; k < 288; k += 64) s_km[DERIV ? w : 0][DERIV ? k : 0] = 0.0;
        // this lane's active-list entries and points, all loaded first
        uint32_t e[U];
        float qx[U], qy[U], qz[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = (j0 + u * 64 + lane < n) ? act[j0 + u * 64 + lane] : kActSlow;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t off = e[u] & ~kActSlow;
            qx[u] = as_global(in.px)[off];
            qy[u] = as_global(in.py)[off];
            qz[u] = as_global(in.pz)[off];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * 64 + lane;
            const bool have = j < n;
            ContactFwd<R> s;
            if (have) {
                Rec r;
                if (e[u] & kActSlow) {
                    r = load_rec(rec + j);   // written by the slow pass
                } else {
                    // fast-path record: the sample from the packed cell, kept for the adjoint
                    double cc[3], fd[3];
                    int idx[3];
                    grid_coords(in, qx[u], qy[u], qz[u], cc);
                    cell_of_c(G, cc, idx, fd);
                    const Corners cr = G.fast ? load_cell(G, idx[0], idx[1], idx[2]) : load_corners(G, idx[0], idx[1], idx[2]);
                    r.a[0] = qx[u]; r.a[1] = qy[u]; r.a[2] = qz[u]; r.a[3] = __uint_as_float(e[u]);
                    record_fast(G, in, cr, fd, r);
                    store_rec(rec + j, r);
                }
                record_decode(r, a.slow_diag, in, s);
                R p6[12];
                contact_forward<R>(I, s, P6 ? p6 : nullptr);
#pragma unroll
                for (int k = 0; k < 12; ++k) acc[k].add(s.f[k]);
                if (P6) {
#pragma unroll
                    for (int k = 0; k < 12; ++k) acc[P6 ? 12 + k : 0].add(p6[k]);
                }
            }
            if (DERIV) {
                // columns of Km (jj < 12) and Dm (jj >= 12); every lane takes part in the sums
                for (int jj = 0; jj < 24; ++jj) {
                    R col[12];
#pragma unroll
                    for (int r = 0; r < 12; ++r) col[r] = (R)0;
                    if (have) contact_jvp<R>(I, s, unit_tangent<R>(jj), col);
                    const int obase = (jj < 12) ? jj : 144 + (jj - 12);
                    double mine = 0.0;   // lane r < 12 keeps row r of the column
#pragma unroll
                    for (int r = 0; r < 12; ++r) {
                        const double v = wave_sum((double)col[r]);
                        mine = (lane == r) ? v : mine;
                    }
                    if (lane < 12) s_km[DERIV ? w : 0][DERIV ? obase + 12 * lane : 0] += mine;
                }
            }
        }
        double* dst = a.partial + (int64_t)c * NACC;
        double lv[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) lv[k] = acc[k].value();
        if (!DERIV) {
            const double col = lds_column_sums<NF>(&s_red[w][0], lv);
            if (lane < NF) dst[lane] = col;
        } else {
            // the item's row. The wrench (and p6) sums in the order of the one-wavefront
            // modes: per lane the records 64 apart in turn (wave order), then the lanes'
            // column sums; the Km/Dm sums in wave order.
#pragma unroll
            for (int k = 0; k < NF; ++k) s_red[w][lane * NF + k] = lv[k];
            __syncthreads();
            if (w == 0) {
#pragma unroll
                for (int k = 0; k < NF; ++k)
                    lv[k] = ((s_red[0][lane * NF + k] + s_red[1][lane * NF + k]) + s_red[2][lane * NF + k]) +
                            s_red[3][lane * NF + k];
                const double col = lds_column_sums<NF>(&s_red[0][0], lv);
                if (lane < NF) dst[lane] = col;
            }
            for (int k = threadIdx.x; k < 288; k += 256)
                dst[NF + k] = ((s_km[0][k] + s_km[DERIV ? 1 : 0][k]) + s_km[DERIV ? 2 : 0][k]) + s_km[DERIV ? 3 : 0][k];
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ fused wrench and adjoint
// Work items of kChunk records, one wavefront each, for a cotangent known when
// the contacts are evaluated (re-evaluation in a backward pass): each record's
// sample is formed as in contact_kernel (the slow pass's records read), its
// wrench summed as contact_kernel's mode 0 sums it and its adjoint as
// vjp_kernel does, per lane over the records 64 apart in turn, then over the
// lanes: the partial row is fm 12 | kVjpOut. No record is written.
template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 1))) void fused_kernel(ContactArgs a) {
    using Rec = typename RecOf<R>::type;
    __shared__ double s_red[4][64 * kVjpOut];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int total = min(a.counters[CNT_CHUNKS], a.chunk_cap);
    for (int c = wave_id(); c < total; c += wave_count()) {
        const ChunkDev ch = as_const(a.chunks)[c];
        const InstDev in = as_const(a.inst)[ch.inst];
        const GridDev G = as_const(a.grids)[in.grid];
        const int n = ch.n;
        const Rec* rec = static_cast<const Rec*>(a.rec) + ch.rec0;
        const auto act = as_global(a.act) + ch.rec0;
        const InstR<R> I = make_inst<R>(in);
        R gb[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) gb[k] = (R)a.g_fm[(int64_t)ch.inst * 12 + k];
        LaneAcc<R> acc[kFmOut];
#pragma unroll
        for (int k = 0; k < kFmOut; ++k) acc[k].zero();
        R vacc[kVjpOut];
#pragma unroll
        for (int k = 0; k < kVjpOut; ++k) vacc[k] = (R)0;
        for (int j = lane; j < n; j += 64) {
            const uint32_t e = act[j];
            Rec r;
            if (e & kActSlow) {
                r = load_rec(rec + j);   // written by the slow pass
            } else {
                const uint32_t off = e;
                const float qx = as_global(in.px)[off], qy = as_global(in.py)[off], qz = as_global(in.pz)[off];
                double cc[3], fd[3];
                int idx[3];
                grid_coords(in, qx, qy, qz, cc);
                cell_of_c(G, cc, idx, fd);
                const Corners cr = G.fast ? load_cell(G, idx[0], idx[1], idx[2]) : load_corners(G, idx[0], idx[1], idx[2]);
                r.a[0] = qx; r.a[1] = qy; r.a[2] = qz; r.a[3] = __uint_as_float(e);
                record_fast(G, in, cr, fd, r);
            }
            ContactFwd<R> s;
            record_decode(r, a.slow_diag, in, s);
            contact_forward<R>(I, s, nullptr);
#pragma unroll
            for (int k = 0; k < kFmOut; ++k) acc[k].add(s.f[k]);
            contact_vjp<R, true>(I, s, gb, vacc);
        }
        double* dst = a.partial + (int64_t)c * (kFmOut + kVjpOut);
        double lf[kFmOut];
#pragma unroll
        for (int k = 0; k < kFmOut; ++k) lf[k] = acc[k].value();
        const double colf = lds_column_sums<kFmOut>(&s_red[w][0], lf);
        if (lane < kFmOut) dst[lane] = colf;
        double lv[kVjpOut];
#pragma unroll
        for (int k = 0; k < kVjpOut; ++k) lv[k] = (double)vacc[k];
        double col = lds_column_sums<kVjpOut>(&s_red[w][0], lv);
        if (lane == 24 || lane == 25) col *= in.kn_chain;   // d/d kn = kn_chain * d/d kn_eff (ForceGeneralSDFContact.cpp:155-156)
        if (lane < kVjpOut) dst[kFmOut + lane] = col;
    }
}

// ------------------------------------------------------------------ batched adjoint
// The adjoint of every record of a work item, with d(loss)/d(wrench) of its
// instance, summed into the item's partial row (kVjpOut).
template <typename R>
__global__ __launch_bounds__(256) void vjp_kernel(VjpArgs a) {
    using Rec = typename RecOf<R>::type;
    __shared__ double s_red[4][64 * kVjpOut];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int total = min(a.counters[CNT_CHUNKS], a.chunk_cap);
    for (int c = wave_id(); c < total; c += wave_count()) {
        const ChunkDev ch = as_const(a.chunks)[c];
        const InstDev in = as_const(a.inst)[ch.inst];
        const int n = ch.n;
        const Rec* rec = static_cast<const Rec*>(a.rec) + ch.rec0;
        const InstR<R> I = make_inst<R>(in);
        R gb[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) gb[k] = (R)a.g_fm[(int64_t)ch.inst * 12 + k];
        R acc[kVjpOut];
#pragma unroll
        for (int k = 0; k < kVjpOut; ++k) acc[k] = (R)0;
        // the next iteration's record is loaded while this one is computed
        Rec cur;
        if (lane < n) cur = load_rec(rec + lane);
        for (int j = lane; j < n; j += 64) {
            Rec nxt;
            if (j + 64 < n) nxt = load_rec(rec + j + 64);
            ContactFwd<R> s;
            record_decode(cur, a.slow_diag, in, s);
            contact_forward<R>(I, s, nullptr);
            contact_vjp<R, true>(I, s, gb, acc);
            cur = nxt;
        }
        double* dst = a.partial + (int64_t)c * kVjpOut;
        double lv[kVjpOut];
#pragma unroll
        for (int k = 0; k < kVjpOut; ++k) lv[k] = (double)acc[k];
        double col = lds_column_sums<kVjpOut>(&s_red[w][0], lv);
        if (lane == 24 || lane == 25) col *= in.kn_chain;   // d/d kn = kn_chain * d/d kn_eff (ForceGeneralSDFContact.cpp:155-156)
        if (lane < kVjpOut) dst[lane] = col;
    }
}

// ------------------------------------------------------------------ per-instance reduction
// One wavefront per instance. The lanes first load the instance's tile entries
// (active count, first work item) 64 tiles at a time; then, lanes over
// components, the work items' partial rows are summed in tile order and each
// tile's items in order, the tile entries broadcast from their lanes, the rows
// listed kRedGroup at a time so that their reads are in flight together. The
// order depends only on the tiling: the same result on every run. Every
// instance's outputs are written (zeros when it has no active contact).
constexpr int kRedGroup = 16;   // partial rows whose reads a reduce wavefront issues together

__global__ __launch_bounds__(256) void reduce_kernel(ReduceArgs a) {
    const int lane = threadIdx.x & 63;
    const int i = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (a.overflow_out && blockIdx.x == 0 && threadIdx.x == 0) *a.overflow_out = a.counters[CNT_OVERFLOW];
    if (i >= a.n_inst) return;
    const InstDev in = as_const(a.inst)[i];
    const int tb = in.tile_begin, nt = in.tile_count;
    const auto part = as_global(a.partial);
    // an evaluation whose records did not fit is incomplete: NaN outputs, n_active -1
    const bool ovf = a.counters[CNT_OVERFLOW] != 0;
    int nact = 0;
    for (int comp0 = 0; comp0 < a.nacc; comp0 += 64) {
        const int comp = comp0 + lane;
        const bool mine = comp < a.nacc;
        double sum = 0.0;
        for (int g0 = 0; g0 < nt; g0 += 64) {
            int cnt = 0, c0 = 0, nch = 0;
            if (g0 + lane < nt) {   // both entries read at once (tile_chunk0 is only meaningful when cnt > 0)
                cnt = a.tile_active[tb + g0 + lane];
                c0 = a.tile_chunk0[tb + g0 + lane];
                nch = (cnt + kChunk - 1) / kChunk;
            }
            if (comp0 == 0) nact += cnt;
            const int ng = min(64, nt - g0);
            // the items of tiles [g0, g0 + ng) in order, kRedGroup at a time: a
            // wave-uniform cursor (tile t, item k) lists their rows, the rows' reads
            // go out together, and the values are added in list order
            int t = 0, k = 0;
            for (;;) {
                int rows[kRedGroup];
                int nr = 0;
#pragma unroll
                for (int j = 0; j < kRedGroup; ++j) {
                    while (t < ng && k >= __builtin_amdgcn_readlane(nch, t)) {
                        ++t;
                        k = 0;
                    }
                    rows[j] = -1;
                    if (t < ng) {
                        rows[j] = __builtin_amdgcn_readlane(c0, t) + k;
                        ++k;
                        nr = j + 1;
                    }
                }
                double v[kRedGroup];
#pragma unroll
                for (int j = 0; j < kRedGroup; ++j)
                    v[j] = (rows[j] >= 0 && mine) ? part[(int64_t)rows[j] * a.nacc + comp] : 0.0;
#pragma unroll
                for (int j = 0; j < kRedGroup; ++j)
                    if (rows[j] >= 0) sum += v[j];
                if (nr < kRedGroup) break;
            }
        }
        if (ovf) sum = __builtin_nan("");
        if (mine) {
            for (int sg = 0; sg < a.nseg; ++sg) {
                const int rel = comp - a.begin[sg];
                if (rel >= 0 && rel < a.width[sg] && a.dst[sg]) a.dst[sg][(int64_t)i * a.width[sg] + rel] = sum;
            }
        }
    }
    if (a.n_active) {
        const int n = (int)wave_sum((double)nact);
        if (lane == 0) a.n_active[i] = ovf ? -1 : n;
    }
    if (a.host_dst) __threadfence_system();   // the stores reach host memory before the launch completes
}

// ------------------------------------------------------------------ grid set-up
__global__ void build_cells_kernel(GridDev G, float4* cells) {
    const int64_t ci = G.n[0], cj = G.n[1], ck = G.n[2], nc = ci * cj * ck;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
        pack_cell(G, (int)(c % ci), (int)((c / ci) % cj), (int)(c / (ci * cj)), cells + 2 * c);
}

__global__ void build_cls_kernel(GridDev G, double col_th, uint8_t* cls) {
    const int64_t ne = (int64_t)cls_bytes(G);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x)
        cls[e] = cls_fill(G, col_th, (uint32_t)e);
}

hipError_t launch_grid_setup(const GridDev& G, float4* cells, hipStream_t st) {
    hipLaunchKernelGGL(build_cells_kernel, dim3(1024), dim3(256), 0, st, G, cells);
    return hipGetLastError();
}

hipError_t launch_cls_setup(const GridDev& G, double col_th, uint8_t* cls, hipStream_t st) {
    hipLaunchKernelGGL(build_cls_kernel, dim3(1024), dim3(256), 0, st, G, col_th, cls);
    return hipGetLastError();
}

// The table's entries: 0 on the lower faces, elsewhere whether block (x-1, y-1, z-1) is open.
__global__ void sat_mark_kernel(GridDev G, const uint8_t* cls, int32_t* sat) {
    const uint32_t dx = sat_dim(G, 0), dy = sat_dim(G, 1), dz = sat_dim(G, 2);
    const int64_t n = (int64_t)dx * dy * dz;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(e % dx), y = (uint32_t)((e / dx) % dy), z = (uint32_t)(e / ((int64_t)dx * dy));
        sat[e] = (x && y && z && sat_block_open(G, cls, x - 1, y - 1, z - 1)) ? 1 : 0;
    }
}

// Running sums along one axis, one lane per line.
__global__ void sat_scan_kernel(GridDev G, int32_t* sat, int axis) {
    const uint32_t d[3] = {sat_dim(G, 0), sat_dim(G, 1), sat_dim(G, 2)};
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int64_t lines = (int64_t)d[a1] * d[a2];
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)d[0] : (int64_t)d[0] * d[1]);
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < lines; l += (int64_t)gridDim.x * blockDim.x) {
        uint32_t c[3];
        c[axis] = 0;
        c[a1] = (uint32_t)(l % d[a1]);
        c[a2] = (uint32_t)(l / d[a1]);
        int32_t* p = sat + sat_index(G, c[0], c[1], c[2]);
        int32_t run = 0;
        for (uint32_t k = 0; k < d[axis]; ++k) {
            run += p[k * stride];
            p[k * stride] = run;
        }
    }
}

hipError_t launch_sat_setup(const GridDev& G, const uint8_t* cls, int32_t* sat, hipStream_t st) {
    hipLaunchKernelGGL(sat_mark_kernel, dim3(512), dim3(256), 0, st, G, cls, sat);
    for (int axis = 0; axis < 3; ++axis) hipLaunchKernelGGL(sat_scan_kernel, dim3(64), dim3(256), 0, st, G, sat, axis);
    return hipGetLastError();
}

// One lane per group, `slots` lanes per tile (so 64 / slots tiles per
// wavefront): the tiles' live-group masks. Block 0 also clears the evaluation's
// counters, which the detection kernel launched after it accumulates into.
__global__ __launch_bounds__(256) void cull_kernel(CullArgs a) {
    if (blockIdx.x == 0 && threadIdx.x < CNT_WORDS) a.counters[threadIdx.x] = 0;
    const int lane = threadIdx.x & 63, S = a.slots;
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int tile = (int)(gi / S), g = (int)(gi % S);
    bool live = false;
    if (tile < a.n_tiles) {
        const TileDev& t = a.tiles[tile];
        const int ng = (t.count + kGroup - 1) / kGroup;
        if (g < ng) {
            live = true;
            const InstDev& in = a.inst[t.inst];
            const GridDev& G = a.grids[t.grid];
            if (G.fast && in.cls && in.csat && in.gbox) {
                const float4* b = reinterpret_cast<const float4*>(in.gbox) + 2 * (t.begin / kGroup + g);
                const float4 l = b[0], h = b[1];
                const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
                live = group_live(G, in, in.csat, lo, hi);
            }
        }
    }
    const uint64_t m = __ballot(live);
    if ((lane & (S - 1)) == 0 && tile < a.n_tiles)
        a.tile_live[tile] = (uint32_t)((m >> lane) & (S >= 32 ? 0xffffffffull : ((1ull << S) - 1ull)));
}

hipError_t launch_cull(const CullArgs& a, hipStream_t st) {
    if (a.n_tiles <= 0) return hipSuccess;
    const int64_t lanes = (int64_t)a.n_tiles * a.slots;
    hipLaunchKernelGGL(cull_kernel, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
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
    if (a.grad) {
        R gb[12], v[kVjpOut];
        for (int k = 0; k < 12; ++k) gb[k] = (R)a.g_fm[k];
        contact_vjp<R>(I, s, gb, v);
        v[24] *= (R)in.kn_chain;   // d/d kn = kn_chain * d/d kn_eff (ForceGeneralSDFContact.cpp:155-156)
        v[25] *= (R)in.kn_chain;
        for (int k = 0; k < kVjpOut; ++k) atomicAdd(a.grad + k, (double)v[k]);
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
        double cc[3];
        grid_coords(in, in.px[o], in.py[o], in.pz[o], cc);
        const double c0 = cc[0], c1 = cc[1], c2 = cc[2];
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
template <int BLOCK>
static const void* detect_ptr_b(int ppl) {
    return ppl == 1 ? (const void*)&detect_kernel<BLOCK, 1>
         : ppl == 2 ? (const void*)&detect_kernel<BLOCK, 2> : (const void*)&detect_kernel<BLOCK, 4>;
}

const void* detect_kernel_ptr(int block, int ppl) {
    if (ppl != 1 && ppl != 2 && ppl != 4) return nullptr;
    switch (block) {
        case 64: return detect_ptr_b<64>(ppl);
        case 128: return detect_ptr_b<128>(ppl);
        case 256: return detect_ptr_b<256>(ppl);
        case 512: return detect_ptr_b<512>(ppl);
        case 1024: return detect_ptr_b<1024>(ppl);
        default: return nullptr;
    }
}

hipError_t launch_detect(const DetectArgs& a, int block, int ppl, int grid, size_t lds, hipStream_t st) {
    const void* f = detect_kernel_ptr(block, ppl);
    if (!f) return hipErrorInvalidValue;
    DetectArgs args = a;
    void* params[] = {&args};
    return hipLaunchKernel(f, dim3(grid), dim3(block), params, lds, st);
}

hipError_t launch_slow(const SlowArgs& a, bool precise, int grid, hipStream_t st) {
    if (precise) hipLaunchKernelGGL(slow_kernel<true>, dim3(grid), dim3(64), 0, st, a);
    else hipLaunchKernelGGL(slow_kernel<false>, dim3(grid), dim3(64), 0, st, a);
    return hipGetLastError();
}

template <int MODE, typename R>
static const void* contact_ptr_t() { return (const void*)&contact_kernel<MODE, R>; }

const void* contact_kernel_ptr(int mode, bool precise) {
    if (precise) return mode == 0 ? contact_ptr_t<0, double>() : mode == 1 ? contact_ptr_t<1, double>() : contact_ptr_t<2, double>();
    return mode == 0 ? contact_ptr_t<0, float>() : mode == 1 ? contact_ptr_t<1, float>() : contact_ptr_t<2, float>();
}

hipError_t launch_contact(const ContactArgs& a, int mode, bool precise, int grid, hipStream_t st) {
    ContactArgs args = a;
    void* params[] = {&args};
    return hipLaunchKernel(contact_kernel_ptr(mode, precise), dim3(grid), dim3(256), params, 0, st);
}

hipError_t launch_vjp(const VjpArgs& a, bool precise, int grid, hipStream_t st) {
    if (precise) hipLaunchKernelGGL(vjp_kernel<double>, dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(vjp_kernel<float>, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_fused(const ContactArgs& a, bool precise, int grid, hipStream_t st) {
    if (precise) hipLaunchKernelGGL(fused_kernel<double>, dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(fused_kernel<float>, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

const void* vjp_kernel_ptr(bool precise) {
    return precise ? (const void*)&vjp_kernel<double> : (const void*)&vjp_kernel<float>;
}

hipError_t launch_reduce(const ReduceArgs& a, hipStream_t st) {
    if (a.n_inst <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_kernel, dim3((a.n_inst + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
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
