// TEST ONLY: the device arithmetic of the contact library (contact_math.h),
// compiled for the host so the CPU test suite can compare it with the FP64
// oracle without a GPU. The loop bodies mirror detect_kernel, contact_kernel,
// vjp_kernel and collision_kernel; accumulation is plain FP64.
#include <hip/hip_runtime.h>
#include <cstring>
#include <vector>
#include "atac.h"
#include "atac_compose.h"
#include "contact_math.h"

using namespace atac;

// The grid with its corner-packed copy and block bounds built on the host by the
// same per-cell / per-block functions the set-up kernels run.
struct HostGrid {
    std::vector<float4> cells;
    std::vector<uint8_t> cls;   // the classification for the instance's col_th
    GridDev G;
};

// the instance's classification table (what the batch builds on the device)
static void host_cls(HostGrid& hg, InstDev& in) {
    hg.cls.resize(cls_bytes(hg.G));
    for (uint32_t e = 0; e < (uint32_t)hg.cls.size(); ++e) hg.cls[e] = cls_fill(hg.G, in.col_th, e);
    in.cls = hg.cls.data();
}

static void host_grid(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3],
                      const double d[3], HostGrid& hg) {
    GridDev& G = hg.G;
    std::memset(&G, 0, sizeof G);
    G.v = vals;
    for (int a = 0; a < 3; ++a) {
        G.n[a] = n[a];
        G.minb[a] = minb[a];
        G.maxb[a] = maxb[a];
        G.d[a] = d[a];
        G.inv_d[a] = 1.0 / d[a];
    }
    set_fd_reach(G);
    {
        const int64_t ci = n[0], cj = n[1], ck = n[2];
        hg.cells.resize(2 * ci * cj * ck);
        for (int64_t c = 0; c < ci * cj * ck; ++c)
            pack_cell(G, (int)(c % ci), (int)((c / ci) % cj), (int)(c / (ci * cj)), &hg.cells[2 * c]);
        G.cells = hg.cells.data();
        G.fast = 1;
    }
}

template <typename R> struct RecOfH;
template <> struct RecOfH<float> { using type = Rec32; };
template <> struct RecOfH<double> { using type = Rec64; };

// One active point through the batched path's record stages: the detection
// pass's fast-path record or slow marking, the slow pass, then the decode the
// contact and adjoint passes start from.
template <typename R>
static bool sample_via_record(const GridDev& G, const InstDev& in, int64_t o, ContactFwd<R>& s) {
    using Rec = typename RecOfH<R>::type;
    Rec r;
    std::memset(&r, 0, sizeof r);
    const float x = in.px[o], y = in.py[o], z = in.pz[o];
    double cc[3], fd[3];
    int idx[3];
    grid_coords(in, x, y, z, cc);
    r.a[0] = x; r.a[1] = y; r.a[2] = z;
    const int32_t off = (int32_t)o;
    std::memcpy(&r.a[3], &off, sizeof off);
    float diag[4] = {0, 0, 0, 0};
    bool slow = false;
    if (fast_cell_c(G, in, cc, idx, fd)) {
        const Corners cr = G.fast ? load_cell(G, idx[0], idx[1], idx[2]) : load_corners(G, idx[0], idx[1], idx[2]);
        record_fast(G, in, cr, fd, r);
    } else {
        slow = true;
        record_slow(G, in, r, diag);
        if constexpr (sizeof(R) == sizeof(float)) {
            const int32_t si = 0;
            std::memcpy(&r.c[3], &si, sizeof si);
        }
    }
    record_decode(r, diag, in, s);
    return slow;
}

template <typename R>
static void eval_t(const GridDev& G, const InstDev& in, uint32_t flags, double* fm, double* p6, double* Km, double* Dm,
                   int64_t* n_active, uint8_t* active, uint8_t* slow, const double* g_fm, double* vjp) {
    const InstR<R> I = make_inst<R>(in);
    std::memset(fm, 0, 12 * sizeof(double));
    if (p6) std::memset(p6, 0, 12 * sizeof(double));
    if (Km) std::memset(Km, 0, 144 * sizeof(double));
    if (Dm) std::memset(Dm, 0, 144 * sizeof(double));
    if (vjp) std::memset(vjp, 0, 26 * sizeof(double));
    int64_t na = 0;
    for (int64_t o = 0; o < in.pt_count; ++o) {
        const bool act = detect_active(G, in, in.px[o], in.py[o], in.pz[o]);
        if (active) active[o] = act ? 1 : 0;
        if (!act) continue;
        ++na;
        ContactFwd<R> s;
        const bool sl = sample_via_record<R>(G, in, o, s);
        if (slow) slow[o] = sl ? 1 : 0;
        R q[12];
        contact_forward<R>(I, s, q);
        for (int k = 0; k < 12; ++k) fm[k] += (double)s.f[k];
        if (p6) for (int k = 0; k < 12; ++k) p6[k] += (double)q[k];
        if (Km && Dm) {
            for (int j = 0; j < 24; ++j) {
                R col[12];
                contact_jvp<R>(I, s, unit_tangent<R>(j), col);
                for (int r = 0; r < 12; ++r) {
                    if (j < 12) Km[12 * r + j] += (double)col[r];
                    else Dm[12 * r + (j - 12)] += (double)col[r];
                }
            }
        }
        if (vjp && g_fm) {
            R gb[12], v[26];
            for (int k = 0; k < 12; ++k) gb[k] = (R)g_fm[k];
            contact_vjp<R>(I, s, gb, v);
            v[24] *= (R)in.kn_chain;
            v[25] *= (R)in.kn_chain;
            for (int k = 0; k < 26; ++k) vjp[k] += (double)v[k];
        }
    }
    if (n_active) *n_active = na;
}

// collision_kernel's body for point i: BodySDFObj::collision values (11) and,
// when o2 is given, its 135 derivative values.
template <typename R>
static void collision_t(const GridDev& G, const InstDev& in, const double* xw, const double* xwdot, double* v, double* o) {
    const InstR<R> I = make_inst<R>(in);
    ContactFwd<R> s;
    contact_sample<R>(G, in, xw[0], xw[1], xw[2], s);
    s.xw1dot = mk3<R>((R)xwdot[0], (R)xwdot[1], (R)xwdot[2]);
    collision_forward<R>(I, s);
    v[0] = s.d; v[1] = s.n.x; v[2] = s.n.y; v[3] = s.n.z; v[4] = s.ddot;
    v[5] = s.tdot.x; v[6] = s.tdot.y; v[7] = s.tdot.z; v[8] = s.xi2.x; v[9] = s.xi2.y; v[10] = s.xi2.z;
    if (!o) return;
    enum { DD_DXW = 0, DD_DQ2 = 3, DN_DXW = 9, DN_DQ2 = 18, DDD_DXW = 36, DDD_DXWDOT = 39, DDD_DQ2 = 42, DDD_DPHI2 = 48,
           DT_DXW = 54, DT_DXWDOT = 63, DT_DQ2 = 72, DT_DPHI2 = 90, DXI_DXW = 108, DXI_DQ2 = 117 };
    for (int c = 0; c < 18; ++c) {
        v3<R> t6[6];
        for (auto& t : t6) t = zero3<R>();
        v3<R>& sl = t6[c / 3];
        if (c % 3 == 0) sl.x = 1; else if (c % 3 == 1) sl.y = 1; else sl.z = 1;
        CollTangent<R> t;
        collision_jvp<R>(I, s, t6[0], t6[1], t6[2], t6[3], t6[4], t6[5], t);
        const int k = c % 3;
        const R dn[3] = {t.dn.x, t.dn.y, t.dn.z}, dt[3] = {t.dtdot.x, t.dtdot.y, t.dtdot.z};
        const R dx[3] = {t.dxi2.x, t.dxi2.y, t.dxi2.z};
        if (c < 3) {
            o[DD_DXW + k] = t.dd;
            o[DDD_DXW + k] = t.dddot;
            for (int r = 0; r < 3; ++r) { o[DN_DXW + 3 * r + k] = dn[r]; o[DT_DXW + 3 * r + k] = dt[r]; o[DXI_DXW + 3 * r + k] = dx[r]; }
        } else if (c < 6) {
            o[DDD_DXWDOT + k] = t.dddot;
            for (int r = 0; r < 3; ++r) o[DT_DXWDOT + 3 * r + k] = dt[r];
        } else if (c < 12) {
            const int col = c - 6;
            o[DD_DQ2 + col] = t.dd;
            o[DDD_DQ2 + col] = t.dddot;
            for (int r = 0; r < 3; ++r) { o[DN_DQ2 + 6 * r + col] = dn[r]; o[DT_DQ2 + 6 * r + col] = dt[r]; o[DXI_DQ2 + 6 * r + col] = dx[r]; }
        } else {
            const int col = c - 12;
            o[DDD_DPHI2 + col] = t.dddot;
            for (int r = 0; r < 3; ++r) o[DT_DPHI2 + 6 * r + col] = dt[r];
        }
    }
}

extern "C" {

// BodySDFObj::collision of n world points against a grid body at pose E2 with
// twist phi2 (the instance atac_sdf_collision builds: body 1 the world frame).
int hc_collision(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3], const double d[3],
                 const double E2[16], const double phi2[6], double col_th, int precise, const double* xw,
                 const double* xwdot, int64_t npts, double* values, double* derivs) {
    HostGrid hg;
    host_grid(vals, n, minb, maxb, d, hg);
    InstDev in;
    std::memset(&in, 0, sizeof in);
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) { in.R1[3 * r + c] = (r == c) ? 1.0 : 0.0; in.R2[3 * r + c] = E2[4 * r + c]; }
        in.p2[r] = E2[4 * r + 3];
        in.w2[r] = phi2[r];
        in.v2[r] = phi2[3 + r];
    }
    in.col_th = col_th;
    in.contact_scale = 1.0;
    in.kn_chain = 1.0;
    set_f32_copies(in);
    for (int64_t i = 0; i < npts; ++i) {
        double* o = derivs ? derivs + 135 * i : nullptr;
        if (precise) collision_t<double>(hg.G, in, xw + 3 * i, xwdot + 3 * i, values + 11 * i, o);
        else collision_t<float>(hg.G, in, xw + 3 * i, xwdot + 3 * i, values + 11 * i, o);
    }
    return 0;
}

// One force object on the host: points are FP32 SoA (as stored on the GPU).
int hc_eval(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3], const double d[3],
            const float* px, const float* py, const float* pz, int64_t npts, const atac_instance* inst, uint32_t flags,
            double* fm, double* p6, double* Km, double* Dm, int64_t* n_active, uint8_t* active, uint8_t* slow,
            const double* g_fm, double* vjp) {
    HostGrid hg;
    host_grid(vals, n, minb, maxb, d, hg);
    const GridDev& G = hg.G;
    InstDev in;
    std::memset(&in, 0, sizeof in);
    compose_instance(*inst, G, px, py, pz, npts, in);
    host_cls(hg, in);
    if (flags & ATAC_EVAL_PRECISE)
        eval_t<double>(G, in, flags, fm, p6, Km, Dm, n_active, active, slow, g_fm, vjp);
    else
        eval_t<float>(G, in, flags, fm, p6, Km, Dm, n_active, active, slow, g_fm, vjp);
    return 0;
}

// Detection values of every point (query_SDF + col_th, +inf outside).
int hc_detect(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3], const double d[3],
              const float* px, const float* py, const float* pz, int64_t npts, const atac_instance* inst, float* out) {
    HostGrid hg;
    host_grid(vals, n, minb, maxb, d, hg);
    InstDev in;
    std::memset(&in, 0, sizeof in);
    compose_instance(*inst, hg.G, px, py, pz, npts, in);
    host_cls(hg, in);
    for (int64_t o = 0; o < npts; ++o) out[o] = detect_value(hg.G, in, px[o], py[o], pz[o]);
    return 0;
}

// Group culling on the host: the summed-area table built as the device builds it
// (block marks, then running sums along x, y, z), and per group of `gsize`
// consecutive points: live[g] = group_live; needed[g] = some point of the group
// is active (the full interpolation's decision); exact[g] = the table's block count
// over the group's range equals the count of open blocks found one by one and
// live[g] is that count's decision (1 when the range is not finite).
int hc_group_cull(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3],
                  const double d[3], const float* px, const float* py, const float* pz, int64_t npts,
                  const atac_instance* inst, int gsize, uint8_t* live, uint8_t* needed, uint8_t* exact) {
    HostGrid hg;
    host_grid(vals, n, minb, maxb, d, hg);
    const GridDev& G = hg.G;
    InstDev in;
    std::memset(&in, 0, sizeof in);
    compose_instance(*inst, G, px, py, pz, npts, in);
    host_cls(hg, in);
    const uint32_t dim[3] = {sat_dim(G, 0), sat_dim(G, 1), sat_dim(G, 2)};
    std::vector<int32_t> sat(sat_count(G), 0);
    for (uint32_t z = 1; z < dim[2]; ++z)
        for (uint32_t y = 1; y < dim[1]; ++y)
            for (uint32_t x = 1; x < dim[0]; ++x)
                sat[sat_index(G, x, y, z)] = sat_block_open(G, in.cls, x - 1, y - 1, z - 1) ? 1 : 0;
    for (uint32_t z = 0; z < dim[2]; ++z)
        for (uint32_t y = 0; y < dim[1]; ++y)
            for (uint32_t x = 1; x < dim[0]; ++x) sat[sat_index(G, x, y, z)] += sat[sat_index(G, x - 1, y, z)];
    for (uint32_t z = 0; z < dim[2]; ++z)
        for (uint32_t y = 1; y < dim[1]; ++y)
            for (uint32_t x = 0; x < dim[0]; ++x) sat[sat_index(G, x, y, z)] += sat[sat_index(G, x, y - 1, z)];
    for (uint32_t z = 1; z < dim[2]; ++z)
        for (uint32_t y = 0; y < dim[1]; ++y)
            for (uint32_t x = 0; x < dim[0]; ++x) sat[sat_index(G, x, y, z)] += sat[sat_index(G, x, y, z - 1)];
    for (int64_t g = 0; g * gsize < npts; ++g) {
        const int64_t o0 = g * gsize, o1 = std::min<int64_t>(npts, o0 + gsize);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        bool finite = true, need = false;
        for (int64_t o = o0; o < o1; ++o) {
            const float c[3] = {px[o], py[o], pz[o]};
            for (int a = 0; a < 3; ++a) {
                finite = finite && std::isfinite(c[a]);
                lo[a] = std::min(lo[a], c[a]);
                hi[a] = std::max(hi[a], c[a]);
            }
            need = need || detect_value(G, in, px[o], py[o], pz[o]) < 0.0f;
        }
        if (!finite)
            for (int a = 0; a < 3; ++a) lo[a] = hi[a] = NAN;
        live[g] = group_live(G, in, sat.data(), lo, hi) ? 1 : 0;
        needed[g] = need ? 1 : 0;
        uint32_t b0[3], b1[3];
        exact[g] = 1;
        if (group_block_range(G, in, lo, hi, b0, b1)) {
            int32_t brute = 0;
            for (uint32_t z = b0[2]; z < b1[2]; ++z)
                for (uint32_t y = b0[1]; y < b1[1]; ++y)
                    for (uint32_t x = b0[0]; x < b1[0]; ++x) brute += sat_block_open(G, in.cls, x, y, z) ? 1 : 0;
            const int32_t count = sat_range_count(G, sat.data(), b0, b1);
            exact[g] = (brute == count && (live[g] != 0) == (count != 0)) ? 1 : 0;
        }
    }
    return 0;
}

// The bounded detection decision (block bounds, corner-packed cells) next to the
// full interpolation's: decided[o] = detect_active, full[o] = detect_value < 0,
// settled[o] = 1 when the block bound alone settled the point.
int hc_detect_decisions(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3],
                        const double d[3], const float* px, const float* py, const float* pz, int64_t npts,
                        const atac_instance* inst, uint8_t* decided, uint8_t* full, uint8_t* settled,
                        uint8_t* batched, uint8_t* slow, uint8_t* slow_ref) {
    HostGrid hg;
    host_grid(vals, n, minb, maxb, d, hg);
    const GridDev& G = hg.G;
    InstDev in;
    std::memset(&in, 0, sizeof in);
    compose_instance(*inst, G, px, py, pz, npts, in);
    host_cls(hg, in);
    // the detection kernel's passes: 4 points at a time through the block bounds
    // (detect_candidates), then each candidate finished (candidate_active),
    // ragged tail included
    for (int64_t o0 = 0; o0 < npts; o0 += 4) {
        float x[4], y[4], z[4];
        unsigned valid = 0;
        for (int k = 0; k < 4; ++k) {
            const int64_t o = o0 + k < npts ? o0 + k : 0;
            x[k] = px[o]; y[k] = py[o]; z[k] = pz[o];
            if (o0 + k < npts) valid |= 1u << k;
        }
        unsigned open;
        const unsigned cand = detect_candidates<4>(G, in, cls_consts(G, in), x, y, z, valid, false, open, [] {}, [] {});
        for (int k = 0; k < 4 && o0 + k < npts; ++k) {
            bool a = false, s = false;
            if ((cand >> k) & 1u) a = candidate_active(G, in, x[k], y[k], z[k], ((open >> k) & 1u) != 0u, s);
            batched[o0 + k] = a ? 1 : 0;
            slow[o0 + k] = s ? 1 : 0;
        }
    }
    for (int64_t o = 0; o < npts; ++o) {
        double cc[3];
        grid_coords(in, px[o], py[o], pz[o], cc);
        slow_ref[o] = batched[o] && !fast_cell_test(G, in, cc);
    }
    for (int64_t o = 0; o < npts; ++o) {
        decided[o] = detect_active(G, in, px[o], py[o], pz[o]) ? 1 : 0;
        full[o] = detect_value(G, in, px[o], py[o], pz[o]) < 0.0f ? 1 : 0;
        settled[o] = 0;
        if (!G.fast) continue;
        double cc[3];
        grid_coords(in, px[o], py[o], pz[o], cc);
        settled[o] = in.cls[cls_offset(G, in, px[o], py[o], pz[o])] != kClsOpen ? 1 : 0;
    }
    return 0;
}
}
