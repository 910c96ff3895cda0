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

static GridDev host_grid(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3],
                         const double d[3]) {
    GridDev G;
    G.v = vals;
    for (int a = 0; a < 3; ++a) {
        G.n[a] = n[a];
        G.minb[a] = minb[a];
        G.maxb[a] = maxb[a];
        G.d[a] = d[a];
        G.inv_d[a] = 1.0 / d[a];
    }
    G.pad = 0;
    return G;
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
        const bool act = detect_value(G, in, in.px[o], in.py[o], in.pz[o]) < 0.0f;
        if (active) active[o] = act ? 1 : 0;
        if (!act) continue;
        ++na;
        ContactFwd<R> s;
        contact_sample<R>(G, in, (double)in.px[o], (double)in.py[o], (double)in.pz[o], s);
        if (slow) slow[o] = s.slow ? 1 : 0;
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

extern "C" {

// One force object on the host: points are FP32 SoA (as stored on the GPU).
int hc_eval(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3], const double d[3],
            const float* px, const float* py, const float* pz, int64_t npts, const atac_instance* inst, uint32_t flags,
            double* fm, double* p6, double* Km, double* Dm, int64_t* n_active, uint8_t* active, uint8_t* slow,
            const double* g_fm, double* vjp) {
    GridDev G = host_grid(vals, n, minb, maxb, d);
    InstDev in;
    std::memset(&in, 0, sizeof in);
    compose_instance(*inst, G, px, py, pz, npts, in);
    if (flags & ATAC_EVAL_PRECISE)
        eval_t<double>(G, in, flags, fm, p6, Km, Dm, n_active, active, slow, g_fm, vjp);
    else
        eval_t<float>(G, in, flags, fm, p6, Km, Dm, n_active, active, slow, g_fm, vjp);
    return 0;
}

// Detection values of every point (query_SDF + col_th, +inf outside).
int hc_detect(const float* vals, const int32_t n[3], const double minb[3], const double maxb[3], const double d[3],
              const float* px, const float* py, const float* pz, int64_t npts, const atac_instance* inst, float* out) {
    GridDev G = host_grid(vals, n, minb, maxb, d);
    InstDev in;
    std::memset(&in, 0, sizeof in);
    compose_instance(*inst, G, px, py, pz, npts, in);
    for (int64_t o = 0; o < npts; ++o) out[o] = detect_value(G, in, px[o], py[o], pz[o]);
    return 0;
}
}
