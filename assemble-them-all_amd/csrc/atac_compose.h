// Host-side composition of the per-instance device record from the C-ABI
// instance (poses, twists, force parameters, grid box and spacing).
#pragma once
#include <cstring>
#include "atac.h"
#include "atac_types.h"
#include "contact_math.h"

namespace atac {

// The FP32 copies of an instance's poses, twists and parameters, which the FP32
// contact arithmetic loads (make_inst<float>); every InstDev the kernels read
// needs them, whoever builds it.
inline void set_f32_copies(InstDev& d) {
    for (int k = 0; k < 9; ++k) { d.fR1[k] = (float)d.R1[k]; d.fR2[k] = (float)d.R2[k]; }
    for (int k = 0; k < 3; ++k) {
        d.fw1[k] = (float)d.w1[k]; d.fv1[k] = (float)d.v1[k];
        d.fw2[k] = (float)d.w2[k]; d.fv2[k] = (float)d.v2[k];
    }
    d.fkn = (float)d.kn; d.fkt = (float)d.kt; d.fmu = (float)d.mu;
    d.fdamping = (float)d.damping; d.fcontact_scale = (float)d.contact_scale;
    d.fpad0 = 0.0f;
    for (float& f : d.fpad) f = 0.0f;
}

inline void compose_instance(const atac_instance& s, const GridDev& G, const float* px, const float* py, const float* pz,
                             int64_t n_points, InstDev& d) {
    double R1[9], p1[3], R2[9], p2[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) { R1[3 * r + c] = s.E1[4 * r + c]; R2[3 * r + c] = s.E2[4 * r + c]; }
        p1[r] = s.E1[4 * r + 3];
        p2[r] = s.E2[4 * r + 3];
    }
    // A = diag(1/d) R2^T R1, b = diag(1/d) (R2^T (p1 - p2) - min_box)
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += R2[3 * k + r] * R1[3 * k + c];
            d.A[3 * r + c] = acc * G.inv_d[r];
        }
        double t = 0.0;
        for (int k = 0; k < 3; ++k) t += R2[3 * k + r] * (p1[k] - p2[k]);
        d.bs[r] = (t - G.minb[r]) * G.inv_d[r] + 1.0;
        d.cmax[r] = (G.maxb[r] - G.minb[r]) * G.inv_d[r];
    }
    std::memcpy(d.R1, R1, sizeof R1);
    std::memcpy(d.R2, R2, sizeof R2);
    std::memcpy(d.p1, p1, sizeof p1);
    std::memcpy(d.p2, p2, sizeof p2);
    for (int k = 0; k < 3; ++k) {
        d.w1[k] = s.phi1[k]; d.v1[k] = s.phi1[3 + k];
        d.w2[k] = s.phi2[k]; d.v2[k] = s.phi2[3 + k];
    }
    const double chain = s.scale * s.contact_scale;   // ForceGeneralSDFContact.cpp:155-156
    d.col_th = s.col_th;
    d.kn = s.kn * chain;
    d.kt = s.kt * chain;
    d.mu = s.mu;
    d.damping = s.damping;
    d.contact_scale = s.contact_scale;
    d.kn_chain = chain;
    d.px = px;
    d.py = py;
    d.pz = pz;
    d.pt_count = n_points;
    d.grid = s.grid;
    d.pad = 0;
    set_f32_copies(d);
    d.cls = nullptr;   // set by the batch (its table for this grid and col_th)
}

}  // namespace atac
