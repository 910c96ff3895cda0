// math helpers of the standalone host layer: the subset of the reference's
// simulation/redmax/Utils.h that the SDF-contact force and its self-tests use.
#pragma once
#include <cstdio>
#include "Common.h"

namespace redmax {

namespace constants {
const dtype eps = (dtype)1e-8;   // Utils.h:9
}

namespace math {

// se(3) exponential of a twist [w; v] (Utils.h:126-159): Rodrigues' formula for the
// rotation and the left Jacobian applied to v for the translation.
inline Matrix4 exp(const Vector6& q) {
    const dtype wx0 = q(0), wy0 = q(1), wz0 = q(2);
    const dtype wlen = std::sqrt(wx0 * wx0 + wy0 * wy0 + wz0 * wz0);
    Matrix4 E = Matrix4::Identity();
    if (wlen > 1e-12) {
        const dtype wx = wx0 / wlen, wy = wy0 / wlen, wz = wz0 / wlen;
        const dtype c = std::cos(wlen), s = std::sin(wlen), c1 = 1.0 - c;
        E(0, 0) = c + wx * wx * c1;  E(0, 1) = -wz * s + wx * wy * c1; E(0, 2) = wy * s + wx * wz * c1;
        E(1, 0) = wz * s + wx * wy * c1; E(1, 1) = c + wy * wy * c1;  E(1, 2) = -wx * s + wy * wz * c1;
        E(2, 0) = -wy * s + wx * wz * c1; E(2, 1) = wx * s + wy * wz * c1; E(2, 2) = c + wz * wz * c1;
        // p = ((w . v) theta) w + (I - R)(w x v), with unit w and v scaled by 1/theta
        const dtype vx = q(3) / wlen, vy = q(4) / wlen, vz = q(5) / wlen;
        const dtype cx = wy * vz - wz * vy, cy = wz * vx - wx * vz, cz = wx * vy - wy * vx;
        const dtype wv = wx * vx + wy * vy + wz * vz;
        const dtype cc[3] = {cx, cy, cz}, w[3] = {wx, wy, wz};
        for (int r = 0; r < 3; ++r) {
            dtype d = cc[r];
            for (int k = 0; k < 3; ++k) d -= E(r, k) * cc[k];
            E(r, 3) = wv * wlen * w[r] + d;
        }
    } else {
        E(0, 3) = q(3); E(1, 3) = q(4); E(2, 3) = q(5);
    }
    return E;
}

template <int R, int K, int C>
inline FixedMatrix<R, C> mul(const FixedMatrix<R, K>& A, const FixedMatrix<K, C>& B) {
    FixedMatrix<R, C> out;
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) {
            dtype s = 0;
            for (int k = 0; k < K; ++k) s += A(r, k) * B(k, c);
            out(r, c) = s;
        }
    return out;
}

}  // namespace math

// print_error (Utils.h:371-381): prints when error_rel > 1e-5 and error > 1e-5 with
// error = ||a - b||_F and error_rel = error / max(min(||a||, ||b||), 1e-4). Returns error_rel.
inline dtype print_error(const std::string& name, const MatrixX& a, const MatrixX& b) {
    dtype e = 0, na = 0, nb = 0;
    for (int r = 0; r < a.rows(); ++r)
        for (int c = 0; c < a.cols(); ++c) {
            const dtype d = a(r, c) - b(r, c);
            e += d * d;
            na += a(r, c) * a(r, c);
            nb += b(r, c) * b(r, c);
        }
    e = std::sqrt(e);
    const dtype lo = std::min(std::sqrt(na), std::sqrt(nb));
    const dtype rel = e / (lo > 1e-4 ? lo : 1e-4);
    if (rel > 1e-5 && e > 1e-5) std::printf("%s: error = %.10lf, error_rel = %.10lf\n", name.c_str(), e, rel);
    return rel;
}

}  // namespace redmax
