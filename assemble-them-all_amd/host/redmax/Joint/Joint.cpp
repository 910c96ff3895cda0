#include "Joint/Joint.h"

#include <cmath>

namespace redmax {

namespace {

// Hyper-dual number v + a e1 + b e2 + ab e1 e2 (e1^2 = e2^2 = 0): evaluating a function
// at x + e1 u + e2 w gives its value, the two directional derivatives and the second
// derivative u^T H w, exactly up to rounding.
struct HD {
    double v = 0, a = 0, b = 0, ab = 0;
};
HD operator+(HD x, HD y) { return {x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab}; }
HD operator-(HD x, HD y) { return {x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab}; }
HD operator*(HD x, HD y) {
    return {x.v * y.v, x.v * y.a + x.a * y.v, x.v * y.b + x.b * y.v, x.v * y.ab + x.a * y.b + x.b * y.a + x.ab * y.v};
}
HD operator*(double s, HD x) { return {s * x.v, s * x.a, s * x.b, s * x.ab}; }
// f(x) from f, f', f'' at x.v
HD chain(HD x, double f, double f1, double f2) { return {f, f1 * x.a, f1 * x.b, f1 * x.ab + f2 * x.a * x.b}; }
HD inv(HD x) { return chain(x, 1.0 / x.v, -1.0 / (x.v * x.v), 2.0 / (x.v * x.v * x.v)); }
HD operator/(HD x, HD y) { return x * inv(y); }
HD hsin(HD x) { return chain(x, std::sin(x.v), std::cos(x.v), -std::sin(x.v)); }
HD hsqrt(HD x) {
    const double s = std::sqrt(x.v);
    return chain(x, s, 0.5 / s, -0.25 / (s * x.v));
}

// sum_k coef[k] (-s)^k by Horner's rule
HD series(HD s, const double* coef, int n) {
    HD r{coef[n - 1], 0, 0, 0};
    const HD ms = -1.0 * s;
    for (int k = n - 2; k >= 0; --k) r = r * ms + HD{coef[k], 0, 0, 0};
    return r;
}

// c = sin(t)/t, a = (1 - cos t)/t^2, b = (t - sin t)/t^3 as functions of s = t^2; Taylor
// series below t = 0.5 (11 terms: the first omitted term is below 1e-20 of the sum)
// 1/(2k+1)!, 1/(2k+2)!, 1/(2k+3)!, k < 11
struct SeriesCoefs {
    double fc[11], fa[11], fb[11];
    SeriesCoefs() {
        double fact[24];
        fact[0] = 1.0;
        for (int n = 1; n < 24; ++n) fact[n] = fact[n - 1] * n;
        for (int k = 0; k < 11; ++k) {
            fc[k] = 1.0 / fact[2 * k + 1];
            fa[k] = 1.0 / fact[2 * k + 2];
            fb[k] = 1.0 / fact[2 * k + 3];
        }
    }
};

void so3_coeffs(HD s, HD& c, HD& a, HD& b) {
    if (s.v < 0.25) {
        static const SeriesCoefs k;
        c = series(s, k.fc, 11);
        a = series(s, k.fa, 11);
        b = series(s, k.fb, 11);
    } else {
        const HD t = hsqrt(s), sn = hsin(t), hs = hsin(0.5 * t);
        c = sn / t;
        a = (2.0 * (hs * hs)) / s;
        b = (t - sn) / (s * t);
    }
}

}  // namespace

void so3_derivatives(const double r[3], So3Derivs& out, bool second) {
    // one evaluation per pair j <= k of directions (the second derivatives are
    // symmetric), or per direction j alone without them
    for (int j = 0; j < 3; ++j)
        for (int k = second ? j : 0; k < (second ? 3 : 1); ++k) {
            HD x[3];
            for (int i = 0; i < 3; ++i) x[i] = {r[i], i == j ? 1.0 : 0.0, second && i == k ? 1.0 : 0.0, 0.0};
            const HD s = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            HD c, a, b;
            so3_coeffs(s, c, a, b);
            // K = [r]x, K^2 = r r^T - s I
            HD K[9] = {HD{}, -1.0 * x[2], x[1], x[2], HD{}, -1.0 * x[0], -1.0 * x[1], x[0], HD{}};
            HD K2[9];
            for (int u = 0; u < 3; ++u)
                for (int w = 0; w < 3; ++w) K2[3 * u + w] = x[u] * x[w] - (u == w ? s : HD{});
            for (int e = 0; e < 9; ++e) {
                const HD I{(e % 4 == 0) ? 1.0 : 0.0, 0, 0, 0};
                const HD R = I + c * K[e] + a * K2[e];
                const HD J = I - a * K[e] + b * K2[e];
                if (j == 0 && k == 0) {
                    out.R[e] = R.v;
                    out.Jr[e] = J.v;
                }
                if (k == (second ? j : 0)) {
                    out.dR[j][e] = R.a;
                    out.dJr[j][e] = J.a;
                }
                if (second) {
                    out.d2R[j][k][e] = out.d2R[k][j][e] = R.ab;
                    out.d2Jr[j][k][e] = out.d2Jr[k][j][e] = J.ab;
                }
            }
        }
}

Joint::Joint(std::string name, int id, Type type, const Matrix3& R, const Vector3& p)
    : _name(std::move(name)), _id(id), _type(type) {
    _ndof = type == Type::Fixed ? 0 : (type == Type::Translational ? 3 : 6);
    _q = VectorX::Zero(_ndof);
    _qdot = VectorX::Zero(_ndof);
    _E_pj_0 = Matrix4::Identity();
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) _E_pj_0(i, j) = R(i, j);
        _E_pj_0(i, 3) = p(i);
    }
    _S_j = MatrixX::Zero(6, _ndof);
    _S_j_dot = MatrixX::Zero(6, _ndof);
    _dSj_dq.assign(_ndof, MatrixX::Zero(6, _ndof));
    _dSjdot_dq.assign(_ndof, MatrixX::Zero(6, _ndof));
    update();
}

void Joint::update(bool derivs) {
    _Q = Matrix4::Identity();
    if (_type == Type::Translational) {
        for (int i = 0; i < 3; ++i) {
            _Q(i, 3) = _q(i);
            _S_j(3 + i, i) = 1.0;   // phi = [0; qdot]: S_j_dot and the q-derivatives are zero
        }
    } else if (_type == Type::Free3DExp) {
        // q = [p; r]: Q = [R(r) p; 0 1], phi = [Jr rdot; R^T pdot]
        const double r[3] = {_q(3), _q(4), _q(5)}, rd[3] = {_qdot(3), _qdot(4), _qdot(5)};
        So3Derivs d;
        so3_derivatives(r, d, derivs);
        double Rdot[9] = {0}, Jdot[9] = {0};
        for (int m = 0; m < 3; ++m)
            for (int e = 0; e < 9; ++e) {
                Rdot[e] += d.dR[m][e] * rd[m];
                Jdot[e] += d.dJr[m][e] * rd[m];
            }
        for (int i = 0; i < 3; ++i) {
            for (int c = 0; c < 3; ++c) {
                _Q(i, c) = d.R[3 * i + c];
                _S_j(i, 3 + c) = d.Jr[3 * i + c];
                _S_j(3 + i, c) = d.R[3 * c + i];
                _S_j_dot(i, 3 + c) = Jdot[3 * i + c];
                _S_j_dot(3 + i, c) = Rdot[3 * c + i];
            }
            _Q(i, 3) = _q(i);
        }
        for (int m = 0; m < 3 && derivs; ++m) {
            MatrixX& dS = _dSj_dq[3 + m];
            MatrixX& dSd = _dSjdot_dq[3 + m];
            for (int i = 0; i < 3; ++i)
                for (int c = 0; c < 3; ++c) {
                    dS(i, 3 + c) = d.dJr[m][3 * i + c];
                    dS(3 + i, c) = d.dR[m][3 * c + i];
                    double sj = 0.0, sr = 0.0;
                    for (int j = 0; j < 3; ++j) {
                        sj += d.d2Jr[j][m][3 * i + c] * rd[j];
                        sr += d.d2R[j][m][3 * c + i] * rd[j];
                    }
                    dSd(i, 3 + c) = sj;
                    dSd(3 + i, c) = sr;
                }
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += _E_pj_0(i, k) * _Q(k, j);
            _E_0j(i, j) = s;
        }
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int k = 0; k < _ndof; ++k) s += _S_j(i, k) * _qdot(k);
        _phi(i) = s;
    }
}

bool Joint::reparam() {
    if (_type != Type::Free3DExp) return false;
    const double kPi = 3.14159265358979323846;
    double r[3] = {_q(3), _q(4), _q(5)};
    double n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(n > 1.5 * kPi)) return false;
    So3Derivs d0, d1;
    so3_derivatives(r, d0);
    while (n > 1.5 * kPi) {
        for (int k = 0; k < 3; ++k) r[k] *= 1.0 - 2.0 * kPi / n;
        n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    }
    so3_derivatives(r, d1);
    // rdot1 = Jr(r1)^-1 Jr(r0) rdot0 (the same angular velocity)
    double w[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) w[i] += d0.Jr[3 * i + c] * _qdot(3 + c);
    double A[9];
    for (int e = 0; e < 9; ++e) A[e] = d1.Jr[e];
    // Cramer's rule on the 3 x 3 system A rdot = w
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                       A[2] * (A[3] * A[7] - A[4] * A[6]);
    double sol[3];
    for (int c = 0; c < 3; ++c) {
        double B[9];
        for (int e = 0; e < 9; ++e) B[e] = A[e];
        for (int i = 0; i < 3; ++i) B[3 * i + c] = w[i];
        sol[c] = (B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                  B[2] * (B[3] * B[7] - B[4] * B[6])) / det;
    }
    for (int k = 0; k < 3; ++k) {
        _q(3 + k) = r[k];
        _qdot(3 + k) = sol[k];
    }
    update();
    return true;
}

void Joint::computeJointForce(VectorX& fr) const {
    for (int i = 0; i < _ndof; ++i) {
        fr(_index[i]) += -_Dr * _qdot(i);
        if (_q(i) < _joint_limit_lower) fr(_index[i]) += _joint_limit_stiffness * (_joint_limit_lower - _q(i));
        if (_q(i) > _joint_limit_upper) fr(_index[i]) += _joint_limit_stiffness * (_joint_limit_upper - _q(i));
    }
}

void Joint::computeJointForceWithDerivative(VectorX& fr, MatrixX& Kr, MatrixX& Dr) const {
    computeJointForce(fr);
    for (int i = 0; i < _ndof; ++i) {
        Dr(_index[i], _index[i]) -= _Dr;
        if (_q(i) < _joint_limit_lower || _q(i) > _joint_limit_upper) Kr(_index[i], _index[i]) -= _joint_limit_stiffness;
    }
}

}  // namespace redmax
