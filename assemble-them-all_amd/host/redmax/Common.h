// Minimal dense types for the standalone host layer.
//
// The reference's redmax namespace types are Eigen matrices with dtype = double
// (simulation/redmax/Common.h:24-56). The SDF-contact force in this directory uses
// only element access, sizes, Zero/Identity and resize on them, so it compiles
// against either these types or the reference's Eigen typedefs unchanged.
#pragma once
#include <algorithm>
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

namespace redmax {

using dtype = double;

// Column-major R x C matrix of doubles (fixed size).
template <int R, int C>
class FixedMatrix {
public:
    FixedMatrix() { setZero(); }
    static FixedMatrix Zero() { return FixedMatrix(); }
    static FixedMatrix Identity() {
        FixedMatrix m;
        for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = 1.0;
        return m;
    }
    static FixedMatrix Ones() {
        FixedMatrix m;
        for (int i = 0; i < R * C; ++i) m.a_[i] = 1.0;
        return m;
    }
    dtype& operator()(int r, int c) { return a_[r + R * c]; }
    dtype operator()(int r, int c) const { return a_[r + R * c]; }
    dtype& operator()(int i) { return a_[i]; }
    dtype operator()(int i) const { return a_[i]; }
    dtype& operator[](int i) { return a_[i]; }
    dtype operator[](int i) const { return a_[i]; }
    int rows() const { return R; }
    int cols() const { return C; }
    int size() const { return R * C; }
    void setZero() {
        for (int i = 0; i < R * C; ++i) a_[i] = 0.0;
    }
    dtype* data() { return a_; }
    const dtype* data() const { return a_; }

private:
    dtype a_[R * C];
};

using Vector3 = FixedMatrix<3, 1>;
using Vector4 = FixedMatrix<4, 1>;
using Vector6 = FixedMatrix<6, 1>;
using RowVector3 = FixedMatrix<1, 3>;
using RowVector6 = FixedMatrix<1, 6>;
using Matrix3 = FixedMatrix<3, 3>;
using Matrix4 = FixedMatrix<4, 4>;
using Matrix6 = FixedMatrix<6, 6>;
using Matrix36 = FixedMatrix<3, 6>;
using SE3 = Matrix4;
using se3 = Vector6;

inline Vector3 make_vector3(dtype x, dtype y, dtype z) {
    Vector3 v;
    v(0) = x; v(1) = y; v(2) = z;
    return v;
}

// Column-major dynamic-size matrix.
class MatrixX {
public:
    MatrixX() = default;
    MatrixX(int r, int c) : r_(r), c_(c), a_((size_t)r * c, 0.0) {}
    static MatrixX Zero(int r, int c) { return MatrixX(r, c); }
    dtype& operator()(int r, int c) { return a_[(size_t)r + (size_t)r_ * c]; }
    dtype operator()(int r, int c) const { return a_[(size_t)r + (size_t)r_ * c]; }
    int rows() const { return r_; }
    int cols() const { return c_; }
    int size() const { return r_ * c_; }
    void setZero() { std::fill(a_.begin(), a_.end(), 0.0); }
    void resize(int r, int c) {
        r_ = r;
        c_ = c;
        a_.assign((size_t)r * c, 0.0);
    }
    dtype* data() { return a_.data(); }
    const dtype* data() const { return a_.data(); }

private:
    int r_ = 0, c_ = 0;
    std::vector<dtype> a_;
};

class VectorX {
public:
    VectorX() = default;
    explicit VectorX(int n) : a_((size_t)n, 0.0) {}
    static VectorX Zero(int n) { return VectorX(n); }
    dtype& operator()(int i) { return a_[(size_t)i]; }
    dtype operator()(int i) const { return a_[(size_t)i]; }
    dtype& operator[](int i) { return a_[(size_t)i]; }
    dtype operator[](int i) const { return a_[(size_t)i]; }
    int size() const { return (int)a_.size(); }
    int rows() const { return (int)a_.size(); }
    int cols() const { return 1; }
    void setZero() { std::fill(a_.begin(), a_.end(), 0.0); }
    void resize(int n) { a_.assign((size_t)n, 0.0); }
    dtype* data() { return a_.data(); }
    const dtype* data() const { return a_.data(); }

private:
    std::vector<dtype> a_;
};

// throw_error (simulation/redmax/Common.h:58-61): prints and throws the message.
inline void throw_error(std::string message) {
    std::cerr << "[Error] " << message << std::endl;
    throw message;
}

}  // namespace redmax
