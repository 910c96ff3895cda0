#include "Body/MeshObj.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace redmax {

namespace {

dtype det3(const Vector3& a, const Vector3& b, const Vector3& c) {
    return a(0) * (b(1) * c(2) - b(2) * c(1)) - a(1) * (b(0) * c(2) - b(2) * c(0)) + a(2) * (b(0) * c(1) - b(1) * c(0));
}

// vertex index of an OBJ face corner ("i", "i/t", "i//n", "i/t/n"), 0-based
int corner_index(const std::string& tok, int n_vertices, const std::string& filename) {
    const long v = std::strtol(tok.c_str(), nullptr, 10);
    const long k = v > 0 ? v - 1 : (long)n_vertices + v;
    if (v == 0 || k < 0 || k >= n_vertices)
        throw_error("load_obj: face index " + tok + " out of range in " + filename);
    return (int)k;
}

}  // namespace

TriMesh load_obj(const std::string& filename, const Vector3& scale) {
    std::ifstream in(filename.c_str());
    if (!in.good()) throw_error("load_obj: cannot read " + filename);
    TriMesh m;
    bool first_shape_done = false;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            double x = 0, y = 0, z = 0;
            ls >> x >> y >> z;
            m.V.push_back(make_vector3(x * scale(0), y * scale(1), z * scale(2)));
        } else if (tag == "f") {
            if (first_shape_done) continue;
            std::vector<int> c;
            std::string tok;
            while (ls >> tok) c.push_back(corner_index(tok, (int)m.V.size(), filename));
            for (size_t k = 1; k + 1 < c.size(); ++k) m.F.push_back({c[0], c[k], c[k + 1]});
        } else if ((tag == "g" || tag == "o") && !m.F.empty()) {
            first_shape_done = true;
        }
    }
    if (m.V.empty() || m.F.empty()) throw_error("load_obj: no vertices or faces in " + filename);
    return m;
}

void compute_mass_property(const TriMesh& m, dtype& volume, Vector3& COM, Matrix3& I) {
    volume = 0.0;
    COM.setZero();
    for (const auto& f : m.F) {
        const Vector3 &a = m.V[f[0]], &b = m.V[f[1]], &c = m.V[f[2]];
        const dtype vol = det3(a, b, c);   // 6x the signed volume of (origin, a, b, c)
        volume += vol;
        for (int k = 0; k < 3; ++k) COM(k) += vol * (a(k) + b(k) + c(k));
    }
    for (int k = 0; k < 3; ++k) COM(k) /= volume * 4.0;
    volume /= 6.0;
    // second moments about the centre of mass, for unit mass
    dtype diag[3] = {0, 0, 0}, offd[3] = {0, 0, 0};
    for (const auto& f : m.F) {
        Vector3 r[3];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) r[v](k) = m.V[f[v]](k) - COM(k);
        const dtype d = det3(r[0], r[1], r[2]);
        for (int j = 0; j < 3; ++j) {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            diag[j] += (r[0](j) * r[1](j) + r[1](j) * r[2](j) + r[2](j) * r[0](j) + r[0](j) * r[0](j) +
                        r[1](j) * r[1](j) + r[2](j) * r[2](j)) * d;
            offd[j] += (r[0](j1) * r[1](j2) + r[1](j1) * r[2](j2) + r[2](j1) * r[0](j2) + r[0](j1) * r[2](j2) +
                        r[1](j1) * r[0](j2) + r[2](j1) * r[1](j2) + r[0](j1) * r[0](j2) * 2 + r[1](j1) * r[1](j2) * 2 +
                        r[2](j1) * r[2](j2) * 2) * d;
        }
    }
    for (int j = 0; j < 3; ++j) {
        diag[j] /= volume * 60.0;
        offd[j] /= volume * 120.0;
    }
    I(0, 0) = diag[1] + diag[2];
    I(1, 1) = diag[0] + diag[2];
    I(2, 2) = diag[0] + diag[1];
    I(0, 1) = I(1, 0) = -offd[2];
    I(0, 2) = I(2, 0) = -offd[1];
    I(1, 2) = I(2, 1) = -offd[0];
}

void symmetric_eigen3(const Matrix3& A_in, Vector3& values, Matrix3& vectors) {
    double A[3][3], V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = 0.5 * (A_in(i, j) + A_in(j, i));
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double on = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-32 * on || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                // rotation in the (p, q) plane that zeroes A[p][q]
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {   // A <- A G
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {   // A <- G^T A
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {   // V <- V G
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int order[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (A[order[j]][order[j]] < A[order[i]][order[i]]) std::swap(order[i], order[j]);
    for (int c = 0; c < 3; ++c) {
        const int s = order[c];
        values(c) = A[s][s];
        double nrm = 0.0;
        for (int k = 0; k < 3; ++k) nrm += V[k][s] * V[k][s];
        nrm = std::sqrt(nrm);
        int big = 0;
        for (int k = 1; k < 3; ++k)
            if (std::fabs(V[k][s]) > std::fabs(V[big][s])) big = k;
        const double sg = V[big][s] < 0 ? -1.0 : 1.0;
        for (int k = 0; k < 3; ++k) vectors(k, c) = sg * V[k][s] / nrm;
    }
}

MassFrame process_mesh(TriMesh& m, dtype density) {
    MassFrame out;
    Vector3 COM;
    Matrix3 I;
    compute_mass_property(m, out.volume, COM, I);
    out.mass = out.volume * density;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) I(i, j) *= out.mass;
    Vector3 ev;
    Matrix3 R;
    symmetric_eigen3(I, ev, R);
    for (int k = 0; k < 3; ++k) {
        out.inertia(k) = ev(k);
        out.inertia(3 + k) = out.mass;
    }
    // right-handed axes
    const dtype handed = (R(1, 0) * R(2, 1) - R(2, 0) * R(1, 1)) * R(0, 2) +
                         (R(2, 0) * R(0, 1) - R(0, 0) * R(2, 1)) * R(1, 2) +
                         (R(0, 0) * R(1, 1) - R(1, 0) * R(0, 1)) * R(2, 2);
    if (handed < 0.0)
        for (int k = 0; k < 3; ++k) R(k, 2) = -R(k, 2);
    out.E_oi = Matrix4::Identity();
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) out.E_oi(i, j) = R(i, j);
        out.E_oi(i, 3) = COM(i);
    }
    // vertices into the body frame: x_i = R^T x_o + p_io, p_io = -R^T COM (the inverse of E_oi)
    Vector3 p_io;
    for (int i = 0; i < 3; ++i) p_io(i) = -(R(0, i) * COM(0) + R(1, i) * COM(1) + R(2, i) * COM(2));
    for (auto& v : m.V) {
        const Vector3 x = v;
        for (int i = 0; i < 3; ++i) v(i) = R(0, i) * x(0) + R(1, i) * x(1) + R(2, i) * x(2) + p_io(i);
    }
    return out;
}

}  // namespace redmax
