// Contact points of a mesh body: the reference's sampling (BodyMeshObj.cpp:184-238),
// kept identical because the point set is part of the contact force's input.
#include "Body/ContactPoints.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace redmax {

std::vector<Vector3> precompute_contact_points(const std::vector<Vector3>& V, const std::vector<std::array<int, 3>>& F,
                                               bool adaptive_sample, int* sample_rate) {
    std::srand(1000);
    const int rate = adaptive_sample ? std::max((int)std::floor(V.size() / 1000.0), 1) : 1;
    if (sample_rate) *sample_rate = rate;
    std::vector<Vector3> out;
    for (size_t i = 0; i < V.size(); ++i)
        if (std::rand() % rate == 0) out.push_back(V[i]);
    if (!adaptive_sample) return out;
    // R2 sequence (Roberts, "The unreasonable effectiveness of quasirandom
    // sequences"): step 1/g and 1/g^2 with g the plastic number, in FP32
    const float g = 1.32471795724474602596f;
    const float a1 = 1.0f / g, a2 = a1 * a1, seed = 0.5f, min_area = 0.5f;
    for (const auto& f : F) {
        if (std::rand() % rate != 0) continue;
        const Vector3 &A = V[(size_t)f[0]], &B = V[(size_t)f[1]], &C = V[(size_t)f[2]];
        auto dist = [](const Vector3& p, const Vector3& q) {
            const dtype x = p(0) - q(0), y = p(1) - q(1), z = p(2) - q(2);
            return std::sqrt(x * x + y * y + z * z);
        };
        const dtype a = dist(B, C), b = dist(C, A), c = dist(A, B);
        const dtype s = 0.5 * (a + b + c);
        const dtype area = std::sqrt(s * (s - a) * (s - b) * (s - c));   // Heron
        if (!(area > min_area)) continue;
        const int n = (int)std::floor(area / min_area);
        for (int j = 0; j < n; ++j) {
            float z1 = std::fmod(seed + a1 * (float)(j + 1), 1.0f);
            float z2 = std::fmod(seed + a2 * (float)(j + 1), 1.0f);
            if (z1 + z2 > 1.0f) {   // fold the unit square onto the triangle
                z1 = 1.0f - z1;
                z2 = 1.0f - z2;
            }
            const dtype w0 = (dtype)(1.0f - z1 - z2), w1 = (dtype)z1, w2 = (dtype)z2;
            out.push_back(make_vector3(w0 * A(0) + w1 * B(0) + w2 * C(0), w0 * A(1) + w1 * B(1) + w2 * C(1),
                                       w0 * A(2) + w1 * B(2) + w2 * C(2)));
        }
    }
    return out;
}

}  // namespace redmax
