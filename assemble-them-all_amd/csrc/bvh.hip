// BVH contact variant: mesh preprocessing on the host, signed-distance and force
// kernels on the GPU (the sampler is bvh.h; the contact math is contact_math.h's,
// with the pseudonormal as gradient and a zero Hessian, BodyBVHObj.cpp:453-566).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <utility>
#include "atac.h"
#include "bvh_host.h"

namespace atac {

// ------------------------------------------------------------------ preprocessing
namespace {

struct V3d {
    double x, y, z;
};
V3d vrow(const std::vector<double>& a, int64_t i) { return V3d{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
V3d vsub(V3d a, V3d b) { return V3d{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3d vcross(V3d a, V3d b) { return V3d{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double vdot(V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
void add_row(std::vector<double>& a, int64_t i, V3d v, double w) {
    a[3 * i] += w * v.x; a[3 * i + 1] += w * v.y; a[3 * i + 2] += w * v.z;
}
void normalize_rows(std::vector<double>& a) {
    for (size_t i = 0; i + 2 < a.size(); i += 3) {
        const double r = std::sqrt(a[i] * a[i] + a[i + 1] * a[i + 1] + a[i + 2] * a[i + 2]);
        if (r > 0.0) { a[i] /= r; a[i + 1] /= r; a[i + 2] /= r; }
    }
}

constexpr int kLeafTris = 4;
constexpr int kMaxDepth = 48;   // < kBvhStack

struct Builder {
    const MeshHost& m;
    std::vector<double> cen;      // centroids, nf x 3
    std::vector<int32_t> idx;     // faces in leaf order
    std::vector<float4> nodes;
    explicit Builder(const MeshHost& mh) : m(mh) {}

    void box_of(int32_t b, int32_t e, float4& lo, float4& hi) const {
        double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int32_t k = b; k < e; ++k)
            for (int c = 0; c < 3; ++c) {
                const int32_t v = m.F[3 * (int64_t)idx[k] + c];
                for (int a = 0; a < 3; ++a) {
                    l[a] = std::min(l[a], m.V[3 * (int64_t)v + a]);
                    h[a] = std::max(h[a], m.V[3 * (int64_t)v + a]);
                }
            }
        lo = make_float4(f32_down(l[0]), f32_down(l[1]), f32_down(l[2]), 0.0f);
        hi = make_float4(f32_up(h[0]), f32_up(h[1]), f32_up(h[2]), 0.0f);
    }
    static float bits(int v) { float f; std::memcpy(&f, &v, sizeof f); return f; }

    // fills node n with faces idx[b, e)
    void build(int32_t n, int32_t b, int32_t e, int depth) {
        float4 lo, hi;
        box_of(b, e, lo, hi);
        if (e - b <= kLeafTris || depth >= kMaxDepth) {
            lo.w = bits(-(b + 1));
            hi.w = bits(e - b);
            nodes[2 * n] = lo; nodes[2 * n + 1] = hi;
            return;
        }
        double cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int32_t k = b; k < e; ++k)
            for (int a = 0; a < 3; ++a) {
                cl[a] = std::min(cl[a], cen[3 * (int64_t)idx[k] + a]);
                ch[a] = std::max(ch[a], cen[3 * (int64_t)idx[k] + a]);
            }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (ch[a] - cl[a] > ch[ax] - cl[ax]) ax = a;
        const int32_t mid = b + (e - b) / 2;
        std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](int32_t p, int32_t q) {
            const double cp = cen[3 * (int64_t)p + ax], cq = cen[3 * (int64_t)q + ax];
            return cp < cq || (cp == cq && p < q);
        });
        const int32_t child = (int32_t)(nodes.size() / 2);
        nodes.resize(nodes.size() + 4);
        lo.w = bits(child);
        hi.w = bits(0);
        nodes[2 * n] = lo; nodes[2 * n + 1] = hi;
        build(child, b, mid, depth + 1);
        build(child + 1, mid, e, depth + 1);
    }
};

}  // namespace

std::string mesh_build(const double* V, int64_t nv, const int64_t* F, int64_t nf, MeshHost& m) {
    if (!V || !F || nv < 3 || nf < 1) return "a mesh needs at least 3 vertices and 1 triangle";
    if (nf >= (int64_t)1 << 30 || nv >= (int64_t)1 << 31) return "mesh too large";
    m.V.assign(V, V + 3 * nv);
    for (int64_t i = 0; i < 3 * nv; ++i)
        if (!std::isfinite(V[i])) return "vertex coordinates must be finite";
    m.F.resize(3 * nf);
    for (int64_t i = 0; i < 3 * nf; ++i) {
        if (F[i] < 0 || F[i] >= nv) return "triangle vertex index out of range";
        m.F[i] = (int32_t)F[i];
    }
    for (int a = 0; a < 3; ++a) { m.bmin[a] = INFINITY; m.bmax[a] = -INFINITY; }
    for (int64_t v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a) {
            m.bmin[a] = std::min(m.bmin[a], V[3 * v + a]);
            m.bmax[a] = std::max(m.bmax[a], V[3 * v + a]);
        }
    // face normals: (v1 - v0) x (v2 - v0), normalised; a face is degenerate when
    // |ab x ac|^2 <= 1e-20 |ab|^2 |ac|^2 (zero area to rounding), and its normal is zero
    m.FN.assign(3 * nf, 0.0);
    std::vector<char> degenerate(nf, 0);
    for (int64_t f = 0; f < nf; ++f) {
        const V3d a = vrow(m.V, m.F[3 * f]), b = vrow(m.V, m.F[3 * f + 1]), c = vrow(m.V, m.F[3 * f + 2]);
        const V3d ab = vsub(b, a), ac = vsub(c, a), n = vcross(ab, ac);
        degenerate[f] = vdot(n, n) <= 1e-20 * vdot(ab, ab) * vdot(ac, ac);
        if (!degenerate[f]) add_row(m.FN, f, n, 1.0);
    }
    normalize_rows(m.FN);
    // edges: the edge opposite corner c of face f is (F[f][c+1], F[f][c+2]); its
    // normal is the normalised sum of the normals of the faces that hold it
    std::map<std::pair<int32_t, int32_t>, int32_t> edge_id;
    m.FE.resize(3 * nf);
    for (int64_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            int32_t u = m.F[3 * f + (c + 1) % 3], w = m.F[3 * f + (c + 2) % 3];
            if (u > w) std::swap(u, w);
            auto it = edge_id.emplace(std::make_pair(u, w), (int32_t)edge_id.size()).first;
            m.FE[3 * f + c] = it->second;
        }
    m.ne = (int64_t)edge_id.size();
    m.EN.assign(3 * m.ne, 0.0);
    for (int64_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) add_row(m.EN, m.FE[3 * f + c], vrow(m.FN, f), 1.0);
    normalize_rows(m.EN);
    // vertex normals: face normals weighted by the face's angle at the vertex
    m.VN.assign(3 * nv, 0.0);
    for (int64_t f = 0; f < nf; ++f)
        for (int c = 0; c < 3; ++c) {
            const V3d p = vrow(m.V, m.F[3 * f + c]);
            const V3d u = vsub(vrow(m.V, m.F[3 * f + (c + 1) % 3]), p), w = vsub(vrow(m.V, m.F[3 * f + (c + 2) % 3]), p);
            const double lu = std::sqrt(vdot(u, u)), lw = std::sqrt(vdot(w, w));
            double cs = (lu > 0.0 && lw > 0.0) ? vdot(u, w) / (lu * lw) : 1.0;
            cs = std::min(1.0, std::max(-1.0, cs));
            add_row(m.VN, m.F[3 * f + c], vrow(m.FN, f), std::acos(cs));
        }
    normalize_rows(m.VN);
    // BVH
    Builder B(m);
    B.cen.resize(3 * nf);
    for (int64_t f = 0; f < nf; ++f)
        for (int a = 0; a < 3; ++a)
            B.cen[3 * f + a] = (m.V[3 * (int64_t)m.F[3 * f] + a] + m.V[3 * (int64_t)m.F[3 * f + 1] + a] +
                                m.V[3 * (int64_t)m.F[3 * f + 2] + a]) / 3.0;
    B.idx.resize(nf);
    std::iota(B.idx.begin(), B.idx.end(), 0);
    B.nodes.resize(2);
    B.build(0, 0, (int32_t)nf, 0);
    m.nodes = std::move(B.nodes);
    m.tri = std::move(B.idx);
    for (int32_t& f : m.tri)
        if (degenerate[f]) f = ~f;
    return "";
}

// ------------------------------------------------------------------ kernels
// Sum over the 64 lanes (every lane active), a fixed butterfly: the same order on
// every run, the result on every lane.
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void mesh_distance_kernel(MeshQueryArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double q0 = a.xw[3 * i] - a.p2[0], q1 = a.xw[3 * i + 1] - a.p2[1], q2 = a.xw[3 * i + 2] - a.p2[2];
    const D3 x = d3(a.R2[0] * q0 + a.R2[3] * q1 + a.R2[6] * q2, a.R2[1] * q0 + a.R2[4] * q1 + a.R2[7] * q2,
                    a.R2[2] * q0 + a.R2[5] * q1 + a.R2[8] * q2);
    double d = 10.0;   // BodyBVHObj.cpp:326-327: points outside the box keep d = 10, normal 0
    D3 nrm = d3(0.0, 0.0, 0.0);
    int f = -1;
    if (mesh_box_filter(a.M, x)) d = mesh_signed_distance(a.M, x, nrm, f) + a.col_th;
    a.d[i] = d;
    if (a.normal) { a.normal[3 * i] = nrm.x; a.normal[3 * i + 1] = nrm.y; a.normal[3 * i + 2] = nrm.z; }
    if (a.face) a.face[i] = f;
}

// One force object over workgroups of 256 (4 wavefronts), kBvhBlockPoints points
// each: wave w of workgroup b takes points b*kBvhBlockPoints + w*64 + lane, + 256,
// ...; a point inside the mesh's box with d = signed distance + col_th < 0 is a
// contact (collision_detection_general_BVH, CollisionDetection.cpp:150-185)
// evaluated with the FP64 contact math. Lane sums in FP64, column sums per wave
// through LDS, the four waves added in order into the workgroup's partial row;
// bvh_reduce_kernel adds the rows in workgroup order: the same result on every run.
__global__ __launch_bounds__(256) void bvh_force_kernel(BvhForceArgs a) {
    constexpr int NF = 24;
    __shared__ double s_red[4][64 * NF];
    __shared__ double s_km[4][288];
    __shared__ double s_row[4][NF];
    __shared__ int s_cnt[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool p6on = (a.flags & ATAC_EVAL_P6) != 0u, deriv = (a.flags & ATAC_EVAL_DERIV) != 0u;
    const InstR<double> I = make_inst<double>(a.in);
    double acc[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) acc[k] = 0.0;
    for (int k = lane; k < 288; k += 64) s_km[w][k] = 0.0;
    int cnt = 0;
    const int64_t b0 = (int64_t)blockIdx.x * kBvhBlockPoints;
    const int64_t b1 = b0 + kBvhBlockPoints < a.n ? b0 + kBvhBlockPoints : a.n;
    for (int64_t base = b0 + (int64_t)w * 64; base < b1; base += 256) {
        const int64_t i = base + lane;
        ContactFwd<double> s;
        bool have = false;
        if (i < b1) {
            double x[3];
            body2_point(a.in, a.xi[3 * i], a.xi[3 * i + 1], a.xi[3 * i + 2], x);
            const D3 xb = d3(x[0], x[1], x[2]);
            if (mesh_box_filter(a.M, xb)) {
                D3 nrm;
                int f;
                const double d = mesh_signed_distance(a.M, xb, nrm, f) + a.in.col_th;
                if (d < 0.0) {
                    have = true;
                    s.xi1 = mk3<double>(a.xi[3 * i], a.xi[3 * i + 1], a.xi[3 * i + 2]);
                    s.x = mk3<double>(x[0], x[1], x[2]);
                    s.d = d;
                    s.g = mk3<double>(nrm.x, nrm.y, nrm.z);
                    for (int k = 0; k < 6; ++k) s.H[k] = 0.0;
                    s.slow = false;
                }
            }
        }
        if (have) {
            double p6[12];
            contact_forward<double>(I, s, p6on ? p6 : nullptr);
#pragma unroll
            for (int k = 0; k < 12; ++k) acc[k] += s.f[k];
            if (p6on) {
#pragma unroll
                for (int k = 0; k < 12; ++k) acc[12 + k] += p6[k];
            }
            ++cnt;
        }
        if (deriv && __ballot(have) != 0ull) {
            // columns of Km (jj < 12) and Dm (jj >= 12): lane r < 12 keeps row r
            for (int jj = 0; jj < 24; ++jj) {
                double col[12];
#pragma unroll
                for (int r = 0; r < 12; ++r) col[r] = 0.0;
                if (have) contact_jvp<double>(I, s, unit_tangent<double>(jj), col);
                const int obase = (jj < 12) ? jj : 144 + (jj - 12);
                double mine = 0.0;
#pragma unroll
                for (int r = 0; r < 12; ++r) {
                    const double v = wave_sum_f64(col[r]);
                    mine = (lane == r) ? v : mine;
                }
                if (lane < 12) s_km[w][obase + 12 * lane] += mine;
            }
        }
    }
    // per-wave column sums, then the waves in order
    double* red = &s_red[w][0];
#pragma unroll
    for (int k = 0; k < NF; ++k) red[lane * NF + k] = acc[k];
    const int wc = (int)wave_sum_f64((double)cnt);
    __syncthreads();
    if (lane < NF) {
        double q[4] = {0.0, 0.0, 0.0, 0.0};
        for (int r = 0; r < 64; r += 4)
            for (int j = 0; j < 4; ++j) q[j] += red[(r + j) * NF + lane];
        s_row[w][lane] = (q[0] + q[1]) + (q[2] + q[3]);
    }
    if (lane == 0) s_cnt[w] = wc;
    __syncthreads();
    double* row = a.part + (int64_t)blockIdx.x * kBvhOut;
    for (int k = threadIdx.x; k < kBvhOut; k += 256) {
        double v;
        if (k < NF) v = ((s_row[0][k] + s_row[1][k]) + s_row[2][k]) + s_row[3][k];
        else v = ((s_km[0][k - NF] + s_km[1][k - NF]) + s_km[2][k - NF]) + s_km[3][k - NF];
        row[k] = v;
    }
    if (threadIdx.x == 0) a.part_cnt[blockIdx.x] = (int64_t)s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

// The workgroups' partial rows added in workgroup order.
__global__ __launch_bounds__(320) void bvh_reduce_kernel(BvhForceArgs a, int nb) {
    const int k = threadIdx.x;
    if (k < kBvhOut) {
        double v = 0.0;
        for (int b = 0; b < nb; ++b) v += a.part[(int64_t)b * kBvhOut + k];
        a.out[k] = v;
    }
    if (k == 0 && a.n_active) {
        int64_t c = 0;
        for (int b = 0; b < nb; ++b) c += a.part_cnt[b];
        *a.n_active = c;
    }
}

hipError_t launch_mesh_distance(const MeshQueryArgs& a, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mesh_distance_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ fixed scratch
// The partial rows and contact counts of the force launches of
// atac_bvh_force_noalloc: kBvhScratchBlocks workgroups' rows (kBvhOut doubles
// each), then their counts (one 8-byte word each). One copy per device, held
// until the process ends. Host code never passes `scratch` itself to a kernel
// (on the host it names a stand-in, not device memory): it uses the address
// hipGetSymbolAddress gave for the device, kept in g_scratch_addr.
static __device__ __align__(256) double scratch[kBvhScratchBlocks * (kBvhOut + 1)];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kScratchDevices] = {};
}  // namespace

hipError_t bvh_scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;   // the lookup needs the device current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_addr[device] = p;
    return hipSuccess;
}

void* bvh_scratch_address(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_addr[device];
}

hipError_t launch_bvh_force(const BvhForceArgs& a, hipStream_t st) {
    const int nb = (int)bvh_force_blocks(a.n);
    hipLaunchKernelGGL(bvh_force_kernel, dim3(nb), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bvh_reduce_kernel, dim3(1), dim3(320), 0, st, a, nb);
    return hipGetLastError();
}

}  // namespace atac
