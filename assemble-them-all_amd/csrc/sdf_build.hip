// Mesh -> signed distance grid on the GPU: the level-set construction of
// sdfgen::make_level_set3 (simulation/externals/sdfgen/makelevelset3.cpp:163-230),
// node for node the same values.
//
// * Band: one wavefront per triangle; lanes over the nodes of its index box
//   widened by the exact band. A node keeps the smallest (distance, triangle)
//   pair (one 64-bit atomic min on distance bits << 32 | triangle): the
//   sequential loop keeps the first triangle reaching the minimum, the lowest
//   index among equal distances, and so does the min. Crossings of the x-rays
//   through the node lines (j, k) with the triangle are counted with atomics.
// * Fast sweeping: within one of the 16 directional sweeps a node reads only its
//   7 neighbours one step back along the sweep axes, all on earlier diagonal
//   planes (a + b + c smaller, a, b, c the steps from the sweep's corner), so the
//   planes run one launch each, every node of a plane in parallel, and the result
//   is that of the sequential sweep.
// * Signs: one thread per x line, parity of the crossings so far.
// FMA contraction is off in this file: every float operation rounds as written.
#pragma clang fp contract(off)

#include "sdf_build.h"
#include "levelset.h"

namespace atac {

using ls::F3;

__device__ __forceinline__ F3 vertex(const float* x, uint32_t v) {
    return ls::f3(x[3 * (int64_t)v], x[3 * (int64_t)v + 1], x[3 * (int64_t)v + 2]);
}
__device__ __forceinline__ F3 grid_node(const LsArgs& a, int i, int j, int k) {
    return ls::f3((float)i * a.dx + a.o[0], (float)j * a.dy + a.o[1], (float)k * a.dz + a.o[2]);
}
__device__ __forceinline__ int64_t node_id(const LsArgs& a, int i, int j, int k) {
    return i + (int64_t)a.ni * (j + (int64_t)a.nj * k);
}

__global__ __launch_bounds__(256) void ls_band_kernel(LsArgs a) {
    const int lane = (int)threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * 4;
    for (int64_t t = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); t < a.nf;
         t += nw) {
        const uint32_t p = a.tri[3 * t], q = a.tri[3 * t + 1], r = a.tri[3 * t + 2];
        const F3 xp = vertex(a.x, p), xq = vertex(a.x, q), xr = vertex(a.x, r);
        // grid coordinates in double
        const double fip = ((double)xp.x - (double)a.o[0]) / (double)a.dx;
        const double fjp = ((double)xp.y - (double)a.o[1]) / (double)a.dy;
        const double fkp = ((double)xp.z - (double)a.o[2]) / (double)a.dz;
        const double fiq = ((double)xq.x - (double)a.o[0]) / (double)a.dx;
        const double fjq = ((double)xq.y - (double)a.o[1]) / (double)a.dy;
        const double fkq = ((double)xq.z - (double)a.o[2]) / (double)a.dz;
        const double fir = ((double)xr.x - (double)a.o[0]) / (double)a.dx;
        const double fjr = ((double)xr.y - (double)a.o[1]) / (double)a.dy;
        const double fkr = ((double)xr.z - (double)a.o[2]) / (double)a.dz;
        const double mi = ls::lmin(fip, ls::lmin(fiq, fir)), Mi = ls::lmax(fip, ls::lmax(fiq, fir));
        const double mj = ls::lmin(fjp, ls::lmin(fjq, fjr)), Mj = ls::lmax(fjp, ls::lmax(fjq, fjr));
        const double mk = ls::lmin(fkp, ls::lmin(fkq, fkr)), Mk = ls::lmax(fkp, ls::lmax(fkq, fkr));
        // distances over the index box widened by the band
        const int i0 = ls::lclamp((int)mi - a.band, 0, a.ni - 1), i1 = ls::lclamp((int)Mi + a.band + 1, 0, a.ni - 1);
        const int j0 = ls::lclamp((int)mj - a.band, 0, a.nj - 1), j1 = ls::lclamp((int)Mj + a.band + 1, 0, a.nj - 1);
        const int k0 = ls::lclamp((int)mk - a.band, 0, a.nk - 1), k1 = ls::lclamp((int)Mk + a.band + 1, 0, a.nk - 1);
        const int nI = i1 - i0 + 1, nJ = j1 - j0 + 1;
        const int64_t tot = (int64_t)nI * nJ * (k1 - k0 + 1);
        for (int64_t e = lane; e < tot; e += 64) {
            const int i = i0 + (int)(e % nI), j = j0 + (int)((e / nI) % nJ), k = k0 + (int)(e / ((int64_t)nI * nJ));
            const float d = ls::tri_dist(grid_node(a, i, j, k), xp, xq, xr);
            if (d < a.phi0)
                atomicMin(a.best + node_id(a, i, j, k), ((unsigned long long)__float_as_uint(d) << 32) | (uint32_t)t);
        }
        // crossings of the x-rays through the node lines (j, k) the triangle spans
        const int cj0 = ls::lclamp((int)ceil(mj), 0, a.nj - 1), cj1 = ls::lclamp((int)floor(Mj), 0, a.nj - 1);
        const int ck0 = ls::lclamp((int)ceil(mk), 0, a.nk - 1), ck1 = ls::lclamp((int)floor(Mk), 0, a.nk - 1);
        if (cj0 > cj1 || ck0 > ck1) continue;
        const int nJ2 = cj1 - cj0 + 1;
        const int64_t tot2 = (int64_t)nJ2 * (ck1 - ck0 + 1);
        for (int64_t e = lane; e < tot2; e += 64) {
            const int j = cj0 + (int)(e % nJ2), k = ck0 + (int)(e / nJ2);
            double ba, bb, bc;
            if (!ls::in_triangle_2d((double)j, (double)k, fjp, fkp, fjq, fkq, fjr, fkr, ba, bb, bc)) continue;
            const double fi = ba * fip + bb * fiq + bc * fir;
            const int ii = (int)ceil(fi);   // the crossing lies in (ii - 1, ii]
            if (ii < 0) atomicAdd(a.count + node_id(a, 0, j, k), 1);
            else if (ii < a.ni) atomicAdd(a.count + node_id(a, ii, j, k), 1);
        }
    }
}

__global__ void ls_unpack_kernel(LsArgs a, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long b = a.best[e];
        const bool none = b == ~0ull;
        a.phi[e] = none ? a.phi0 : __uint_as_float((uint32_t)(b >> 32));
        a.closest[e] = none ? -1 : (int32_t)(uint32_t)(b & 0xffffffffull);
    }
}

// one diagonal plane `p` of the sweep in direction (di, dj, dk)
__global__ void ls_sweep_kernel(LsArgs a, int di, int dj, int dk, int p) {
    const int A = a.ni - 1, B = a.nj - 1;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)B * (a.nk - 1)) return;
    const int sb = (int)(e % B), sc = (int)(e / B), sa = p - sb - sc;
    if (sa < 0 || sa >= A) return;
    const int i = di > 0 ? 1 + sa : a.ni - 2 - sa;
    const int j = dj > 0 ? 1 + sb : a.nj - 2 - sb;
    const int k = dk > 0 ? 1 + sc : a.nk - 2 - sc;
    const F3 gx = grid_node(a, i, j, k);
    const int64_t id = node_id(a, i, j, k);
    float ph = a.phi[id];
    int32_t ct = a.closest[id];
    auto check = [&](int i1, int j1, int k1) {
        const int32_t c = a.closest[node_id(a, i1, j1, k1)];
        if (c < 0) return;
        const uint32_t* t = a.tri + 3 * (int64_t)c;
        const float d = ls::tri_dist(gx, vertex(a.x, t[0]), vertex(a.x, t[1]), vertex(a.x, t[2]));
        if (d < ph) {
            ph = d;
            ct = c;
        }
    };
    check(i - di, j, k);
    check(i, j - dj, k);
    check(i - di, j - dj, k);
    check(i, j, k - dk);
    check(i - di, j, k - dk);
    check(i, j - dj, k - dk);
    check(i - di, j - dj, k - dk);
    a.phi[id] = ph;
    a.closest[id] = ct;
}

__global__ void ls_sign_kernel(LsArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)a.nj * a.nk) return;
    const int j = (int)(e % a.nj), k = (int)(e / a.nj);
    int total = 0;
    for (int i = 0; i < a.ni; ++i) {
        const int64_t id = node_id(a, i, j, k);
        total += a.count[id];
        if (total % 2 == 1) a.phi[id] = -a.phi[id];
    }
}

hipError_t ls_build(const LsArgs& a, hipStream_t st) {
    const int64_t n = (int64_t)a.ni * a.nj * a.nk;
    if (a.nf > 0) {
        const int64_t blocks = std::min<int64_t>((a.nf + 3) / 4, 65536);
        hipLaunchKernelGGL(ls_band_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(ls_unpack_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)), dim3(256), 0, st, a, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int A = a.ni - 1, B = a.nj - 1, C = a.nk - 1;
    if (A > 0 && B > 0 && C > 0) {
        static const int dirs[8][3] = {{+1, +1, +1}, {-1, -1, -1}, {+1, +1, -1}, {-1, -1, +1},
                                       {+1, -1, +1}, {-1, +1, -1}, {+1, -1, -1}, {-1, +1, +1}};
        const unsigned blocks = (unsigned)(((int64_t)B * C + 255) / 256);
        for (int pass = 0; pass < 2; ++pass)
            for (int s = 0; s < 8; ++s)
                for (int p = 0; p <= A + B + C - 3; ++p) {
                    hipLaunchKernelGGL(ls_sweep_kernel, dim3(blocks), dim3(256), 0, st, a, dirs[s][0], dirs[s][1],
                                       dirs[s][2], p);
                    e = hipGetLastError();
                    if (e != hipSuccess) return e;
                }
    }
    hipLaunchKernelGGL(ls_sign_kernel, dim3((unsigned)(((int64_t)a.nj * a.nk + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace atac
