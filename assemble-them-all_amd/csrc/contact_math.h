// Per-point and per-contact arithmetic of the point-vs-SDF contact path.
//
// Everything here is __host__ __device__ so the same code runs inside the gfx950
// kernels and in the host-side check library the CPU tests load. The functions
// restate, in their own form, the reference routines:
//   trilinear query / FD gradient / FD Hessian   Body/BodySDFObj.cpp:123-205
//   collision values                             Body/BodySDFObj.cpp:239-274
//   collision derivatives                        Body/BodySDFObj.cpp:278-377
//   penalty + damping + friction wrench          Force/ForceGeneralSDFContact.cpp:147-219
//   Km / Dm (as a tangent map)                   Force/ForceGeneralSDFContact.cpp:221-447
//   contact-scale column of dfm_dp               Force/ForceGeneralSDFContact.cpp:693-699 (+ friction p6 lines)
//
// SDF sampling has two paths that agree in exact arithmetic:
//  * fast (FP32): when every point of the reference's FD stencil (x +- 1e-5 along
//    each axis, and the half-step mixed stencil) lies in the cell of x, the
//    central difference of the trilinear interpolant IS its partial derivative
//    and the mixed stencil IS the bilinear cross term, while the diagonal
//    stencil is exactly 0. These are read off the 8 corners in FP32.
//  * slow (FP64): otherwise (x within 1e-5 of a cell face or of the box) the
//    stencils are evaluated operation for operation in FP64, reproducing the
//    reference's blend of neighbouring cells and its outside-the-box rule.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include "atac_types.h"

#define ATAC_HD __host__ __device__ __forceinline__

namespace atac {

// Address-space views for device code. Data that no kernel writes while it runs
// (instances, grids, tiles) is read through the constant space, so wave-uniform
// reads become scalar loads into SGPRs; grid and point arrays are read through
// the global space (plain global loads instead of generic flat ones).
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T*;
template <typename T>
using global_ptr = const __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ const_ptr<T> as_const(const T* p) { return (const_ptr<T>)p; }
template <typename T>
__device__ __forceinline__ global_ptr<T> as_global(const T* p) { return (global_ptr<T>)p; }
#define ATAC_GLOBAL(p) ::atac::as_global(p)
#else
// host pass (and the host-compiled check library): plain pointers
template <typename T>
using const_ptr = const T*;
template <typename T>
using global_ptr = const T*;
template <typename T>
ATAC_HD const T* as_const(const T* p) { return p; }
template <typename T>
ATAC_HD const T* as_global(const T* p) { return p; }
#define ATAC_GLOBAL(p) (p)
#endif

// Program-order pins for device code: `pin(v)` makes v computed at this point
// (the compiler may not sink its computation below it); no instruction is
// scheduled across `sched_fence()`. No-ops on the host.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ void pin(unsigned& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }
#else
ATAC_HD void pin(unsigned&) {}
ATAC_HD void sched_fence() {}
#endif

// ------------------------------------------------------------------ small vectors
template <typename T>
struct v3 {
    T x, y, z;
};
template <typename T> ATAC_HD v3<T> mk3(T x, T y, T z) { return v3<T>{x, y, z}; }
template <typename T> ATAC_HD v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> ATAC_HD v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> ATAC_HD v3<T> operator-(v3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> ATAC_HD v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> ATAC_HD v3<T> operator*(v3<T> a, T s) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> ATAC_HD v3<T>& operator+=(v3<T>& a, v3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <typename T> ATAC_HD T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> ATAC_HD v3<T> cross(v3<T> a, v3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> ATAC_HD T norm(v3<T> a) { return sqrt(dot(a, a)); }
template <typename T> ATAC_HD v3<T> zero3() { return {T(0), T(0), T(0)}; }

// 3x3, row-major
template <typename T>
struct m3 {
    T a[9];
};
template <typename T> ATAC_HD v3<T> mv(const m3<T>& M, v3<T> v) {
    return {M.a[0] * v.x + M.a[1] * v.y + M.a[2] * v.z, M.a[3] * v.x + M.a[4] * v.y + M.a[5] * v.z,
            M.a[6] * v.x + M.a[7] * v.y + M.a[8] * v.z};
}
template <typename T> ATAC_HD v3<T> mtv(const m3<T>& M, v3<T> v) {
    return {M.a[0] * v.x + M.a[3] * v.y + M.a[6] * v.z, M.a[1] * v.x + M.a[4] * v.y + M.a[7] * v.z,
            M.a[2] * v.x + M.a[5] * v.y + M.a[8] * v.z};
}
// symmetric 3x3 stored as xx yy zz xy xz yz
template <typename T> ATAC_HD v3<T> symv(const T H[6], v3<T> v) {
    return {H[0] * v.x + H[3] * v.y + H[4] * v.z, H[3] * v.x + H[1] * v.y + H[5] * v.z,
            H[4] * v.x + H[5] * v.y + H[2] * v.z};
}

// ------------------------------------------------------------------ SDF sampling
constexpr double kFdEps = 1e-5;        // BodySDFObj.cpp:178,191
constexpr double kCellMargin = 1e-9;   // cell units kept clear of the stencil edge

ATAC_HD float lerpf(float a, float b, float t) { return (1.0f - t) * a + t * b; }
ATAC_HD double lerpd(double a, double b, double t) { return (1.0 - t) * a + t * b; }

ATAC_HD int64_t grid_index(const GridDev& G, int i, int j, int k) {
    return (int64_t)i + (int64_t)G.n[0] * ((int64_t)j + (int64_t)G.n[1] * (int64_t)k);
}

// Trilinear interpolation at x (SDF body frame) in FP64, in the reference's
// i -> j -> k order and with its i_pos clamp at the last node
// (BodySDFObj.cpp:130-155). Caller guarantees x is inside [min_box, max_box].
ATAC_HD double trilinear64_inside(const GridDev& G, double x0, double x1, double x2) {
    double ci = (x0 - G.minb[0]) / G.d[0], cj = (x1 - G.minb[1]) / G.d[1], ck = (x2 - G.minb[2]) / G.d[2];
    int i = (int)floor(ci), j = (int)floor(cj), k = (int)floor(ck);
    ci = ci - i; cj = cj - j; ck = ck - k;
    i = i < 0 ? 0 : (i > G.n[0] - 1 ? G.n[0] - 1 : i);
    j = j < 0 ? 0 : (j > G.n[1] - 1 ? G.n[1] - 1 : j);
    k = k < 0 ? 0 : (k > G.n[2] - 1 ? G.n[2] - 1 : k);
    int ip = (i == G.n[0] - 1) ? i : i + 1;
    int jp = (j == G.n[1] - 1) ? j : j + 1;
    int kp = (k == G.n[2] - 1) ? k : k + 1;
    const float* v = G.v;
    double a = lerpd((double)v[grid_index(G, i, j, k)], (double)v[grid_index(G, ip, j, k)], ci);
    double b = lerpd((double)v[grid_index(G, i, jp, k)], (double)v[grid_index(G, ip, jp, k)], ci);
    double c = lerpd((double)v[grid_index(G, i, j, kp)], (double)v[grid_index(G, ip, j, kp)], ci);
    double d = lerpd((double)v[grid_index(G, i, jp, kp)], (double)v[grid_index(G, ip, jp, kp)], ci);
    return lerpd(lerpd(a, b, cj), lerpd(c, d, cj), ck);
}

// query_SDF(x, outside_accurate) in FP64 (BodySDFObj.cpp:123-174)
ATAC_HD double query_sdf64(const GridDev& G, double x0, double x1, double x2, bool outside_accurate) {
    bool inside = x0 >= G.minb[0] && x0 <= G.maxb[0] && x1 >= G.minb[1] && x1 <= G.maxb[1] &&
                  x2 >= G.minb[2] && x2 <= G.maxb[2];
    if (inside) return trilinear64_inside(G, x0, x1, x2);
    if (!outside_accurate) return (double)FLT_MAX;
    double b0 = fmin(fmax(x0, G.minb[0]), G.maxb[0]);
    double b1 = fmin(fmax(x1, G.minb[1]), G.maxb[1]);
    double b2 = fmin(fmax(x2, G.minb[2]), G.maxb[2]);
    double e0 = x0 - b0, e1 = x1 - b1, e2 = x2 - b2;
    return sqrt(e0 * e0 + e1 * e1 + e2 * e2) + trilinear64_inside(G, b0, b1, b2);
}

// The FD stencils of query_dSDF / query_ddSDF (BodySDFObj.cpp:176-205) as 20
// evaluation points: 0 = x (outside_accurate = false, the value), 1 = x, 2..7 =
// x +- h along x, y, z, 8..19 = the half-step mixed stencils of xy, xz, yz
// (+ +, - -, + -, - +). h = 1e-5.
constexpr int kStencil = 20;

ATAC_HD void stencil_point(const double x[3], int s, double p[3], bool& accurate) {
    const double h = kFdEps, hh = kFdEps / 2;
    p[0] = x[0]; p[1] = x[1]; p[2] = x[2];
    accurate = s != 0;
    if (s >= 2 && s < 8) {
        const int ax = (s - 2) >> 1;
        p[ax] = (s & 1) ? x[ax] - h : x[ax] + h;
    } else if (s >= 8) {
        const int pair = (s - 8) >> 2, k = (s - 8) & 3;   // pair 0 = xy, 1 = xz, 2 = yz
        const int a0 = pair == 2 ? 1 : 0, a1 = pair == 0 ? 1 : 2;
        const double s0 = (k == 0 || k == 2) ? hh : -hh, s1 = (k == 0 || k == 3) ? hh : -hh;
        p[a0] = x[a0] + s0;
        p[a1] = x[a1] + s1;
    }
}

// Value, gradient and Hessian from the 20 stencil values, in the reference's
// operation order.
ATAC_HD void stencil_combine(const double v[kStencil], double& val, double g[3], double H[6]) {
    const double h = kFdEps;
    val = v[0];
    const double xs = v[1];
    g[0] = (v[2] - v[3]) / (2 * h);
    g[1] = (v[4] - v[5]) / (2 * h);
    g[2] = (v[6] - v[7]) / (2 * h);
    H[0] = (v[2] + v[3] - 2 * xs) / (h * h);
    H[1] = (v[4] + v[5] - 2 * xs) / (h * h);
    H[2] = (v[6] + v[7] - 2 * xs) / (h * h);
    H[3] = (v[8] + v[9] - v[10] - v[11]) / (h * h);
    H[4] = (v[12] + v[13] - v[14] - v[15]) / (h * h);
    H[5] = (v[16] + v[17] - v[18] - v[19]) / (h * h);
}

// FP64 value, FD gradient and FD Hessian at x, operation for operation as
// query_SDF / query_dSDF / query_ddSDF (BodySDFObj.cpp:123-205); the axis-stencil
// values are shared between gradient and Hessian diagonal, as they are bitwise
// the same evaluations in the reference.
ATAC_HD void sdf_stencil64(const GridDev& G, const double x[3], double& val, double g[3], double H[6]) {
    double v[kStencil];
    for (int s = 0; s < kStencil; ++s) {
        double p[3];
        bool accurate;
        stencil_point(x, s, p, accurate);
        v[s] = query_sdf64(G, p[0], p[1], p[2], accurate);
    }
    stencil_combine(v, val, g, H);
}

// Cell of x when the whole FD stencil stays inside it (fast path), else false.
// f: the fractions in FP32, fd: the same in FP64.
ATAC_HD bool fast_cell(const GridDev& G, const double x[3], int idx[3], float f[3], double fd[3]) {
    const double h = kFdEps;
    for (int a = 0; a < 3; ++a) {
        if (!(x[a] - h >= G.minb[a] && x[a] + h <= G.maxb[a])) return false;
        // a product with 1/d instead of the reference's division: the result differs by
        // an ulp at most, and cells are only accepted 2e-4 or more away from a face
        double c = (x[a] - G.minb[a]) * G.inv_d[a];
        double fl = floor(c);
        if (!(fl >= 0.0 && fl <= (double)(G.n[a] - 2))) return false;
        double fr = c - fl;
        double m = h * G.inv_d[a] + kCellMargin;
        if (fr < m || fr > 1.0 - m) return false;
        idx[a] = (int)fl;
        f[a] = (float)fr;
        fd[a] = fr;
    }
    return true;
}


struct Corners {
    float v000, v100, v010, v110, v001, v101, v011, v111;
};

ATAC_HD Corners load_corners(const GridDev& G, int i, int j, int k) {
    const auto v = ATAC_GLOBAL(G.v);
    int64_t b = grid_index(G, i, j, k);
    int64_t sj = G.n[0], sk = (int64_t)G.n[0] * G.n[1];
    Corners c;
    c.v000 = v[b];
    c.v100 = v[b + 1];
    c.v010 = v[b + sj];
    c.v110 = v[b + sj + 1];
    c.v001 = v[b + sk];
    c.v101 = v[b + sk + 1];
    c.v011 = v[b + sk + sj];
    c.v111 = v[b + sk + sj + 1];
    return c;
}

// The corners of cell (i, j, k) from the corner-packed copy.
ATAC_HD Corners load_cell(const GridDev& G, int i, int j, int k) {
    const auto q = ATAC_GLOBAL(G.cells) + 2 * ((int64_t)i + (int64_t)G.n[0] * ((int64_t)j + (int64_t)G.n[1] * k));
    const float4 lo = q[0], hi = q[1];
    Corners c;
    c.v000 = lo.x; c.v100 = lo.y; c.v010 = lo.z; c.v110 = lo.w;
    c.v001 = hi.x; c.v101 = hi.y; c.v011 = hi.z; c.v111 = hi.w;
    return c;
}

// Set-up of the corner-packed copy and the block bounds (one cell / one block).
// Cell (i, j, k), 0 <= i < n: the 8 nodes detect_value reads for a point in it,
// the offsets clamped at the last node (a cell of the last layer repeats it).
ATAC_HD void pack_cell(const GridDev& G, int i, int j, int k, float4 out[2]) {
    const auto v = G.v;
    const int64_t b = grid_index(G, i, j, k);
    const int64_t sj = G.n[0], sk = (int64_t)G.n[0] * G.n[1];
    const int64_t di = i < G.n[0] - 1 ? 1 : 0;
    const int64_t dj = j < G.n[1] - 1 ? sj : 0;
    const int64_t dk = k < G.n[2] - 1 ? sk : 0;
    out[0] = make_float4(v[b], v[b + di], v[b + dj], v[b + dj + di]);
    out[1] = make_float4(v[b + dk], v[b + dk + di], v[b + dk + dj], v[b + dk + dj + di]);
}

// Saturating FP64 -> int32 truncation (NaN -> 0), the hardware conversion's rule.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ int cvt_sat_i32(double c) {
    int r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(c));
    return r;
}
// a * b + c for a, b below 2^24 (the full-rate 24-bit multiply-add)
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
#else
inline int cvt_sat_i32(double c) {
    if (c != c) return 0;
    if (c >= 2147483647.0) return 2147483647;
    if (c <= -2147483648.0) return -2147483647 - 1;
    return (int)c;
}
inline uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) { return (a & 0xffffffu) * (b & 0xffffffu) + c; }
#endif
// v clamped to [0, hi] (one v_med3_i32)
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ int med3_i32(int v, int hi) {
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v), "v"(hi));
    return r;
}
#else
inline int med3_i32(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }
#endif

// FP32 roundings of an FP64 value toward -inf and toward +inf.
ATAC_HD float f32_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
ATAC_HD float f32_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

// Margin of the cell bounds (relative to the magnitude of the values they
// bound); the classification's comment gives the reason for its size.
constexpr double kBoundTol = 4e-6;

// Cell classification of a grid for one col_th, one byte per cell of the grid
// padded by one cell on every side (padded index i' = i + 1 in [0, n + 1] per
// axis, i the cell index floor(c)): kClsInactive (no point of the cell is
// active), kClsActive (every inside point is), kClsOpen (the exact path decides).
// The trilinear value of a point is a convex combination of its cell's 8 nodes,
// so it lies in their (min, max), up to the FP32 rounding of the 7 lerps and the
// col_th add: at most 13 units of 2^-24 of M + |col_th|, M = max|node| (each lerp
// adds <= 4 units of its largest input; the add 1 unit). A cell is inactive when
// min - kBoundTol*M > -col_th + kBoundTol*|col_th| and active when
// max + kBoundTol*M < -col_th - kBoundTol*|col_th| (rounded outward): the value
// clears the threshold by 4e-6 (M + |col_th|), 5x the rounding bound, so the
// class decides the point exactly as the interpolation. The stream finds the cell
// with the shifted map (cls_offset), the same cell grid_coords gives. A cell the
// box face cuts, and the clamped layer beyond the last node, are open, so the box
// test of their points is the exact one; cells outside the box are inactive; a
// cell with a non-finite node is open.
constexpr uint8_t kClsInactive = 0, kClsActive = 1, kClsOpen = 2;

ATAC_HD int32_t cls_pad(const GridDev& G, int a) { return G.n[a] + 2; }
ATAC_HD uint32_t cls_bricks(const GridDev& G, int a) { return ((uint32_t)cls_pad(G, a) + 3u) >> 2; }
ATAC_HD uint32_t cls_bytes(const GridDev& G) { return cls_bricks(G, 0) * cls_bricks(G, 1) * cls_bricks(G, 2) * 64u; }
ATAC_HD double grid_cmax(const GridDev& G, int a) { return (G.maxb[a] - G.minb[a]) * G.inv_d[a]; }   // as compose_instance

// Byte offset of padded cell (i', j', k'): bricks of 4x4x4 cells (one 64-byte
// line), bricks x-fastest, cells in a brick x-fastest, so the 64 Morton
// neighbours a load instruction covers share one or two lines.
ATAC_HD uint32_t cls_entry(const GridDev& G, uint32_t i, uint32_t j, uint32_t k) {
    const uint32_t brick = mad24(mad24(k >> 2, cls_bricks(G, 1), j >> 2), cls_bricks(G, 0), i >> 2);
    return (brick << 6) | ((k & 3u) << 4) | ((j & 3u) << 2) | (i & 3u);
}

// The classification of byte e of the table for col_th (cells past the padded
// grid in the last bricks are inactive).
ATAC_HD uint8_t cls_fill(const GridDev& G, double col_th, uint32_t e) {
    const uint32_t brick = e >> 6, b0 = cls_bricks(G, 0), b1 = cls_bricks(G, 1);
    const int ip[3] = {(int)(4u * (brick % b0) + (e & 3u)), (int)(4u * ((brick / b0) % b1) + ((e >> 2) & 3u)),
                       (int)(4u * (brick / (b0 * b1)) + ((e >> 4) & 3u))};
    bool open = false;
    for (int a = 0; a < 3; ++a) {
        if (ip[a] >= cls_pad(G, a)) return kClsInactive;    // brick padding
        const int i = ip[a] - 1;
        const double cm = grid_cmax(G, a);
        if (i < 0 || (double)i > cm) return kClsInactive;   // outside the box
        if (i >= G.n[a] || (double)i + 1.0 > cm) open = true;   // cut by the box face, or beyond the last node
    }
    if (open) return kClsOpen;
    const int c[3] = {ip[0] - 1, ip[1] - 1, ip[2] - 1};
    float lo = INFINITY, hi = -INFINITY;
    bool finite = true;
    for (int kk = c[2]; kk <= (c[2] + 1 < G.n[2] - 1 ? c[2] + 1 : G.n[2] - 1); ++kk)
        for (int jj = c[1]; jj <= (c[1] + 1 < G.n[1] - 1 ? c[1] + 1 : G.n[1] - 1); ++jj)
            for (int ii = c[0]; ii <= (c[0] + 1 < G.n[0] - 1 ? c[0] + 1 : G.n[0] - 1); ++ii) {
                const float v = G.v[grid_index(G, ii, jj, kk)];
                finite = finite && std::isfinite(v);
                lo = fminf(lo, v);
                hi = fmaxf(hi, v);
            }
    if (!finite) return kClsOpen;
    const double m = kBoundTol * fmax(fabs((double)lo), fabs((double)hi));
    const double cth = (double)(float)col_th, mt = kBoundTol * fabs(cth);   // col_th as the detection adds it
    if ((double)f32_down((double)lo - m) > (double)f32_up(mt - cth)) return kClsInactive;
    if ((double)f32_up((double)hi + m) < (double)f32_down(-cth - mt)) return kClsActive;
    return kClsOpen;
}

// FP64 trilinear value of a cell from its FP32 corners, in the reference's
// i -> j -> k order (BodySDFObj.cpp:143-153).
ATAC_HD double cell_value64(const Corners& c, const double f[3]) {
    const double a = lerpd((double)c.v000, (double)c.v100, f[0]), b = lerpd((double)c.v010, (double)c.v110, f[0]);
    const double cc = lerpd((double)c.v001, (double)c.v101, f[0]), d = lerpd((double)c.v011, (double)c.v111, f[0]);
    return lerpd(lerpd(a, b, f[1]), lerpd(cc, d, f[1]), f[2]);
}

// FP32 value, closed-form gradient and mixed partials of the trilinear cell.
ATAC_HD void sdf_cell32(const GridDev& G, const Corners& c, const float f[3], float& val, float g[3], float H[6]) {
    const float fx = f[0], fy = f[1], fz = f[2];
    // value in the reference's i -> j -> k order
    float a = lerpf(c.v000, c.v100, fx), b = lerpf(c.v010, c.v110, fx);
    float cc = lerpf(c.v001, c.v101, fx), d = lerpf(c.v011, c.v111, fx);
    val = lerpf(lerpf(a, b, fy), lerpf(cc, d, fy), fz);
    const float idx = (float)G.inv_d[0], idy = (float)G.inv_d[1], idz = (float)G.inv_d[2];
    // axis differences
    float x00 = c.v100 - c.v000, x10 = c.v110 - c.v010, x01 = c.v101 - c.v001, x11 = c.v111 - c.v011;
    float y00 = c.v010 - c.v000, y10 = c.v110 - c.v100, y01 = c.v011 - c.v001, y11 = c.v111 - c.v101;
    float z00 = c.v001 - c.v000, z10 = c.v101 - c.v100, z01 = c.v011 - c.v010, z11 = c.v111 - c.v110;
    g[0] = lerpf(lerpf(x00, x10, fy), lerpf(x01, x11, fy), fz) * idx;
    g[1] = lerpf(lerpf(y00, y10, fx), lerpf(y01, y11, fx), fz) * idy;
    g[2] = lerpf(lerpf(z00, z10, fx), lerpf(z01, z11, fx), fy) * idz;
    H[0] = 0.0f;
    H[1] = 0.0f;
    H[2] = 0.0f;
    H[3] = lerpf(x10 - x00, x11 - x01, fz) * (idx * idy);
    H[4] = lerpf(x01 - x00, x11 - x10, fy) * (idx * idz);
    H[5] = lerpf(y01 - y00, y11 - y10, fx) * (idy * idz);
}

// Detection test of one point, collision_detection_general_SDF's loop body
// (CollisionDetection.cpp:135-141 -> BodySDFObj::distance, BodySDFObj.cpp:207-213):
// the point maps into grid index space through the FP64 affine map
// c = A xi + b, the cell and fraction are taken in FP64, the 8 corners and the
// trilinear value in FP32. Returns the value query_SDF(x) + col_th, or +inf when
// x is outside [min_box, max_box] (the reference returns FLT_MAX + col_th).
ATAC_HD void grid_coords(const InstDev& in, float px, float py, float pz, double c[3]) {
    const double x = px, y = py, z = pz;   // three fused FP64 steps per axis, host and device alike
    // the shifted map c' = A xi + (b + 1), then c = c' - 1: exact whenever c' >= 1
    // (c' and 1 are multiples of ulp(c')), so floor(c) = floor(c') - 1 for every
    // point the box can hold and the stream's cell (cls_offset) is this one
    c[0] = fma(in.A[2], z, fma(in.A[1], y, fma(in.A[0], x, in.bs[0]))) - 1.0;
    c[1] = fma(in.A[5], z, fma(in.A[4], y, fma(in.A[3], x, in.bs[1]))) - 1.0;
    c[2] = fma(in.A[8], z, fma(in.A[7], y, fma(in.A[6], x, in.bs[2]))) - 1.0;
}

ATAC_HD float detect_value(const GridDev& G, const InstDev& in, float px, float py, float pz) {
    double cc[3];
    grid_coords(in, px, py, pz, cc);
    const double c0 = cc[0], c1 = cc[1], c2 = cc[2];
    const bool inside = c0 >= 0.0 && c0 <= in.cmax[0] && c1 >= 0.0 && c1 <= in.cmax[1] && c2 >= 0.0 && c2 <= in.cmax[2];
    if (!inside) return INFINITY;
    int i = (int)c0, j = (int)c1, k = (int)c2;  // c >= 0: truncation is floor
    i = i < G.n[0] - 1 ? i : G.n[0] - 1;
    j = j < G.n[1] - 1 ? j : G.n[1] - 1;
    k = k < G.n[2] - 1 ? k : G.n[2] - 1;
    const float fx = (float)(c0 - (double)i), fy = (float)(c1 - (double)j), fz = (float)(c2 - (double)k);
    const int64_t sj = G.n[0], sk = (int64_t)G.n[0] * G.n[1];
    const int64_t di = i < G.n[0] - 1 ? 1 : 0;
    const int64_t dj = j < G.n[1] - 1 ? sj : 0;
    const int64_t dk = k < G.n[2] - 1 ? sk : 0;
    const auto v = ATAC_GLOBAL(G.v) + ((int64_t)i + sj * j + sk * k);
    const float a = lerpf(v[0], v[di], fx), b = lerpf(v[dj], v[dj + di], fx);
    const float c = lerpf(v[dk], v[dk + di], fx), d = lerpf(v[dk + dj], v[dk + dj + di], fx);
    return lerpf(lerpf(a, b, fy), lerpf(c, d, fy), fz) + (float)in.col_th;
}

// Nontemporal load of a V at a 32-bit byte offset from a wave-uniform base.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename V, typename T>
__device__ __forceinline__ V nt_load_at(global_ptr<T> base, uint32_t byte_off) {
    using GV = const __attribute__((address_space(1))) V*;
    using GC = const __attribute__((address_space(1))) char*;
    return __builtin_nontemporal_load(reinterpret_cast<GV>(reinterpret_cast<GC>(base) + byte_off));
}
#else
template <typename V, typename T>
ATAC_HD V nt_load_at(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_off);
}
#endif

// The trilinear detection value of a cell's 8 corners (corner-packed order, the
// lerp order of detect_value) at fractions (fx, fy, fz), plus col_th.
ATAC_HD float cell_value(float4 lo, float4 hi, float fx, float fy, float fz, float cth) {
    const float a = lerpf(lo.x, lo.y, fx), b = lerpf(lo.z, lo.w, fx);
    const float c = lerpf(hi.x, hi.y, fx), d = lerpf(hi.z, hi.w, fx);
    return lerpf(lerpf(a, b, fy), lerpf(c, d, fy), fz) + cth;
}

// Exact detection of a point the block bounds left open: the value is interpolated
// as detect_value does, from the corner-packed cell (whose last layer holds the
// clamped corners), or from the 8 nodes on a grid without the table.
ATAC_HD bool detect_interp(const GridDev& G, const InstDev& in, float px, float py, float pz, const double cc[3]) {
    const double c0 = cc[0], c1 = cc[1], c2 = cc[2];
    const bool inside = c0 >= 0.0 && c0 <= in.cmax[0] && c1 >= 0.0 && c1 <= in.cmax[1] && c2 >= 0.0 && c2 <= in.cmax[2];
    if (!inside) return false;
    if (!G.fast) return detect_value(G, in, px, py, pz) < 0.0f;
    int i = (int)c0, j = (int)c1, k = (int)c2;
    i = i < G.n[0] - 1 ? i : G.n[0] - 1;
    j = j < G.n[1] - 1 ? j : G.n[1] - 1;
    k = k < G.n[2] - 1 ? k : G.n[2] - 1;
    const float fx = (float)(c0 - (double)i), fy = (float)(c1 - (double)j), fz = (float)(c2 - (double)k);
    const auto q = ATAC_GLOBAL(G.cells) + 2 * ((int64_t)i + (int64_t)G.n[0] * ((int64_t)j + (int64_t)G.n[1] * k));
    return cell_value(q[0], q[1], fx, fy, fz, (float)in.col_th) < 0.0f;
}

// Groups and their culling. A point set's segment is cut into groups of kGroup
// consecutive (Morton-sorted) points, each with the box of its FP32 coordinates.
// Before detection, a group whose box can only reach cells where no point is
// active (cell_may_be_active) is dropped: none of its points would be found
// active, so the decisions are the same without it. The test counts the
// non-inactive blocks of 2 x 2 x 2 padded cells over the box's cell range with a
// summed-area table: S(x, y, z) (x, y, z from 0 to the block counts, x fastest)
// is the number of such blocks with indices below (x, y, z), zero on the lower faces.
constexpr int kGroup = 256;
ATAC_HD uint32_t sat_dim(const GridDev& G, int a) { return ((uint32_t)cls_pad(G, a) + 1u) / 2u + 1u; }
ATAC_HD size_t sat_count(const GridDev& G) { return (size_t)sat_dim(G, 0) * sat_dim(G, 1) * sat_dim(G, 2); }
ATAC_HD size_t sat_index(const GridDev& G, uint32_t x, uint32_t y, uint32_t z) {
    return ((size_t)z * sat_dim(G, 1) + y) * sat_dim(G, 0) + x;
}

// Can a point in padded cell (i, j, k) be active? Its table entry says so, except
// in the layer beyond the last node (padded index n) of an axis whose box ends
// on that node: the table keeps that layer open for its points exactly on the box
// face, whose values are convex combinations of face nodes of the interior
// neighbour (padded n - 1), while the rest of the layer lies outside the box; so
// there the neighbour's entry decides. (Only the culling uses this; detection
// reads the table itself.)
ATAC_HD bool cell_may_be_active(const GridDev& G, const uint8_t* cls, uint32_t i, uint32_t j, uint32_t k) {
    uint32_t c[3] = {i, j, k};
    for (int a = 0; a < 3; ++a)
        if (c[a] == (uint32_t)G.n[a] && grid_cmax(G, a) == (double)(G.n[a] - 1)) c[a] -= 1u;
    return cls[cls_entry(G, c[0], c[1], c[2])] != kClsInactive;
}

// Is block (bx, by, bz) of the padded grid not inactive (any of its cells)?
ATAC_HD bool sat_block_open(const GridDev& G, const uint8_t* cls, uint32_t bx, uint32_t by, uint32_t bz) {
    bool any = false;
    for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t i = 2 * bx + (c & 1u), j = 2 * by + ((c >> 1) & 1u), k = 2 * bz + (c >> 2);
        if (i < (uint32_t)cls_pad(G, 0) && j < (uint32_t)cls_pad(G, 1) && k < (uint32_t)cls_pad(G, 2))
            any = any || cell_may_be_active(G, cls, i, j, k);
    }
    return any;
}

// Can a point in the box [lo, hi] (body frame) read a classification entry that
// is not inactive? Over the box, coordinate a of the map c' = A x + bs spans
// [bs_a + sum_j min(A_aj lo_j, A_aj hi_j), bs_a + sum_j max(A_aj lo_j, A_aj hi_j)];
// widened by 1e-12 of the terms' magnitude (a point's own FP64 map and this
// interval are each within a few roundings, under 1e-15 of it, of the exact
// values), then truncated and clamped to the padded grid as cls_offset does, it
// holds the cell of every point of the box. The box is live when a block over
// those cells is not inactive, and whenever an input or the interval is not
// finite.
// The cells [c0, c1] (per axis, padded indices) a point of the box can fall in;
// false when an input or the interval is not finite.
ATAC_HD bool group_cell_range(const GridDev& G, const InstDev& in, const float lo[3], const float hi[3],
                              uint32_t c0[3], uint32_t c1[3]) {
    for (int a = 0; a < 3; ++a) {
        double sl = in.bs[a], sh = in.bs[a], mag = fabs(in.bs[a]);
        for (int j = 0; j < 3; ++j) {
            if (!(fabs((double)lo[j]) <= 3.4e38) || !(fabs((double)hi[j]) <= 3.4e38)) return false;
            const double p = in.A[3 * a + j] * (double)lo[j], q = in.A[3 * a + j] * (double)hi[j];
            sl += p < q ? p : q;
            sh += p < q ? q : p;
            mag += fmax(fabs(p), fabs(q));
        }
        const double eps = 1e-12 * mag;
        const double cl = floor(sl - eps), ch = floor(sh + eps);
        if (!(fabs(cl) <= 1e300) || !(fabs(ch) <= 1e300)) return false;   // NaN or inf
        const double top = (double)(G.n[a] + 1);
        c0[a] = (uint32_t)(cl < 0.0 ? 0.0 : (cl > top ? top : cl));
        c1[a] = (uint32_t)(ch < 0.0 ? 0.0 : (ch > top ? top : ch));
    }
    return true;
}

// The summed-area table's blocks [b0, b1) over those cells.
ATAC_HD bool group_block_range(const GridDev& G, const InstDev& in, const float lo[3], const float hi[3],
                               uint32_t b0[3], uint32_t b1[3]) {
    uint32_t c0[3], c1[3];
    if (!group_cell_range(G, in, lo, hi, c0, c1)) return false;
    for (int a = 0; a < 3; ++a) {
        b0[a] = c0[a] >> 1;
        b1[a] = (c1[a] >> 1) + 1u;
    }
    return true;
}

// Non-inactive blocks in [b0, b1) (per axis) by the summed-area table.
ATAC_HD int32_t sat_range_count(const GridDev& G, const int32_t* sat, const uint32_t b0[3], const uint32_t b1[3]) {
    return sat[sat_index(G, b1[0], b1[1], b1[2])] - sat[sat_index(G, b0[0], b1[1], b1[2])] -
           sat[sat_index(G, b1[0], b0[1], b1[2])] - sat[sat_index(G, b1[0], b1[1], b0[2])] +
           sat[sat_index(G, b0[0], b0[1], b1[2])] + sat[sat_index(G, b0[0], b1[1], b0[2])] +
           sat[sat_index(G, b1[0], b0[1], b0[2])] - sat[sat_index(G, b0[0], b0[1], b0[2])];
}

ATAC_HD bool group_live(const GridDev& G, const InstDev& in, const int32_t* sat, const float lo[3], const float hi[3]) {
    uint32_t b0[3], b1[3];
    if (!group_block_range(G, in, lo, hi, b0, b1)) return true;
    return sat_range_count(G, sat, b0, b1) != 0;
}

// The values cls_offset takes besides the map's matrix: the map's addends, the
// padded grid's last cell index per axis and the brick counts along x and y.
struct ClsConsts {
    double b[3];
    int32_t hi[3];
    uint32_t nb0, nb1;
};

ATAC_HD ClsConsts cls_consts(const GridDev& G, const InstDev& in) {
    ClsConsts c;
    for (int a = 0; a < 3; ++a) {
        c.b[a] = in.bs[a];
        c.hi[a] = G.n[a] + 1;
    }
    c.nb0 = cls_bricks(G, 0);
    c.nb1 = cls_bricks(G, 1);
    return c;
}

// cls_consts in vector registers, for a loop over many points: a VOP3
// instruction reads at most one scalar operand, so uniform values left in scalar
// registers are copied into vector registers at every use.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double in_vgpr(double v) {
    double r;
    asm volatile("v_mov_b64 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ uint32_t in_vgpr(uint32_t v) {
    uint32_t r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ ClsConsts cls_consts_vgpr(const GridDev& G, const InstDev& in) {
    ClsConsts c = cls_consts(G, in);
    for (int a = 0; a < 3; ++a) {
        c.b[a] = in_vgpr(c.b[a]);
        c.hi[a] = (int32_t)in_vgpr((uint32_t)c.hi[a]);
    }
    c.nb0 = in_vgpr(c.nb0);
    c.nb1 = in_vgpr(c.nb1);
    return c;
}
#else
ATAC_HD ClsConsts cls_consts_vgpr(const GridDev& G, const InstDev& in) { return cls_consts(G, in); }
#endif

// The table byte offset of a point: its grid coordinates by the shifted map
// c' = A xi + (b + 1), truncated per axis (a negative or NaN coordinate gives
// 0, the outside layer) and clamped to the padded grid; the brick layout of
// cls_entry.
ATAC_HD uint32_t cls_offset(const InstDev& in, const ClsConsts& kc, float px, float py, float pz) {
    const double x = px, y = py, z = pz;
    const double c0 = fma(in.A[2], z, fma(in.A[1], y, fma(in.A[0], x, kc.b[0])));
    const double c1 = fma(in.A[5], z, fma(in.A[4], y, fma(in.A[3], x, kc.b[1])));
    const double c2 = fma(in.A[8], z, fma(in.A[7], y, fma(in.A[6], x, kc.b[2])));
    const uint32_t i = (uint32_t)med3_i32(cvt_sat_i32(c0), kc.hi[0]);
    const uint32_t j = (uint32_t)med3_i32(cvt_sat_i32(c1), kc.hi[1]);
    const uint32_t k = (uint32_t)med3_i32(cvt_sat_i32(c2), kc.hi[2]);
    const uint32_t brick = mad24(mad24(k >> 2, kc.nb1, j >> 2), kc.nb0, i >> 2);
    return (brick << 6) | ((k & 3u) << 4) | ((j & 3u) << 2) | (i & 3u);
}

ATAC_HD uint32_t cls_offset(const GridDev& G, const InstDev& in, float px, float py, float pz) {
    return cls_offset(in, cls_consts(G, in), px, py, pz);
}

// The detection decision detect_value(...) < 0, bit for bit, through the cell
// classification (the interpolation only for open cells).
ATAC_HD bool detect_active(const GridDev& G, const InstDev& in, float px, float py, float pz, double cc[3]) {
    grid_coords(in, px, py, pz, cc);
    if (!G.fast || !in.cls) return detect_value(G, in, px, py, pz) < 0.0f;
    const uint8_t e = in.cls[cls_offset(G, in, px, py, pz)];
    if (e != kClsOpen) return e == kClsActive;
    return detect_interp(G, in, px, py, pz, cc);
}

ATAC_HD bool detect_active(const GridDev& G, const InstDev& in, float px, float py, float pz) {
    double cc[3];
    return detect_active(G, in, px, py, pz, cc);
}

// fast_cell_c without the outputs: does the FD stencil of c stay in its cell?
ATAC_HD bool fast_cell_test(const GridDev& G, const InstDev& in, const double c[3]);

// The FD stencil's reach in cell units (GridDev::fd_m).
ATAC_HD void set_fd_reach(GridDev& G) {
    for (int a = 0; a < 3; ++a) {
        G.fd_m[a] = kFdEps * G.inv_d[a] + kCellMargin;
        G.fd_1m[a] = 1.0 - G.fd_m[a];
    }
}

// fast_cell in grid-index coordinates c = (x - min_box) / d (the detection map):
// the cell is interior, and the FD stencil (+-1e-5 along each axis) stays inside
// both the cell and [min_box, max_box] with kCellMargin cell units to spare.
// Straight-line (no early exit): for c >= 0 the truncation is the floor, and a
// saturated or NaN conversion fails the tests as the floor would.
ATAC_HD bool fast_cell_c(const GridDev& G, const InstDev& in, const double c[3], int idx[3], double fd[3]) {
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        const int i = cvt_sat_i32(c[a]);
        const double fr = c[a] - (double)i, m = G.fd_m[a];
        ok = ok && c[a] >= 0.0 && i <= G.n[a] - 2 && fr >= m && fr <= G.fd_1m[a] && c[a] + m <= in.cmax[a];
        idx[a] = i;
        fd[a] = fr;
    }
    return ok;
}

ATAC_HD bool fast_cell_test(const GridDev& G, const InstDev& in, const double c[3]) {
    int idx[3];
    double fd[3];
    return fast_cell_c(G, in, c, idx, fd);
}

// 16 bytes at a 32-bit byte offset from a wave-uniform base (one scalar base
// and a 32-bit vector offset: no 64-bit address arithmetic per lane).
template <typename T>
ATAC_HD float4 load_f4_at(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <typename T>
ATAC_HD float2 load_f2_at(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) + byte_off);
}
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
__device__ __forceinline__ float4 load_f4_at(global_ptr<T> base, uint32_t byte_off) {
    using G4 = const __attribute__((address_space(1))) float4*;
    using GC = const __attribute__((address_space(1))) char*;
    return *reinterpret_cast<G4>(reinterpret_cast<GC>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ float2 load_f2_at(global_ptr<T> base, uint32_t byte_off) {
    using G2 = const __attribute__((address_space(1))) float2*;
    using GC = const __attribute__((address_space(1))) char*;
    return *reinterpret_cast<G2>(reinterpret_cast<GC>(base) + byte_off);
}
#endif

// A pointer the wavefront holds in scalar registers.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
__device__ __forceinline__ const T* uniform_ptr(const T* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const T*)(((uint64_t)hi << 32) | lo);
}
#else
template <typename T>
ATAC_HD const T* uniform_ptr(const T* p) { return p; }
#endif

// Byte `off` of a classification table of `bytes` bytes: a buffer load with the
// table as a scalar resource and a 32-bit lane offset (bytes past the end read 0).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint32_t cls_load(const uint8_t* tab, uint32_t bytes, uint32_t off) {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void*)uniform_ptr(tab), (short)0, (int)bytes, 0x00020000);
    return __builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, 0);
}
#else
inline uint32_t cls_load(const uint8_t* tab, uint32_t bytes, uint32_t off) { return off < bytes ? tab[off] : 0u; }
#endif

// Stream pass of the detection of PPL points (detect_active without the cell
// reads): once the points' cell-bound offsets are computed, `before()` runs (the
// caller issues reads that can share this step's wait), the cell-bound reads are
// issued, then `issue()` (the caller's next loads go out behind them, so waiting
// for the bounds does not wait for them), then the bounds are applied. Returns
// the candidates (bit k): the points they settle as active and the ones they
// leave open (also in `open`); candidate_active finishes both. Only what every
// point needs is done here, without branches or box tests, so a wavefront's lanes
// stay busy: the padded table gives every point an entry (outside the box:
// kBoundOutside, on a cell the box face cuts: kBoundOpen), and a point that is not
// valid reads entry 0, which is outside. The per-candidate work runs later,
// densely, one candidate per lane. With all_active every valid point is a
// settled candidate and nothing is read; on a grid without the table every valid
// point is open.
template <int PPL, typename F0, typename F>
ATAC_HD unsigned detect_candidates(const GridDev& G, const InstDev& in, const ClsConsts& kc, const float (&x)[PPL],
                                   const float (&y)[PPL], const float (&z)[PPL], unsigned valid, bool all_active,
                                   unsigned& open, F0&& before, F&& issue) {
    uint32_t boff[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) boff[k] = ((valid >> k) & 1u) ? cls_offset(in, kc, x[k], y[k], z[k]) : 0u;
    sched_fence();   // the points are in: the reads below wait for nothing earlier
    before();
    uint32_t cl[PPL];
    const bool use_tab = G.fast && in.cls && !all_active;
    if (use_tab) {
        const uint32_t bytes = cls_bytes(G);
#pragma unroll
        for (int k = 0; k < PPL; ++k) cl[k] = cls_load(in.cls, bytes, boff[k]);
    }
    issue();
    if (all_active) {
        open = 0u;
        return valid;
    }
    if (!use_tab) {
        open = valid;
        return valid;
    }
    unsigned act = 0, opn = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        act |= (cl[k] == kClsActive ? 1u : 0u) << k;
        opn |= (cl[k] == kClsOpen ? 1u : 0u) << k;
    }
    open = opn;
    return act | opn;
}

// A candidate of detect_candidates on a grid with the cell table, in two
// halves so its cell read can share a wait: candidate_read computes the
// coordinates and the FD-stencil test and issues the cell read of an open
// candidate; candidate_finish decides it: active (settled, or open and
// interpolated from its cell as detect_value does), with `slow` set if it is
// active and its FD stencil leaves the cell (!fast_cell_test).
struct CandRead {
    float4 lo, hi;
    float f[3];
    bool open, off_cell, inside;
};

// candidate_read without the cell read: returns the byte offset of the open
// candidate's cell in the table (0 for a settled one, whose cell is not read).
ATAC_HD uint32_t candidate_locate(const GridDev& G, const InstDev& in, float x, float y, float z, bool open,
                                  CandRead& r) {
    // a candidate is inside the box (c >= 0): the truncation is the floor. One
    // truncation per axis serves the cell (clamped at the last node, as
    // detect_value clamps), its fractions and fast_cell_c's test: the fraction in
    // the clamped cell is fast_cell_c's whenever that test can pass (i <= n - 2).
    // An open candidate may lie on a cell the box face cuts: the exact box test
    // decides it (c >= 0 holds: its cell index is not negative).
    double c[3];
    grid_coords(in, x, y, z, c);
    bool ok = true, inside = true;
    uint32_t idx[3];
    for (int a = 0; a < 3; ++a) {
        const int i = cvt_sat_i32(c[a]);
        const int ic = i < G.n[a] - 1 ? i : G.n[a] - 1;
        const double fr = c[a] - (double)ic, m = G.fd_m[a];
        ok = ok && c[a] >= 0.0 && i <= G.n[a] - 2 && fr >= m && fr <= G.fd_1m[a] && c[a] + m <= in.cmax[a];
        inside = inside && c[a] >= 0.0 && c[a] <= in.cmax[a];
        idx[a] = (uint32_t)ic;
        r.f[a] = (float)fr;
    }
    r.off_cell = !ok;
    r.open = open;
    r.inside = inside;
    return open ? (idx[0] + (uint32_t)G.n[0] * mad24(idx[2], (uint32_t)G.n[1], idx[1])) * 32u : 0u;
}

ATAC_HD void candidate_read(const GridDev& G, const InstDev& in, float x, float y, float z, bool open, CandRead& r) {
    const uint32_t off = candidate_locate(G, in, x, y, z, open, r);
    if (open) {
        r.lo = load_f4_at(ATAC_GLOBAL(G.cells), off);
        r.hi = load_f4_at(ATAC_GLOBAL(G.cells), off + 16u);
    }
}

ATAC_HD bool candidate_finish(const CandRead& r, float cth, bool& slow) {
    const bool act = r.open ? r.inside && cell_value(r.lo, r.hi, r.f[0], r.f[1], r.f[2], cth) < 0.0f : true;
    slow = act && r.off_cell;
    return act;
}

ATAC_HD bool candidate_active(const GridDev& G, const InstDev& in, float x, float y, float z, bool open,
                              bool& slow) {
    CandRead r;
    candidate_read(G, in, x, y, z, open, r);
    return candidate_finish(r, (float)in.col_th, slow);
}

// Cell and fractions of grid coordinates c known to pass fast_cell_c (the
// detection pass tested it); the index is clamped to the interior cells, so
// the corner reads stay inside the grid whatever c is.
ATAC_HD void cell_of_c(const GridDev& G, const double c[3], int idx[3], double fd[3]) {
    for (int a = 0; a < 3; ++a) {
        const double fl = floor(c[a]);
        fd[a] = c[a] - fl;
        int i = (int)fl;
        i = i < 0 ? 0 : (i > G.n[a] - 2 ? G.n[a] - 2 : i);
        idx[a] = i;
    }
}

// FP64 value, closed-form gradient and mixed partials of the trilinear cell, from
// the same FP32 corners (the FP64 contact path).
ATAC_HD void sdf_cell64(const GridDev& G, const Corners& c, const double f[3], double& val, double g[3], double H[6]) {
    const double fx = f[0], fy = f[1], fz = f[2];
    const double v000 = c.v000, v100 = c.v100, v010 = c.v010, v110 = c.v110;
    const double v001 = c.v001, v101 = c.v101, v011 = c.v011, v111 = c.v111;
    val = cell_value64(c, f);
    const double x00 = v100 - v000, x10 = v110 - v010, x01 = v101 - v001, x11 = v111 - v011;
    const double y00 = v010 - v000, y10 = v110 - v100, y01 = v011 - v001, y11 = v111 - v101;
    const double z00 = v001 - v000, z10 = v101 - v100, z01 = v011 - v010, z11 = v111 - v110;
    g[0] = lerpd(lerpd(x00, x10, fy), lerpd(x01, x11, fy), fz) * G.inv_d[0];
    g[1] = lerpd(lerpd(y00, y10, fx), lerpd(y01, y11, fx), fz) * G.inv_d[1];
    g[2] = lerpd(lerpd(z00, z10, fx), lerpd(z01, z11, fx), fy) * G.inv_d[2];
    H[0] = H[1] = H[2] = 0.0;
    H[3] = lerpd(x10 - x00, x11 - x01, fz) * (G.inv_d[0] * G.inv_d[1]);
    H[4] = lerpd(x01 - x00, x11 - x10, fy) * (G.inv_d[0] * G.inv_d[2]);
    H[5] = lerpd(y01 - y00, y11 - y10, fx) * (G.inv_d[1] * G.inv_d[2]);
}

// ------------------------------------------------------------------ contact math
// Quotient a / b of the contact math: FP64 divides. FP32 multiplies by a reciprocal
// instead of running the ten-instruction IEEE division sequence: on the GPU the
// hardware one (v_rcp_f32, within 1 ulp of 1/b), in the host build the correctly
// rounded 1.0f / b, so host checks of this code see the same two roundings
// (reciprocal, then product; at most about 2 ulp from a / b in total) and differ from
// the GPU only by v_rcp_f32's last bit.
ATAC_HD double qdiv(double a, double b) { return a / b; }
ATAC_HD float qdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rcpf(b);
#else
    return a * (1.0f / b);
#endif
}

template <typename R>
struct InstR {
    m3<R> R1, R2;
    v3<R> w1, v1, w2, v2;
    R kn, kt, mu, c, cs;
};

template <typename R>
ATAC_HD InstR<R> make_inst(const InstDev& in) {
    InstR<R> I;
    for (int k = 0; k < 9; ++k) { I.R1.a[k] = (R)in.R1[k]; I.R2.a[k] = (R)in.R2[k]; }
    I.w1 = mk3<R>((R)in.w1[0], (R)in.w1[1], (R)in.w1[2]);
    I.v1 = mk3<R>((R)in.v1[0], (R)in.v1[1], (R)in.v1[2]);
    I.w2 = mk3<R>((R)in.w2[0], (R)in.w2[1], (R)in.w2[2]);
    I.v2 = mk3<R>((R)in.v2[0], (R)in.v2[1], (R)in.v2[2]);
    I.kn = (R)in.kn; I.kt = (R)in.kt; I.mu = (R)in.mu; I.c = (R)in.damping; I.cs = (R)in.contact_scale;
    return I;
}

// FP32: the rounded copies kept in the instance record (the same values as the
// conversions above, loaded as scalars instead of converted per lane)
template <>
ATAC_HD InstR<float> make_inst<float>(const InstDev& in) {
    InstR<float> I;
    for (int k = 0; k < 9; ++k) { I.R1.a[k] = in.fR1[k]; I.R2.a[k] = in.fR2[k]; }
    I.w1 = mk3<float>(in.fw1[0], in.fw1[1], in.fw1[2]);
    I.v1 = mk3<float>(in.fv1[0], in.fv1[1], in.fv1[2]);
    I.w2 = mk3<float>(in.fw2[0], in.fw2[1], in.fw2[2]);
    I.v2 = mk3<float>(in.fv2[0], in.fv2[1], in.fv2[2]);
    I.kn = in.fkn; I.kt = in.fkt; I.mu = in.fmu; I.c = in.fdamping; I.cs = in.fcontact_scale;
    return I;
}

// Forward state of one active contact (values of BodySDFObj::collision and of
// the force loop body, plus what the tangent and adjoint maps reuse).
template <typename R>
struct ContactFwd {
    v3<R> xi1, x;       // contact point (body 1 frame), its SDF-body-frame position
    R d;                // query_SDF(x) + col_th
    v3<R> g;            // grad phi (not normalised)
    R H[6];             // Hessian of phi (xx yy zz xy xz yz)
    v3<R> u1, xw1dot, e, n, xi2, a2, xdot, s2, vw, tdot, n1, ni1a, ni2a, t1, t2, T1a, T2a;
    R gn, ddot, lam, fcn, tn, rho;
    bool fric, stat;
    bool slow;
    R f[12];            // wrench contribution [tau1 f1 tau2 f2]
};

// The contact point in the SDF body frame, formed as the reference does:
// xw1 = R1 xi1 + p1 (CollisionDetection.cpp:136), x = R2^T (xw1 - p2)
// (BodySDFObj.cpp:251).
ATAC_HD void body2_point(const InstDev& in, double px, double py, double pz, double x[3]) {
    const double xi[3] = {px, py, pz};
    double xw[3];
    for (int r = 0; r < 3; ++r) xw[r] = in.R1[3 * r] * xi[0] + in.R1[3 * r + 1] * xi[1] + in.R1[3 * r + 2] * xi[2] + in.p1[r];
    const double q0 = xw[0] - in.p2[0], q1 = xw[1] - in.p2[1], q2 = xw[2] - in.p2[2];
    for (int r = 0; r < 3; ++r) x[r] = in.R2[r] * q0 + in.R2[3 + r] * q1 + in.R2[6 + r] * q2;
}

// Sample phi at the contact point xi1 (FP32 as stored) of instance `in`.
template <typename R>
ATAC_HD void contact_sample(const GridDev& G, const InstDev& in, double px, double py, double pz, ContactFwd<R>& s) {
    double x[3];
    body2_point(in, px, py, pz, x);
    s.xi1 = mk3<R>((R)px, (R)py, (R)pz);
    s.x = mk3<R>((R)x[0], (R)x[1], (R)x[2]);
    int idx[3];
    float f[3];
    double fd[3];
    if (sizeof(R) == sizeof(double) && fast_cell(G, x, idx, f, fd)) {
        // FP64 contact path: value and partials in FP64 from the same corners
        Corners c = load_corners(G, idx[0], idx[1], idx[2]);
        double val, g[3], H[6];
        sdf_cell64(G, c, fd, val, g, H);
        val += in.col_th;
        s.d = (R)val;
        s.g = mk3<R>((R)g[0], (R)g[1], (R)g[2]);
        for (int k = 0; k < 6; ++k) s.H[k] = (R)H[k];
        s.slow = false;
    } else if (fast_cell(G, x, idx, f, fd)) {
        Corners c = load_corners(G, idx[0], idx[1], idx[2]);
        float val, g[3], H[6];
        sdf_cell32(G, c, f, val, g, H);
        // the penetration depth of an active contact is taken in FP64 from the
        // same corners: the FP32 interpolant's ~1e-8 absolute error would be
        // ~1e-5 of a shallow contact's depth
        const double dval = cell_value64(c, fd) + in.col_th;
        s.d = (R)dval;
        s.g = mk3<R>((R)g[0], (R)g[1], (R)g[2]);
        for (int k = 0; k < 6; ++k) s.H[k] = (R)H[k];
        s.slow = false;
    } else {
        double val, g[3], H[6];
        sdf_stencil64(G, x, val, g, H);
        s.d = (R)(val + in.col_th);
        s.g = mk3<R>((R)g[0], (R)g[1], (R)g[2]);
        for (int k = 0; k < 6; ++k) s.H[k] = (R)H[k];
        s.slow = true;
    }
}

// ------------------------------------------------------------------ contact records
// Fast-path record of an active point from its grid coordinates c (detection
// map) and cell: d in FP64 from the FP32 corners (stored rounded to the record's
// precision), gradient and mixed partials in the record's precision. The
// diagonal Hessian of a fast-path sample is exactly 0.
ATAC_HD void record_fast(const GridDev& G, const InstDev& in, const Corners& cr, const double fd[3], Rec32& r) {
    const float f[3] = {(float)fd[0], (float)fd[1], (float)fd[2]};
    float val, g[3], H[6];
    sdf_cell32(G, cr, f, val, g, H);
    r.b[0] = (float)(cell_value64(cr, fd) + in.col_th);
    r.b[1] = g[0]; r.b[2] = g[1]; r.b[3] = g[2];
    r.c[0] = H[3]; r.c[1] = H[4]; r.c[2] = H[5];
    r.c[3] = __builtin_bit_cast(float, -1);
}

ATAC_HD void record_fast(const GridDev& G, const InstDev& in, const Corners& cr, const double fd[3], Rec64& r) {
    double val, g[3], H[6];
    sdf_cell64(G, cr, fd, val, g, H);
    r.d = val + in.col_th;
    for (int k = 0; k < 3; ++k) r.g[k] = g[k];
    for (int k = 0; k < 6; ++k) r.H[k] = H[k];
}

// Slow-path sample (FD stencils operation for operation in FP64) of a record's
// point; fills the record (and, for FP32 records, the diagonal Hessian entry).
ATAC_HD void record_slow_fill(const InstDev& in, double val, const double g[3], const double H[6], Rec32& r,
                              float* diag) {
    r.b[0] = (float)(val + in.col_th);
    r.b[1] = (float)g[0]; r.b[2] = (float)g[1]; r.b[3] = (float)g[2];
    r.c[0] = (float)H[3]; r.c[1] = (float)H[4]; r.c[2] = (float)H[5];
    diag[0] = (float)H[0]; diag[1] = (float)H[1]; diag[2] = (float)H[2]; diag[3] = 0.0f;
}

ATAC_HD void record_slow_fill(const InstDev& in, double val, const double g[3], const double H[6], Rec64& r, float*) {
    r.d = val + in.col_th;
    for (int k = 0; k < 3; ++k) r.g[k] = g[k];
    for (int k = 0; k < 6; ++k) r.H[k] = H[k];
}

template <typename Rec>
ATAC_HD void record_slow(const GridDev& G, const InstDev& in, Rec& r, float* diag) {
    double x[3], val, g[3], H[6];
    body2_point(in, (double)r.a[0], (double)r.a[1], (double)r.a[2], x);
    sdf_stencil64(G, x, val, g, H);
    record_slow_fill(in, val, g, H, r, diag);
}

// Forward state of a record: the contact point, its SDF-frame position (formed
// from xi1 as the reference does) and the sampled d, grad phi, Hessian.
ATAC_HD void record_decode(const Rec32& r, const float* slow_diag, const InstDev& in, ContactFwd<float>& s) {
    double x[3];
    body2_point(in, (double)r.a[0], (double)r.a[1], (double)r.a[2], x);
    s.xi1 = mk3<float>(r.a[0], r.a[1], r.a[2]);
    s.x = mk3<float>((float)x[0], (float)x[1], (float)x[2]);
    s.d = r.b[0];
    s.g = mk3<float>(r.b[1], r.b[2], r.b[3]);
    s.H[3] = r.c[0]; s.H[4] = r.c[1]; s.H[5] = r.c[2];
    const int slow = __builtin_bit_cast(int, r.c[3]);
    s.slow = slow >= 0;
    if (s.slow) {
        s.H[0] = slow_diag[4 * (int64_t)slow];
        s.H[1] = slow_diag[4 * (int64_t)slow + 1];
        s.H[2] = slow_diag[4 * (int64_t)slow + 2];
    } else {
        s.H[0] = s.H[1] = s.H[2] = 0.0f;
    }
}

ATAC_HD void record_decode(const Rec64& r, const float*, const InstDev& in, ContactFwd<double>& s) {
    double x[3];
    body2_point(in, (double)r.a[0], (double)r.a[1], (double)r.a[2], x);
    s.xi1 = mk3<double>(r.a[0], r.a[1], r.a[2]);
    s.x = mk3<double>(x[0], x[1], x[2]);
    s.d = r.d;
    s.g = mk3<double>(r.g[0], r.g[1], r.g[2]);
    for (int k = 0; k < 6; ++k) s.H[k] = r.H[k];
    s.slow = false;
}

// Values of BodySDFObj::collision (BodySDFObj.cpp:251-273) given s.x, s.d, s.g
// and the world velocity s.xw1dot of the contact point.
template <typename R>
ATAC_HD void collision_forward(const InstR<R>& I, ContactFwd<R>& s) {
    s.gn = norm(s.g);
    s.e = s.g * qdiv((R)1, s.gn);
    s.n = mv(I.R2, s.e);
    s.xi2 = s.x - s.d * s.e;
    s.a2 = mtv(I.R2, s.xw1dot);
    s.xdot = cross(s.x, I.w2) + s.a2 - I.v2;       // skew(w2)^T x + R2^T xw_dot - v2
    s.ddot = dot(s.g, s.xdot);
    s.s2 = cross(I.w2, s.xi2) + I.v2;
    s.vw = s.xw1dot - mv(I.R2, s.s2);
    s.tdot = s.vw - s.n * dot(s.n, s.vw);          // (I - n n^T) vw
}

// Wrench of one contact (ForceGeneralSDFContact.cpp:157-218 with
// BodySDFObj::collision, :239-274). Writes s.f and, when p6 != nullptr, the
// reference's contact-scale column contribution.
template <typename R>
ATAC_HD void contact_forward(const InstR<R>& I, ContactFwd<R>& s, R* p6) {
    const R eps = (R)1e-8;  // constants::eps, Utils.h:9
    s.u1 = cross(I.w1, s.xi1) + I.v1;              // skew(xi1)^T w1 + v1
    s.xw1dot = mv(I.R1, s.u1);
    collision_forward(I, s);
    s.n1 = mtv(I.R1, s.n);
    s.ni1a = cross(s.xi1, s.n1);                   // GTRT1 n = [xi1 x n1; n1]
    s.ni2a = cross(s.xi2, s.e);                    // GTRT2 n = [xi2 x e; e]
    s.lam = -I.kn * s.d + I.c * s.ddot * s.d;
    s.f[0] = s.lam * s.ni1a.x; s.f[1] = s.lam * s.ni1a.y; s.f[2] = s.lam * s.ni1a.z;
    s.f[3] = s.lam * s.n1.x;   s.f[4] = s.lam * s.n1.y;   s.f[5] = s.lam * s.n1.z;
    s.f[6] = -s.lam * s.ni2a.x; s.f[7] = -s.lam * s.ni2a.y; s.f[8] = -s.lam * s.ni2a.z;
    s.f[9] = -s.lam * s.e.x;    s.f[10] = -s.lam * s.e.y;   s.f[11] = -s.lam * s.e.z;
    if (p6) {
        const R k = qdiv(I.kn * s.d, I.cs);
        p6[0] = -k * s.ni1a.x; p6[1] = -k * s.ni1a.y; p6[2] = -k * s.ni1a.z;
        p6[3] = -k * s.n1.x;   p6[4] = -k * s.n1.y;   p6[5] = -k * s.n1.z;
        p6[6] = k * s.ni2a.x;  p6[7] = k * s.ni2a.y;  p6[8] = k * s.ni2a.z;
        p6[9] = k * s.e.x;     p6[10] = k * s.e.y;    p6[11] = k * s.e.z;
    }
    s.fric = I.mu > eps;
    s.stat = true;
    s.fcn = s.tn = s.rho = (R)0;
    if (s.fric) {
        R ss = 0;
        for (int k = 0; k < 6; ++k) ss += s.f[k] * s.f[k];
        s.fcn = sqrt(ss);                          // fc1.norm()
        s.tn = norm(s.tdot);
        s.t1 = mtv(I.R1, s.tdot);
        s.t2 = mtv(I.R2, s.tdot);
        s.T1a = cross(s.xi1, s.t1);
        s.T2a = cross(s.xi2, s.t2);
        s.stat = I.mu * s.fcn >= I.kt * s.tn - eps;
        R a1, a2;                                  // friction = -a1 [T1a; t1], +a2 [T2a; t2]
        if (s.stat) {
            a1 = I.kt;
            a2 = I.kt;
        } else {
            s.rho = qdiv(s.fcn, s.tn);
            a1 = I.mu * s.rho;
            a2 = I.mu * s.rho;
        }
        s.f[0] -= a1 * s.T1a.x; s.f[1] -= a1 * s.T1a.y; s.f[2] -= a1 * s.T1a.z;
        s.f[3] -= a1 * s.t1.x;  s.f[4] -= a1 * s.t1.y;  s.f[5] -= a1 * s.t1.z;
        s.f[6] += a2 * s.T2a.x; s.f[7] += a2 * s.T2a.y; s.f[8] += a2 * s.T2a.z;
        s.f[9] += a2 * s.t2.x;  s.f[10] += a2 * s.t2.y; s.f[11] += a2 * s.t2.z;
        if (p6) {
            const R b1 = qdiv(a1, I.cs), b2 = qdiv(a2, I.cs);
            p6[0] -= b1 * s.T1a.x; p6[1] -= b1 * s.T1a.y; p6[2] -= b1 * s.T1a.z;
            p6[3] -= b1 * s.t1.x;  p6[4] -= b1 * s.t1.y;  p6[5] -= b1 * s.t1.z;
            p6[6] += b2 * s.T2a.x; p6[7] += b2 * s.T2a.y; p6[8] += b2 * s.T2a.z;
            p6[9] += b2 * s.t2.x;  p6[10] += b2 * s.t2.y; p6[11] += b2 * s.t2.z;
        }
    }
}

// Perturbation direction: right-multiplied pose twists E <- E exp([om; v]) of
// both bodies (the convention of every reference FD test, e.g.
// ForceGeneralSDFContact.cpp:1066), body-frame twist increments, and kn.
template <typename R>
struct Tangent {
    v3<R> om1, v1, om2, v2, w1p, v1p, w2p, v2p;
    R kn;
};

// Column j of [Km | Dm] (j < 12: q1 then q2; 12 <= j < 24: phi1 then phi2).
template <typename R>
ATAC_HD Tangent<R> unit_tangent(int j) {
    Tangent<R> t;
    t.om1 = t.v1 = t.om2 = t.v2 = t.w1p = t.v1p = t.w2p = t.v2p = zero3<R>();
    t.kn = 0;
    v3<R>* slots[8] = {&t.om1, &t.v1, &t.om2, &t.v2, &t.w1p, &t.v1p, &t.w2p, &t.v2p};
    v3<R>& s = *slots[j / 3];
    int c = j % 3;
    if (c == 0) s.x = 1; else if (c == 1) s.y = 1; else s.z = 1;
    return t;
}

// Tangents of the collision values (the 14 derivative blocks of
// BodySDFObj::collision, BodySDFObj.cpp:314-376, applied to one direction).
template <typename R>
struct CollTangent {
    R dd, dddot;
    v3<R> de, dn, dxi2, dtdot;
};

template <typename R>
ATAC_HD void collision_jvp(const InstR<R>& I, const ContactFwd<R>& s, v3<R> dxw, v3<R> dxwdot, v3<R> om2, v3<R> v2,
                           v3<R> w2p, v3<R> v2p, CollTangent<R>& o) {
    v3<R> dx = mtv(I.R2, dxw) + cross(s.x, om2) - v2;          // dx_dxw, dx_dw2 = skew(x), dx_dv2 = -I
    o.dd = dot(s.g, dx);
    v3<R> dg = symv(s.H, dx);
    o.de = (dg - s.e * dot(s.e, dg)) * qdiv((R)1, s.gn);           // de_ddd_dx * H * dx
    o.dn = mv(I.R2, o.de + cross(om2, s.e));
    o.dxi2 = dx - s.e * o.dd - s.d * o.de;
    v3<R> dxdot = cross(dx, I.w2) + cross(s.x, w2p) + cross(s.a2, om2) + mtv(I.R2, dxwdot) - v2p;
    o.dddot = dot(dg, s.xdot) + dot(s.g, dxdot);
    v3<R> dxw2dot = mv(I.R2, cross(om2, s.s2) + cross(I.w2, o.dxi2) + cross(w2p, s.xi2) + v2p);
    v3<R> dvw = dxwdot - dxw2dot;
    o.dtdot = dvw - s.n * dot(s.n, dvw) - o.dn * dot(s.n, s.vw) - s.n * dot(s.vw, o.dn);
}

// Directional derivative of the contact wrench (the tangent form of the
// reference's Km/Dm assembly, ForceGeneralSDFContact.cpp:240-446).
template <typename R>
ATAC_HD void contact_jvp(const InstR<R>& I, const ContactFwd<R>& s, const Tangent<R>& t, R df[12]) {
    v3<R> dxw1 = mv(I.R1, cross(t.om1, s.xi1) + t.v1);                          // dxw1_dq1
    v3<R> dxw1dot = mv(I.R1, cross(t.om1, s.u1) + cross(t.w1p, s.xi1) + t.v1p);  // dxw1dot_dq1, _dphi1
    CollTangent<R> c;
    collision_jvp(I, s, dxw1, dxw1dot, t.om2, t.v2, t.w2p, t.v2p, c);
    const R dd = c.dd, dddot = c.dddot;
    const v3<R> de = c.de, dn = c.dn, dxi2 = c.dxi2, dtdot = c.dtdot;
    v3<R> dn1 = mtv(I.R1, dn) + cross(s.n1, t.om1);
    R dlam = -t.kn * s.d - I.kn * dd + I.c * (dddot * s.d + s.ddot * dd);
    v3<R> a1 = dlam * s.ni1a + s.lam * cross(s.xi1, dn1);
    v3<R> l1 = dlam * s.n1 + s.lam * dn1;
    v3<R> a2 = -(dlam * s.ni2a + s.lam * (cross(dxi2, s.e) + cross(s.xi2, de)));
    v3<R> l2 = -(dlam * s.e + s.lam * de);
    if (s.fric) {
        v3<R> dt1 = mtv(I.R1, dtdot) + cross(s.t1, t.om1);
        v3<R> dt2 = mtv(I.R2, dtdot) + cross(s.t2, t.om2);
        v3<R> dT1a = cross(s.xi1, dt1);
        v3<R> dT2a = cross(dxi2, s.t2) + cross(s.xi2, dt2);
        if (s.stat) {
            a1 = a1 - I.kt * dT1a; l1 = l1 - I.kt * dt1;
            a2 = a2 + I.kt * dT2a; l2 = l2 + I.kt * dt2;
        } else {
            // fc1 = lam [ni1a; n1]; its tangent is [a1; l1] before friction is added
            R fc_dot = s.lam * (dot(s.ni1a, a1) + dot(s.n1, l1));
            R dfcn = qdiv(fc_dot, s.fcn);
            R dtn = qdiv(dot(s.tdot, dtdot), s.tn);
            R drho = qdiv(dfcn, s.tn) - qdiv(s.fcn * dtn, s.tn * s.tn);
            a1 = a1 - I.mu * (drho * s.T1a + s.rho * dT1a);
            l1 = l1 - I.mu * (drho * s.t1 + s.rho * dt1);
            a2 = a2 + I.mu * (drho * s.T2a + s.rho * dT2a);
            l2 = l2 + I.mu * (drho * s.t2 + s.rho * dt2);
        }
    }
    df[0] = a1.x; df[1] = a1.y; df[2] = a1.z; df[3] = l1.x; df[4] = l1.y; df[5] = l1.z;
    df[6] = a2.x; df[7] = a2.y; df[8] = a2.z; df[9] = l2.x; df[10] = l2.y; df[11] = l2.z;
}

// Adjoint of contact_jvp: given gb = d(loss)/d(wrench), returns
// out[0:6] = gb^T Km[:, q1], out[6:12] = gb^T Km[:, q2], out[12:18] = gb^T Dm[:, phi1],
// out[18:24] = gb^T Dm[:, phi2], out[24] = gb^T d f / d kn, out[25] = gb^T d f / d kt
// (kn, kt effective). Hand-written reverse sweep of the tangent map above.
// ACC: add into out instead of assigning.
template <typename R, bool ACC = false>
ATAC_HD void contact_vjp(const InstR<R>& I, const ContactFwd<R>& s, const R gb[12], R out[26]) {
    const v3<R> f1a = mk3<R>(gb[0], gb[1], gb[2]), f1l = mk3<R>(gb[3], gb[4], gb[5]);
    const v3<R> f2a = mk3<R>(gb[6], gb[7], gb[8]), f2l = mk3<R>(gb[9], gb[10], gb[11]);
    v3<R> om1b = zero3<R>(), v1b = zero3<R>(), om2b = zero3<R>(), v2b = zero3<R>();
    v3<R> w1pb = zero3<R>(), v1pb = zero3<R>(), w2pb = zero3<R>(), v2pb = zero3<R>();
    v3<R> tdotb = zero3<R>(), xi2b = zero3<R>(), nb = zero3<R>(), eb = zero3<R>();
    R knb = 0, ktb = 0;
    v3<R> fc1ab = f1a, fc1lb = f1l;   // adjoint of the normal force on body 1
    if (s.fric) {
        v3<R> T1ab, t1b, T2ab, t2b;
        if (s.stat) {
            T1ab = -I.kt * f1a; t1b = -I.kt * f1l; T2ab = I.kt * f2a; t2b = I.kt * f2l;
            ktb += -(dot(f1a, s.T1a) + dot(f1l, s.t1)) + dot(f2a, s.T2a) + dot(f2l, s.t2);
        } else {
            R rhob = -I.mu * (dot(f1a, s.T1a) + dot(f1l, s.t1)) + I.mu * (dot(f2a, s.T2a) + dot(f2l, s.t2));
            T1ab = -I.mu * s.rho * f1a; t1b = -I.mu * s.rho * f1l;
            T2ab = I.mu * s.rho * f2a;  t2b = I.mu * s.rho * f2l;
            R fcnb = qdiv(rhob, s.tn);
            R tnb = qdiv(-rhob * s.fcn, s.tn * s.tn);
            R k = qdiv(fcnb * s.lam, s.fcn);       // fc1 = lam [ni1a; n1]
            fc1ab += k * s.ni1a;
            fc1lb += k * s.n1;
            tdotb += qdiv(tnb, s.tn) * s.tdot;
        }
        t1b += cross(T1ab, s.xi1);
        tdotb += mv(I.R1, t1b);
        om1b += cross(t1b, s.t1);
        xi2b += cross(s.t2, T2ab);
        t2b += cross(T2ab, s.xi2);
        tdotb += mv(I.R2, t2b);
        om2b += cross(t2b, s.t2);
    }
    R lamb = dot(fc1ab, s.ni1a) + dot(fc1lb, s.n1) - dot(f2a, s.ni2a) - dot(f2l, s.e);
    v3<R> n1b = s.lam * (cross(fc1ab, s.xi1) + fc1lb);
    v3<R> ni2ab = -s.lam * f2a, ni2lb = -s.lam * f2l;
    xi2b += cross(s.e, ni2ab);
    eb += cross(ni2ab, s.xi2) + ni2lb;
    nb += mv(I.R1, n1b);
    om1b += cross(n1b, s.n1);
    knb += -lamb * s.d;
    R db = lamb * (-I.kn + I.c * s.ddot);
    R ddotb = lamb * I.c * s.d;
    v3<R> vwb = tdotb - s.n * dot(s.n, tdotb);
    nb += -dot(s.n, s.vw) * tdotb - dot(s.n, tdotb) * s.vw;
    v3<R> xw1dotb = vwb;
    v3<R> z = mtv(I.R2, -vwb);                     // R2^T (adjoint of xw2dot)
    om2b += cross(s.s2, z);
    xi2b += cross(z, I.w2);
    w2pb += cross(s.xi2, z);
    v2pb += z;
    v3<R> gtb = ddotb * s.xdot;
    v3<R> xdotb = ddotb * s.g;
    v3<R> xb = cross(I.w2, xdotb);
    w2pb += cross(xdotb, s.x);
    om2b += cross(xdotb, s.a2);
    xw1dotb += mv(I.R2, xdotb);
    v2pb += -xdotb;
    xb += xi2b;
    db += -dot(s.e, xi2b);
    eb += -s.d * xi2b;
    v3<R> m = mtv(I.R2, nb);
    eb += m;
    om2b += cross(s.e, m);
    gtb += (eb - s.e * dot(s.e, eb)) * qdiv((R)1, s.gn);
    xb += symv(s.H, gtb);
    xb += db * s.g;
    v3<R> xw1b = mv(I.R2, xb);
    om2b += cross(xb, s.x);
    v2b += -xb;
    v3<R> y = mtv(I.R1, xw1dotb);
    om1b += cross(s.u1, y);
    w1pb += cross(s.xi1, y);
    v1pb += y;
    v3<R> y2 = mtv(I.R1, xw1b);
    om1b += cross(s.xi1, y2);
    v1b += y2;
    const v3<R> o[8] = {om1b, v1b, om2b, v2b, w1pb, v1pb, w2pb, v2pb};
    if (ACC) {
        for (int k = 0; k < 8; ++k) { out[3 * k] += o[k].x; out[3 * k + 1] += o[k].y; out[3 * k + 2] += o[k].z; }
        out[24] += knb;
        out[25] += ktb;
    } else {
        for (int k = 0; k < 8; ++k) { out[3 * k] = o[k].x; out[3 * k + 1] = o[k].y; out[3 * k + 2] = o[k].z; }
        out[24] = knb;
        out[25] = ktb;
    }
}

}  // namespace atac
