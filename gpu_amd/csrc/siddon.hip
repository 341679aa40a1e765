// siddon.hip - ray-driven Siddon projector for gfx950.
//
// One ray per lane.  A 256-thread workgroup owns a 16x16 tile of detector
// pixels and each wave64 an 8x8 sub-tile, so the 64 rays of a wave stay
// within a few voxels of each other along the whole traversal and their
// gathers share cache lines.
//
// The traversal restates forward_ray_driven_3d_kernel_correction
// (V63/kernel_IterativeRecon_universal.cu:14-551; multi-source variant
// TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:1-545) with the same FP
// operation order, so with contraction off the FP32 path reproduces the
// reference arithmetic.  The reference's seven tie branches (:307-549) are
// one branch-free step here: every axis whose next crossing equals the
// minimum steps, which is the same set of cases without wave divergence.
#include "ct_internal.hpp"
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <mutex>
#include <string>

namespace ct {

template <typename T> __device__ __forceinline__ T dfloor(T x);
template <> __device__ __forceinline__ float dfloor<float>(float x) { return floorf(x); }
template <> __device__ __forceinline__ double dfloor<double>(double x) { return floor(x); }
template <typename T> __device__ __forceinline__ T dmin(T a, T b);
template <> __device__ __forceinline__ float dmin<float>(float a, float b) { return fminf(a, b); }
template <> __device__ __forceinline__ double dmin<double>(double a, double b) { return fmin(a, b); }
template <typename T> __device__ __forceinline__ T dmax(T a, T b);
template <> __device__ __forceinline__ float dmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double dmax<double>(double a, double b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T dabs(T a);
template <> __device__ __forceinline__ float dabs<float>(float a) { return fabsf(a); }
template <> __device__ __forceinline__ double dabs<double>(double a) { return fabs(a); }
template <typename T> __device__ __forceinline__ T dsqrt(T a);
template <> __device__ __forceinline__ float dsqrt<float>(float a) { return sqrtf(a); }
template <> __device__ __forceinline__ double dsqrt<double>(double a) { return sqrt(a); }
template <typename T> __device__ __forceinline__ T dinf();
template <> __device__ __forceinline__ float dinf<float>() { return __builtin_inff(); }
template <> __device__ __forceinline__ double dinf<double>() { return __builtin_inf(); }
template <typename T> __device__ __forceinline__ T loop_eps();
template <> __device__ __forceinline__ float loop_eps<float>() { return 1e-16f; }
template <> __device__ __forceinline__ double loop_eps<double>() { return 1e-16; }

// Traverses the ray detector pixel (col,row) -> source `src`; visit(index, len)
// per segment.  Returns the summed length.
template <typename T, typename F>
__device__ __forceinline__ T siddon_traverse(const DGeo<T> &d, T s, T c, int col, int row, int src, F &&visit) {
    const int M = d.M, N = d.N, Z = d.Z;
    const T u = d.ymin + (T)col * d.du;
    const T x1x = d.dod * c - u * s;
    const T x1y = d.dod * s + u * c;
    const T x1z = d.zmin + (T)row * d.dv;
    T x2x, x2y, x2z;
    ray_source(d, s, c, u, x1z, src, x2x, x2y, x2z);
    const T ixd = (T)1 / (x2x - x1x);
    const T iyd = (T)1 / (x2y - x1y);
    const T izd = (T)1 / (x2z - x1z);

    T a0, a1;
    a0 = (d.bx - x1x) * ixd;
    a1 = (d.bx + d.vx * (T)M - x1x) * ixd;
    const T axmin = dmin(a0, a1), axmax = dmax(a0, a1);
    a0 = (d.by - x1y) * iyd;
    a1 = (d.by + d.vy * (T)N - x1y) * iyd;
    const T aymin = dmin(a0, a1), aymax = dmax(a0, a1);
    a0 = (d.bz - x1z) * izd;
    a1 = (d.bz + d.vz * (T)Z - x1z) * izd;
    const T azmin = dmin(a0, a1), azmax = dmax(a0, a1);
    const T amin = dmax(dmax(axmin, aymin), dmax(aymin, azmin));
    const T amax = dmin(dmin(axmax, aymax), dmin(aymax, azmax));

    T len = 0;
    if (amax <= amin) return len;

    int imin, imax, jmin, jmax, kmin, kmax;
    const T ex = x2x - x1x, ey = x2y - x1y, ez = x2z - x1z;
    if (x1x < x2x) {
        imin = (amin == axmin) ? 1 : (int)dfloor((amin * ex + x1x - d.bx) * d.ivx) + 1;
        imax = (amax == axmax) ? M : (int)dfloor((amax * ex + x1x - d.bx) * d.ivx);
    } else {
        imax = (amin == axmin) ? M - 1 : (int)dfloor((amin * ex + x1x - d.bx) * d.ivx);
        imin = (amax == axmax) ? 0 : (int)dfloor((amax * ex + x1x - d.bx) * d.ivx) + 1;
    }
    if (x1y < x2y) {
        jmin = (amin == aymin) ? 1 : (int)dfloor((amin * ey + x1y - d.by) * d.ivy) + 1;
        jmax = (amax == aymax) ? N : (int)dfloor((amax * ey + x1y - d.by) * d.ivy);
    } else {
        jmax = (amin == aymin) ? N - 1 : (int)dfloor((amin * ey + x1y - d.by) * d.ivy);
        jmin = (amax == aymax) ? 0 : (int)dfloor((amax * ey + x1y - d.by) * d.ivy) + 1;
    }
    const bool zpar = dabs(x1z - x2z) < d.vz * (T)1e-6;
    if (zpar) {
        kmin = (int)dfloor((x1z - d.bz) * d.ivz) + 1;
        kmax = (int)dfloor((x1z - d.bz) * d.ivz);
    } else if (x1z < x2z) {
        kmin = (amin == azmin) ? 1 : (int)dfloor((amin * ez + x1z - d.bz) * d.ivz) + 1;
        kmax = (amax == azmax) ? Z : (int)dfloor((amax * ez + x1z - d.bz) * d.ivz);
    } else {
        kmax = (amin == azmin) ? Z - 1 : (int)dfloor((amin * ez + x1z - d.bz) * d.ivz);
        kmin = (amax == azmax) ? 0 : (int)dfloor((amax * ez + x1z - d.bz) * d.ivz) + 1;
    }

    // next-plane index n* and the crossing a* of that plane (:218-266)
    const int sgx = (x1x < x2x) ? 1 : ((x1x > x2x) ? -1 : 0);
    const int sgy = (x1y < x2y) ? 1 : ((x1y > x2y) ? -1 : 0);
    const int sgz = (x1z < x2z) ? 1 : ((x1z > x2z) ? -1 : 0);
    int nx, ny, nz;
    T ax, ay, az;
    if (dabs(x1x - x2x) < d.vx * (T)1e-6) { ax = dinf<T>(); nx = imin; }
    else { nx = (sgx > 0) ? imin : imax; ax = (d.vx * (T)nx + d.bx - x1x) * ixd; }
    if (dabs(x1y - x2y) < d.vy * (T)1e-6) { ay = dinf<T>(); ny = jmin; }
    else { ny = (sgy > 0) ? jmin : jmax; ay = (d.vy * (T)ny + d.by - x1y) * iyd; }
    if (zpar) { az = dinf<T>(); nz = kmin; }
    else { nz = (sgz > 0) ? kmin : kmax; az = (d.vz * (T)nz + d.bz - x1z) * izd; }
    int vi = (x1x < x2x) ? imin - 1 : imax;   // (:271-284)
    int vj = (x1y < x2y) ? jmin - 1 : jmax;
    int vk = (x1z < x2z) ? kmin - 1 : kmax;

    const T dist = dsqrt(ex * ex + ey * ey + ez * ez);
    T ac = amin;
    // the traversal visits at most M+N+Z+1 voxels; the cap only guarantees
    // termination for non-finite inputs
    int guard = M + N + Z + 4;
    while (amax - ac > loop_eps<T>() && guard-- > 0) {
        if ((unsigned)vi >= (unsigned)M || (unsigned)vj >= (unsigned)N || (unsigned)vk >= (unsigned)Z) break;
        const T m = dmin(dmin(ax, ay), az);
        const T seg = dist * (m - ac);
        len += seg;
        visit(vk * (M * N) + vj * M + vi, seg);
        ac = m;
        if (ax == m) { vi += sgx; nx += sgx; ax = (d.vx * (T)nx + d.bx - x1x) * ixd; }
        if (ay == m) { vj += sgy; ny += sgy; ay = (d.vy * (T)ny + d.by - x1y) * iyd; }
        if (az == m) { vk += sgz; nz += sgz; az = (d.vz * (T)nz + d.bz - x1z) * izd; }
    }
    return len;
}

// Ray setup shared by the batched forward sum: the reference's entry point,
// plane indices and first crossings (:20-292), kept operation for operation.
template <typename T> struct RayState {
    T amax, ac, ax, ay, az, dist, x1x, x1y, x1z, x2x, x2y, x2z, ixd, iyd, izd;
    int nx, ny, nz, vi, vj, vk, sgx, sgy, sgz;
    bool hit;
};

template <typename T>
__device__ __forceinline__ RayState<T> ray_setup(const DGeo<T> &d, T s, T c, int col, int row, int src) {
    RayState<T> r;
    const int M = d.M, N = d.N, Z = d.Z;
    const T u = d.ymin + (T)col * d.du;
    const T x1x = d.dod * c - u * s;
    const T x1y = d.dod * s + u * c;
    const T x1z = d.zmin + (T)row * d.dv;
    T x2x, x2y, x2z;
    ray_source(d, s, c, u, x1z, src, x2x, x2y, x2z);
    r.x2x = x2x; r.x2y = x2y; r.x2z = x2z;
    const T ixd = (T)1 / (x2x - x1x);
    const T iyd = (T)1 / (x2y - x1y);
    const T izd = (T)1 / (x2z - x1z);
    r.x1x = x1x; r.x1y = x1y; r.x1z = x1z; r.ixd = ixd; r.iyd = iyd; r.izd = izd;
    T a0, a1;
    a0 = (d.bx - x1x) * ixd;
    a1 = (d.bx + d.vx * (T)M - x1x) * ixd;
    const T axmin = dmin(a0, a1), axmax = dmax(a0, a1);
    a0 = (d.by - x1y) * iyd;
    a1 = (d.by + d.vy * (T)N - x1y) * iyd;
    const T aymin = dmin(a0, a1), aymax = dmax(a0, a1);
    a0 = (d.bz - x1z) * izd;
    a1 = (d.bz + d.vz * (T)Z - x1z) * izd;
    const T azmin = dmin(a0, a1), azmax = dmax(a0, a1);
    const T amin = dmax(dmax(axmin, aymin), dmax(aymin, azmin));
    const T amax = dmin(dmin(axmax, aymax), dmin(aymax, azmax));
    r.amax = amax; r.ac = amin;
    r.hit = !(amax <= amin);
    if (!r.hit) return r;
    int imin, imax, jmin, jmax, kmin, kmax;
    const T ex = x2x - x1x, ey = x2y - x1y, ez = x2z - x1z;
    if (x1x < x2x) {
        imin = (amin == axmin) ? 1 : (int)dfloor((amin * ex + x1x - d.bx) * d.ivx) + 1;
        imax = (amax == axmax) ? M : (int)dfloor((amax * ex + x1x - d.bx) * d.ivx);
    } else {
        imax = (amin == axmin) ? M - 1 : (int)dfloor((amin * ex + x1x - d.bx) * d.ivx);
        imin = (amax == axmax) ? 0 : (int)dfloor((amax * ex + x1x - d.bx) * d.ivx) + 1;
    }
    if (x1y < x2y) {
        jmin = (amin == aymin) ? 1 : (int)dfloor((amin * ey + x1y - d.by) * d.ivy) + 1;
        jmax = (amax == aymax) ? N : (int)dfloor((amax * ey + x1y - d.by) * d.ivy);
    } else {
        jmax = (amin == aymin) ? N - 1 : (int)dfloor((amin * ey + x1y - d.by) * d.ivy);
        jmin = (amax == aymax) ? 0 : (int)dfloor((amax * ey + x1y - d.by) * d.ivy) + 1;
    }
    const bool zpar = dabs(x1z - x2z) < d.vz * (T)1e-6;
    if (zpar) {
        kmin = (int)dfloor((x1z - d.bz) * d.ivz) + 1;
        kmax = (int)dfloor((x1z - d.bz) * d.ivz);
    } else if (x1z < x2z) {
        kmin = (amin == azmin) ? 1 : (int)dfloor((amin * ez + x1z - d.bz) * d.ivz) + 1;
        kmax = (amax == azmax) ? Z : (int)dfloor((amax * ez + x1z - d.bz) * d.ivz);
    } else {
        kmax = (amin == azmin) ? Z - 1 : (int)dfloor((amin * ez + x1z - d.bz) * d.ivz);
        kmin = (amax == azmax) ? 0 : (int)dfloor((amax * ez + x1z - d.bz) * d.ivz) + 1;
    }
    r.sgx = (x1x < x2x) ? 1 : ((x1x > x2x) ? -1 : 0);
    r.sgy = (x1y < x2y) ? 1 : ((x1y > x2y) ? -1 : 0);
    r.sgz = (x1z < x2z) ? 1 : ((x1z > x2z) ? -1 : 0);
    if (dabs(x1x - x2x) < d.vx * (T)1e-6) { r.ax = dinf<T>(); r.nx = imin; }
    else { r.nx = (r.sgx > 0) ? imin : imax; r.ax = (d.vx * (T)r.nx + d.bx - x1x) * ixd; }
    if (dabs(x1y - x2y) < d.vy * (T)1e-6) { r.ay = dinf<T>(); r.ny = jmin; }
    else { r.ny = (r.sgy > 0) ? jmin : jmax; r.ay = (d.vy * (T)r.ny + d.by - x1y) * iyd; }
    if (zpar) { r.az = dinf<T>(); r.nz = kmin; }
    else { r.nz = (r.sgz > 0) ? kmin : kmax; r.az = (d.vz * (T)r.nz + d.bz - x1z) * izd; }
    r.vi = (x1x < x2x) ? imin - 1 : imax;
    r.vj = (x1y < x2y) ? jmin - 1 : jmax;
    r.vk = (x1z < x2z) ? kmin - 1 : kmax;
    r.dist = dsqrt(ex * ex + ey * ey + ez * ez);
    return r;
}

// Ray sum with B steps generated per batch and their B gathers issued
// together: the steps, segment lengths and the order of the two sums are
// those of siddon_traverse, so with COMP = false the result is identical; only
// the loads are in flight together instead of one at a time.  COMP = true
// (fast mode) adds Kahan compensation to the ray sum and length, which keeps
// long FP32 rays at FP64-like accuracy.
template <typename T, typename SV, int B, bool COMP>
__device__ __forceinline__ T siddon_sum(const DGeo<T> &d, T s, T c, int col, int row, int src,
                                        const SV *__restrict__ f, T &len_out) {
    RayState<T> r = ray_setup(d, s, c, col, row, src);
    T sum = 0, len = 0, csum = 0, clen = 0;
    if (r.hit) {
        const unsigned M = d.M, N = d.N, Z = d.Z, MN = M * N;
        int guard = d.M + d.N + d.Z + 4;
        bool alive = true;
        while (alive) {
            unsigned idx[B];
            T seg[B];
            bool act[B];
            // step generation, branch-free: every state update is a select
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const bool ok = alive && (r.amax - r.ac > loop_eps<T>()) && (guard > 0) &&
                                (unsigned)r.vi < M && (unsigned)r.vj < N && (unsigned)r.vk < Z;
                const T m = dmin(dmin(r.ax, r.ay), r.az);
                act[u] = ok;
                seg[u] = r.dist * (m - r.ac);
                idx[u] = ok ? (unsigned)r.vk * MN + (unsigned)r.vj * M + (unsigned)r.vi : 0u;
                const bool sx = ok && (r.ax == m), sy = ok && (r.ay == m), sz = ok && (r.az == m);
                const int nx1 = r.nx + r.sgx, ny1 = r.ny + r.sgy, nz1 = r.nz + r.sgz;
                const T ax1 = (d.vx * (T)nx1 + d.bx - r.x1x) * r.ixd;
                const T ay1 = (d.vy * (T)ny1 + d.by - r.x1y) * r.iyd;
                const T az1 = (d.vz * (T)nz1 + d.bz - r.x1z) * r.izd;
                r.vi = sx ? r.vi + r.sgx : r.vi;
                r.nx = sx ? nx1 : r.nx;
                r.ax = sx ? ax1 : r.ax;
                r.vj = sy ? r.vj + r.sgy : r.vj;
                r.ny = sy ? ny1 : r.ny;
                r.ay = sy ? ay1 : r.ay;
                r.vk = sz ? r.vk + r.sgz : r.vk;
                r.nz = sz ? nz1 : r.nz;
                r.az = sz ? az1 : r.az;
                r.ac = ok ? m : r.ac;
                guard -= ok ? 1 : 0;
                alive = ok;
            }
            T fv[B];
#pragma unroll
            for (int u = 0; u < B; ++u) fv[u] = (T)ldf(f, idx[u]);
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if constexpr (COMP) {
                    const T yl = seg[u] - clen;
                    const T tl = len + yl;
                    const T cl = (tl - len) - yl;
                    const T ys = seg[u] * fv[u] - csum;
                    const T ts = sum + ys;
                    const T cs = (ts - sum) - ys;
                    len = act[u] ? tl : len;
                    clen = act[u] ? cl : clen;
                    sum = act[u] ? ts : sum;
                    csum = act[u] ? cs : csum;
                } else {
                    len = act[u] ? len + seg[u] : len;
                    sum = act[u] ? sum + seg[u] * fv[u] : sum;
                }
            }
        }
    }
    len_out = len;
    return sum;
}

// Fast-mode FP32 traversal (CT_MATH_FAST).  Same walk as siddon_traverse
// with the instruction count of one step cut from ~59 to ~30 VALU:
//  * the crossing of plane n is fma(n, vx*ixd, (bx - x1x)*ixd) (one rounding
//    per term instead of four), with n kept as a float counter;
//  * one linear voxel offset (in units of ES bytes) with per-axis strides
//    replaces the three voxel coordinates; the first voxel is bounds-checked
//    once (a ray entering outside the grid is empty, as in the reference),
//    and the walk ends at alpha_max, clipped to the grid's exit planes as the
//    walk itself computes them, instead of bounds-checking every step (steps
//    past the end have length 0; the forward still reads them in bounds or
//    gets 0 from the range-checked load, the adjoint skips them);
//  * min/min3 are issued directly (no per-use NaN canonicalisation);
//  * steps carry alpha lengths; the forward multiplies its sum by dist once;
//  * the ray length is the closed form dist*(alpha_max - alpha_min), the
//    telescoped sum of the segments without their rounding.
// The forward (siddon_walk_fast) and the matched adjoint (k_adjoint) both
// step through this walker, so their segments are the same numbers.
//
// ZQ: the volume is in the z-quad layout of the SART sweep,
// [z/4][y][x][z%4] (four z-neighbours share 16 bytes, so the 64 rays of a
// wave - a few voxels apart in the lateral and z directions - hit a few cache
// lines per gather instead of one each).  An x or y step moves the offset by
// a constant; a z step moves it by +-1 inside a quad and by +-(4MN - 3)
// across a quad boundary, i.e. when the plane crossed has an index that is a
// multiple of 4.
// One axis of the fast walk's plane model: plane n is crossed at
// alpha = fma(n, k, b0) with k = v/(x2 - x1) and b0 = (b - x1)/(x2 - x1).  An
// axis whose ray component is below 1e-6 of a voxel (the reference's parallel
// test) never steps; the ray then stays in slab i0 = floor((x1 - b)/v).
// The matched gather adjoint evaluates the same planes per detector column
// and row, so every segment is the same number in both.
struct PlaneAxis {
    float k, b0;
    int sg, i0;
    bool par;
    // clip [lo, hi] to the grid's planes 0 and n
    __device__ __forceinline__ void clip(float &lo, float &hi, int n) const {
        if (par) return;
        const float g0 = b0, gn = fmaf((float)n, k, b0);
        lo = hmax(lo, hmin(g0, gn));
        hi = hmin(hi, hmax(g0, gn));
    }
    // the slab holding alpha lo, starting from the reference setup's index
    __device__ __forceinline__ bool start(int v0, float lo, int n, int &v) const {
        if (par) { v = i0; return (unsigned)v < (unsigned)n; }
        v = v0;
        if (fmaf((float)(v + (sg < 0)), k, b0) > lo) v -= sg;        // lo lies before the entry plane
        else if (fmaf((float)(v + (sg > 0)), k, b0) < lo) v += sg;   // lo lies past the exit plane
        return (unsigned)v < (unsigned)n;
    }
};
__device__ __forceinline__ PlaneAxis plane_axis(float v, float iv, float b, float x1, float id, int sg, bool par) {
    PlaneAxis a;
    a.k = v * id; a.b0 = (b - x1) * id; a.sg = sg; a.par = par;
    a.i0 = (int)floorf((x1 - b) * iv);
    return a;
}

template <int ES, bool ZQ = false> struct FastWalk {
    float kx, ky, kz, bx0, by0, bz0, sgxf, sgyf, sgzf, nfx, nfy, nfz, ax, ay, az, ac, amax, dist, len;
    int dix, diy, diz, dizb;
    unsigned idx;
    // returns false for an empty ray (the forward yields 0 for it)
    __device__ __forceinline__ bool init(const DGeo<float> &d, float s, float c, int col, int row, int src) {
        RayState<float> r = ray_setup(d, s, c, col, row, src);
        if (!r.hit) return false;
        const unsigned M = d.M, N = d.N, Z = d.Z, MN = M * N;
        // The reference's walk begins in the voxel its FP32 setup names and
        // stops before the first segment when that voxel lies outside the
        // grid, so the ray comes out as 0 (V63/kernel_IterativeRecon_universal.cu
        // :271-284, :297-303, :554-563).  This happens at exact ties, where
        // the entry point rounds onto the outer side of an edge plane; such
        // rays are empty here too, in the forward and the matched adjoint.
        if ((unsigned)r.vi >= M || (unsigned)r.vj >= N || (unsigned)r.vk >= Z) return false;
        const PlaneAxis X = plane_axis(d.vx, d.ivx, d.bx, r.x1x, r.ixd, r.sgx, d.vx * 1e-6f > fabsf(r.x1x - r.x2x));
        const PlaneAxis Y = plane_axis(d.vy, d.ivy, d.by, r.x1y, r.iyd, r.sgy, d.vy * 1e-6f > fabsf(r.x1y - r.x2y));
        const PlaneAxis Zx = plane_axis(d.vz, d.ivz, d.bz, r.x1z, r.izd, r.sgz, d.vz * 1e-6f > fabsf(r.x1z - r.x2z));
        // the walk runs over [alpha_min, alpha_max] of the reference setup,
        // clipped to the grid's entry and exit planes as this walk computes
        // them (so no step past the grid carries a nonzero length)
        float lo = r.ac, hi = r.amax;
        X.clip(lo, hi, (int)M); Y.clip(lo, hi, (int)N); Zx.clip(lo, hi, (int)Z);
        if (!(hi > lo)) return false;
        // start voxel: the slab that holds alpha_min under these planes
        int vi, vj, vk;
        if (!X.start(r.vi, lo, (int)M, vi) || !Y.start(r.vj, lo, (int)N, vj) || !Zx.start(r.vk, lo, (int)Z, vk))
            return false;
        // a parallel axis gets k = 0, b0 = +inf: its pending crossing is +inf
        kx = X.par ? 0.0f : X.k; bx0 = X.par ? __builtin_inff() : X.b0;
        ky = Y.par ? 0.0f : Y.k; by0 = Y.par ? __builtin_inff() : Y.b0;
        kz = Zx.par ? 0.0f : Zx.k; bz0 = Zx.par ? __builtin_inff() : Zx.b0;
        sgxf = (float)r.sgx; sgyf = (float)r.sgy; sgzf = (float)r.sgz;
        nfx = (float)(vi + (r.sgx > 0)); nfy = (float)(vj + (r.sgy > 0)); nfz = (float)(vk + (r.sgz > 0));
        ax = fmaf(nfx, kx, bx0); ay = fmaf(nfy, ky, by0); az = fmaf(nfz, kz, bz0);
        if constexpr (ZQ) {
            dix = r.sgx * 4 * ES; diy = r.sgy * (int)M * 4 * ES;
            diz = r.sgz * ES; dizb = r.sgz * (4 * (int)MN - 3) * ES;
            idx = (((unsigned)(vk >> 2) * MN + (unsigned)vj * M + (unsigned)vi) * 4u + ((unsigned)vk & 3u)) * (unsigned)ES;
        } else {
            dix = r.sgx * ES; diy = r.sgy * (int)M * ES; diz = r.sgz * (int)MN * ES; dizb = diz;
            idx = ((unsigned)vk * MN + (unsigned)vj * M + (unsigned)vi) * (unsigned)ES;
        }
        ac = lo; amax = hi; dist = r.dist;
        len = dist * (hi - lo);
        return true;
    }
    // one step: the segment's alpha length (times dist = its length in mm)
    // and the offset of the voxel it lies in.  alpha_min <= every pending
    // crossing, so once ac reaches alpha_max every later step has length 0;
    // the caller tests for the end of the ray once per batch (ended()).
    __device__ __forceinline__ float step(unsigned &id) {
        // the crossings are fma results (canonical), so the compiler's
        // minimum is the bare v_min3/v_min pair with no hazard padding
        const float m = __builtin_fminf(__builtin_fminf(__builtin_fminf(ax, ay), az), amax);
        const float dl = m - ac;
        id = idx;
        const bool sx = ax == m, sy = ay == m, sz = az == m;
        // ZQ: the z plane about to be crossed bounds a quad
        const bool qb = ZQ && __builtin_amdgcn_fractf(nfz * 0.25f) == 0.0f;
        // the stepped axes move to their next plane; every pending crossing
        // is re-evaluated from its plane index (the same value when unmoved)
        nfx += sx ? sgxf : 0.0f; nfy += sy ? sgyf : 0.0f; nfz += sz ? sgzf : 0.0f;
        ax = fmaf(nfx, kx, bx0); ay = fmaf(nfy, ky, by0); az = fmaf(nfz, kz, bz0);
        idx += (unsigned)((sx ? dix : 0) + (sy ? diy : 0) + (sz ? (qb ? dizb : diz) : 0));
        ac = m;
        return dl;
    }
    __device__ __forceinline__ bool ended() const { return !(amax - ac > loop_eps<float>()); }
    // continue the walk from alpha a in [ac, amax]: the slab of each
    // stepping axis that holds a (as start() finds it), its pending crossing
    // and the voxel offset
    __device__ __forceinline__ void restart(const DGeo<float> &d, float a) {
        const int vi = restart_axis(kx, bx0, sgxf, nfx, a, d.M);
        const int vj = restart_axis(ky, by0, sgyf, nfy, a, d.N);
        const int vk = restart_axis(kz, bz0, sgzf, nfz, a, d.Z);
        ax = fmaf(nfx, kx, bx0); ay = fmaf(nfy, ky, by0); az = fmaf(nfz, kz, bz0);
        const unsigned M = d.M, MN = (unsigned)d.M * (unsigned)d.N;
        if constexpr (ZQ)
            idx = (((unsigned)(vk >> 2) * MN + (unsigned)vj * M + (unsigned)vi) * 4u + ((unsigned)vk & 3u)) * (unsigned)ES;
        else
            idx = ((unsigned)vk * MN + (unsigned)vj * M + (unsigned)vi) * (unsigned)ES;
        ac = a;
    }
    __device__ __forceinline__ static int restart_axis(float k, float b0, float sg, float &nf, float a, int n) {
        const int up = sg > 0.0f;
        if (k == 0.0f) return (int)nf - up;   // parallel axis: its slab never changes
        const int sgi = up ? 1 : -1;
        int v = min(max((int)floorf((a - b0) / k), 0), n - 1);
        if (fmaf((float)(v + 1 - up), k, b0) > a) v -= sgi;     // a lies before the entry plane
        else if (fmaf((float)(v + up), k, b0) < a) v += sgi;    // a lies past the exit plane
        v = min(max(v, 0), n - 1);
        nf = (float)(v + up);
        return v;
    }
};

constexpr int FAST_B = 16;  // steps per batch: their gathers are in flight together

// Fast-mode ray sum: batches of FAST_B steps whose loads are issued together;
// each batch is summed in FP32 and the batch sums accumulate in FP32 with
// Kahan compensation.
// Volumes under 4 GiB are read with range-checked buffer loads on 32-bit
// byte offsets; larger ones use 64-bit addresses with an explicit range test.
// The batch loop of the fast ray sum over an initialised walker: batches of
// FAST_B steps whose loads are issued together, each batch summed in FP32 and
// the batch sums accumulated Kahan-compensated in FP32.  Returns the sum of
// alpha-length x value
// (times dist = the ray sum).  Volumes under 4 GiB are read with
// range-checked buffer loads on 32-bit byte offsets; larger ones use 64-bit
// addresses with an explicit range test.
template <typename SV, bool BUF, bool ZQ>
__device__ __forceinline__ float walk_sum(FastWalk<BUF ? (int)sizeof(SV) : 1, ZQ> &w, const DGeo<float> &d,
                                          const SV *__restrict__ f) {
    // elements of the array (the z-quad layout pads z to a multiple of 4)
    const unsigned nvox = (unsigned)d.M * (unsigned)d.N * (ZQ ? (unsigned)((d.Z + 3) & ~3) : (unsigned)d.Z);
    // (for the 64-bit path the resource is built but never used)
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(f, BUF ? nvox * (unsigned)sizeof(SV) : 0u);
    float sum = 0.0f, comp = 0.0f;   // Kahan-compensated total of the batch sums
    int guard = d.M + d.N + d.Z + 4;
    do {
        unsigned ids[FAST_B];
        float dl[FAST_B];
#pragma unroll
        for (int u = 0; u < FAST_B; ++u) dl[u] = w.step(ids[u]);
        guard -= FAST_B;
        float fv[FAST_B];
#pragma unroll
        for (int u = 0; u < FAST_B; ++u) {
            if constexpr (BUF) fv[u] = bload<SV>(rs, ids[u]);
            else fv[u] = ids[u] < nvox ? (float)ldf(f, (int64_t)ids[u]) : 0.0f;
        }
        float ps = dl[0] * fv[0];
#pragma unroll
        for (int u = 1; u < FAST_B; ++u) ps = fmaf(dl[u], fv[u], ps);
        const float y = ps - comp, t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    } while (!w.ended() && guard > 0);
    return sum - comp;
}

template <typename SV, bool BUF, bool ZQ = false>
__device__ __forceinline__ float siddon_walk_fast(const DGeo<float> &d, float s, float c, int col, int row, int src,
                                                  const SV *__restrict__ f, float &len_out) {
    FastWalk<BUF ? (int)sizeof(SV) : 1, ZQ> w;
    len_out = 0.0f;
    if (!w.init(d, s, c, col, row, src)) return 0.0f;
    len_out = w.len;
    return walk_sum<SV, BUF, ZQ>(w, d, f) * w.dist;
}

template <typename SV>
__device__ __forceinline__ float siddon_sum_fast(const DGeo<float> &d, float s, float c, int col, int row, int src,
                                                 const SV *__restrict__ f, float &len_out) {
    if ((uint64_t)d.M * d.N * d.Z * sizeof(SV) < (1ull << 32))
        return siddon_walk_fast<SV, true>(d, s, c, col, row, src, f, len_out);
    return siddon_walk_fast<SV, false>(d, s, c, col, row, src, f, len_out);
}

// 16x16-pixel tile per workgroup (XCD-aware tile order), 8x8 per wave;
// returns the tile's image slot
__device__ __forceinline__ int tile_coords(int &col, int &row) {
    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63;
    const int3 t = xcd_tile();
    col = t.x * 16 + (w & 1) * 8 + (l & 7);
    row = t.y * 16 + (w >> 1) * 8 + (l >> 3);
    return t.z;
}

// CMD 0: ray sum; CMD 1: SART correction (b - sum)/len
// (V63/kernel_IterativeRecon_universal.cu:554-563).  image slot = q*NS + src
// for batch slot q; view = first + q*stride; meas is indexed by view - mv0.
template <typename T, typename SV, typename SO, typename SM, int CMD, bool COMP>
__global__ __launch_bounds__(256) void k_forward(DGeo<T> d, const T *__restrict__ ang,
                                                 const SV *__restrict__ f, SO *__restrict__ out,
                                                 const SM *__restrict__ meas, int first, int stride,
                                                 int mv0, int z0) {
    int col, row;
    const int slot = tile_coords(col, row);
    if (col >= d.R || row >= d.ZP) return;
    const int img = slot + z0;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = first + q * stride;
    const T s = ang[2 * v], c = ang[2 * v + 1];
    T len;
    T sum;
    if constexpr (COMP && std::is_same<T, float>::value) sum = siddon_sum_fast<SV>(d, s, c, col, row, src, f, len);
    else sum = siddon_sum<T, SV, 4, false>(d, s, c, col, row, src, f, len);
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t pix = (int64_t)row * d.R + col;
    T r;
    if (len < d.vz * (T)1e-6) r = 0;
    else if (CMD == 0) r = sum;
    else {
        const T b = (T)ldf(meas, ((int64_t)(v - mv0) * d.NS + src) * P + pix);
        r = (b - sum) / len;
    }
    St<SO>::st(out, (int64_t)img * P + pix, r);
}

// SART correction (CMD 1) of the sweep on a z-quad-layout volume (fast
// mode, FP32 arithmetic, volume under 4 GiB): the same rays, segments and
// output as k_forward<float, SV, float, SM, 1, true>.
// Projection on a z-quad-layout volume (fast mode, FP32 arithmetic, volume
// under 4 GiB): CMD 0 ray sums, CMD 1 SART corrections (b - sum)/len, the same
// rays and segments as k_forward.  S parts per ray: a workgroup holds 256 / S
// rays (16x16 pixels for S = 1, 16x8 for 2, 8x8 for 4); part p walks alpha in
// [lo + (hi - lo) p / S, lo + (hi - lo) (p + 1) / S] and the partial sums meet
// in LDS, added in part order.  Splitting gives more waves per view and parts
// of equal length, so more gathers are in flight; it also splits the segment
// that holds each cut point, so the SART correction uses it (S = 2) and the ray
// sums of the autograd pair do not (their segments stay those of the matched
// adjoint).
constexpr int CORR_SPLIT = 2;

template <typename SV, typename SO, typename SM, int CMD, int S>
__global__ __launch_bounds__(256) void k_project_zq(DGeo<float> d, const float *__restrict__ ang,
                                                    const SV *__restrict__ fq, SO *__restrict__ out,
                                                    const SM *__restrict__ meas, int first, int stride, int mv0,
                                                    int z0) {
    constexpr int NR = 256 / S, WPP = 4 / S;   // rays and waves per part
    constexpr int TW = WPP >= 2 ? 16 : 8, TH = WPP == 4 ? 16 : 8;
    __shared__ float part[S > 1 ? 256 : 1];
    const int tid = threadIdx.x, p = tid / NR, rr = tid % NR;
    const int w = rr >> 6, l = rr & 63;
    const int3 t = xcd_tile();
    const int col = t.x * TW + (w & 1) * 8 + (l & 7), row = t.y * TH + (w >> 1) * 8 + (l >> 3);
    const int img = t.z + z0;
    const bool inb = col < d.R && row < d.ZP;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = first + q * stride;
    float sum = 0.0f;
    float len = 0.0f, dist = 0.0f;
    if (inb) {
        const float s = ang[2 * v], c = ang[2 * v + 1];
        FastWalk<(int)sizeof(SV), true> fw;
        if (fw.init(d, s, c, col, row, src)) {
            len = fw.len;
            dist = fw.dist;
            if constexpr (S > 1) {
                const float lo = fw.ac, span = fw.amax - fw.ac;
                if (p + 1 < S) fw.amax = fmaf(span, (float)(p + 1) / (float)S, lo);
                if (p > 0) fw.restart(d, fmaf(span, (float)p / (float)S, lo));
            }
            sum = walk_sum<SV, true, true>(fw, d, fq);
        }
    }
    if constexpr (S > 1) {
        part[tid] = sum;
        __syncthreads();
        if (p != 0) return;
#pragma unroll
        for (int k = 1; k < S; ++k) sum += part[k * NR + rr];
    }
    if (!inb) return;
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t pix = (int64_t)row * d.R + col;
    float r = 0.0f;
    if (!(len < d.vz * 1e-6f)) {
        const float ax = sum * dist;
        if constexpr (CMD == 0) r = ax;
        else r = ((float)ldf(meas, ((int64_t)(v - mv0) * d.NS + src) * P + pix) - ax) / len;
    }
    St<SO>::st(out, (int64_t)img * P + pix, r);
}

// Column walk (fast mode, z-quad volume).  Every ray of one detector column
// runs from a point of the column to the column's source point, and for all
// three beam styles those points share their x and y: the rays of a column
// lie in one vertical plane and cross the x and y planes at the same alpha
// (fma(n, k, b0) over the same k and b0, so the same FP32 numbers).  A
// workgroup takes 256 rows of one column and merges the column's x and y
// crossings once into an LDS list; each list segment is one x/y voxel, and at
// a segment all 256 rays read that voxel column of z, within a window the
// list also records (the extreme rows at the segment's ends, one voxel of
// margin).  The workgroup stages those windows into LDS with coalesced
// 16-byte quad loads, double-buffered over batches of segments, and every
// lane walks the list in step with the others, clipped to its own
// [alpha_min, alpha_max], tracking only its own z slab and reading its voxels
// from the staged window.  A segment holds at most one z crossing when
// |dz|/vz < max(|dx|/vx, |dy|/vy) along every ray (the host checks a bound);
// the part after the crossing reads the next z voxel.  The segments are those
// of the per-ray walk (same alphas, same voxels; a tie there is one step, here
// a zero-length segment), and the sums keep its FP32 batches with a
// Kahan-compensated FP32 total.  Per segment ~14 VALU (staging included), one
// LDS pair read and no per-lane gather, against ~25 VALU and one scattered
// gather per step of the per-ray walk.  Ray-sum batches use it for views
// whose rays run along x (see col_walk_view); the SART sweep for every view,
// on a transposed copy of its volume for views whose rays run along y.
constexpr int COL_ROWS = 256;   // rays per workgroup: rows of one column
constexpr int COL_B = 16;       // list padding: a multiple of every staged batch (col_batch)
constexpr int COL_NQ = 48;      // largest staged window, in z-quads
// LDS elements per staged window: COL_NQ quads and one quad of padding, so
// that the windows of a batch's segments start on different banks (the
// staging writes of eight segments' same quad are then conflict-free)
constexpr int COL_W = 4 * COL_NQ + 4;
// List segments per staged batch: 16 (one barrier per 16 segments, and a
// staging load's eight-lane groups cover 16 segments), or 8 for BF16
// volumes, whose FP32 windows at 16 would leave four workgroups per CU
// instead of six (correction per view at 512^3 / 1024^2: FP32 0.382 -> 0.342
// ms and FP16 0.286 -> 0.261 ms from 8 to 16; BF16 0.263 ms at either).
template <typename SV> constexpr int col_batch() { return std::is_same<SV, __hip_bfloat16>::value ? 8 : 16; }
// staged quads per thread and batch
template <typename SV> constexpr int col_sp() { return (col_batch<SV>() * COL_NQ + COL_ROWS - 1) / COL_ROWS; }

// LDS: two staging buffers of col_batch windows in a static array (fixed
// address and stride, so every staged read has a constant offset; FP32
// elements, or the volume's own halves for FP16); the list in dynamic LDS as
// three arrays of at most M + N - 1 entries padded to a multiple of COL_B:
// alpha, the window's first quad times the staged bytes per quad (its byte
// offset in a staged row), and the x/y byte offset for the staging loads.
// (The entry count is a multiple of 4, so the alpha and window arrays stay
// 16-byte aligned for their vector reads; col_lds_bytes bounds the largest
// workgroup.)
static inline size_t col_list_entries(int M, int N) { return (size_t)((M + N + COL_B + 3) & ~3); }
static inline size_t col_dyn_bytes(int M, int N) { return 12 * col_list_entries(M, N); }
static inline size_t col_lds_bytes(int M, int N) {
    return sizeof(float) * 2 * COL_B * COL_W + col_dyn_bytes(M, N);
}

// Element type of the staged windows: FP16 volumes stage their halves as
// loaded (the walk's FMAs take FP16 operands directly), the others FP32.
template <typename SV> struct StageOf { using type = float; };
template <> struct StageOf<__half> { using type = _Float16; };

// TR: the volume is in the transposed z-quad layout [z/4][x][y][z%4], for
// views whose rays run along y (the staged quads of consecutive segments then
// share cache lines, as they do along x in the plain z-quad layout).
template <typename SV, typename SO, typename SM, int CMD, bool TR = false>
__global__ __launch_bounds__(COL_ROWS) void k_project_col(DGeo<float> d, const float *__restrict__ ang,
                                                          const SV *__restrict__ fq, SO *__restrict__ out,
                                                          const SM *__restrict__ meas, int first, int stride,
                                                          int mv0, int z0, int NQ) {
    using SE = typename StageOf<SV>::type;
    constexpr int SB = (int)sizeof(SE);                // staged bytes per voxel
    constexpr int QSH = SB == 4 ? 4 : 3;               // log2 of the staged bytes per quad
    constexpr int CB = col_batch<SV>(), SP = col_sp<SV>();
    __shared__ __align__(16) SE stg[2 * CB * COL_W];
    extern __shared__ __align__(16) unsigned char col_smem[];
    const int LE = (d.M + d.N + COL_B + 3) & ~3;   // col_list_entries
    float *const la = reinterpret_cast<float *>(col_smem);             // alpha
    int *const lw = reinterpret_cast<int *>(col_smem + 4 * LE);          // first quad x 4 SB
    unsigned *const ls = reinterpret_cast<unsigned *>(col_smem + 8 * LE); // x/y byte offset
    constexpr int ES = (int)sizeof(SV);
    const int tid = (int)threadIdx.x;
    const int3 t = xcd_tile();
    const int col = t.x, row0 = t.y * COL_ROWS, row = row0 + tid;
    const int rowL = min(row0 + COL_ROWS, d.ZP) - 1;
    const int img = t.z + z0;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = first + q * stride;
    const float s = ang[2 * v], c = ang[2 * v + 1];
    const unsigned M = d.M, MN = (unsigned)d.M * (unsigned)d.N;
    const int ZQ4 = (d.Z + 3) >> 2;                        // quads per (x, y)
    const unsigned QS = MN * 4u * (unsigned)ES;            // bytes between z-quad planes

    // the column's x/y walk (the x and y terms of ray_setup and FastWalk::init)
    const float u = d.ymin + (float)col * d.du;
    const float x1x = d.dod * c - u * s, x1y = d.dod * s + u * c;
    float x2x, x2y, x2z0, x2zL;
    const float z1a = d.zmin + (float)row0 * d.dv, z1b = d.zmin + (float)rowL * d.dv;
    ray_source(d, s, c, u, z1a, src, x2x, x2y, x2z0);
    ray_source(d, s, c, u, z1b, src, x2x, x2y, x2zL);
    const float ixd = 1.0f / (x2x - x1x), iyd = 1.0f / (x2y - x1y);
    float a0 = (d.bx - x1x) * ixd, a1 = (d.bx + d.vx * (float)d.M - x1x) * ixd;
    const float axmin = fminf(a0, a1), axmax = fmaxf(a0, a1);
    a0 = (d.by - x1y) * iyd; a1 = (d.by + d.vy * (float)d.N - x1y) * iyd;
    const float aymin = fminf(a0, a1), aymax = fmaxf(a0, a1);
    const int sgx = (x1x < x2x) ? 1 : ((x1x > x2x) ? -1 : 0);
    const int sgy = (x1y < x2y) ? 1 : ((x1y > x2y) ? -1 : 0);
    const PlaneAxis X = plane_axis(d.vx, d.ivx, d.bx, x1x, ixd, sgx, d.vx * 1e-6f > fabsf(x1x - x2x));
    const PlaneAxis Y = plane_axis(d.vy, d.ivy, d.by, x1y, iyd, sgy, d.vy * 1e-6f > fabsf(x1y - x2y));
    float lxy = fmaxf(axmin, aymin), hxy = fminf(axmax, aymax);
    X.clip(lxy, hxy, d.M); Y.clip(lxy, hxy, d.N);
    int vi0 = X.i0, vj0 = Y.i0, NX = 0, NY = 0;
    bool cmiss = !(hxy > lxy);
    if (!cmiss) {
        float nf;
        if (!X.par) {
            vi0 = FastWalk<ES, true>::restart_axis(X.k, X.b0, (float)sgx, nf, lxy, d.M);
            NX = max((FastWalk<ES, true>::restart_axis(X.k, X.b0, (float)sgx, nf, hxy, d.M) - vi0) * sgx, 0);
        }
        if (!Y.par) {
            vj0 = FastWalk<ES, true>::restart_axis(Y.k, Y.b0, (float)sgy, nf, lxy, d.N);
            NY = max((FastWalk<ES, true>::restart_axis(Y.k, Y.b0, (float)sgy, nf, hxy, d.N) - vj0) * sgy, 0);
        }
        cmiss = (unsigned)vi0 >= M || (unsigned)vj0 >= (unsigned)d.N;
    }
    const int K = cmiss ? 0 : NX + NY;                       // crossings; entry K ends at hxy
    const int KP = cmiss ? 0 : (K + COL_B) / COL_B * COL_B;  // entries incl. padding
    if (!cmiss) {
        const float px0 = (float)(vi0 + (sgx > 0)), py0 = (float)(vj0 + (sgy > 0));
        const float fsx = (float)sgx, fsy = (float)sgy;
        auto XA = [&](int i) { return fmaf(fmaf((float)i, fsx, px0), X.k, X.b0); };
        auto YA = [&](int j) { return fmaf(fmaf((float)j, fsy, py0), Y.k, Y.b0); };
        auto OFF = [&](int i, int j) {
            if constexpr (TR) return (unsigned)(((vi0 + i * sgx) * d.N + (vj0 + j * sgy)) * 4 * ES);
            else return (unsigned)(((vj0 + j * sgy) * (int)M + (vi0 + i * sgx)) * 4 * ES);
        };
        // first staged quad of a segment [ap, a]: the z voxel range of the
        // workgroup's first and last rows at both ends, one voxel of margin
        auto QLO = [&](float ap, float a) {
            const float za = fminf(fmaf(ap, x2z0 - z1a, z1a), fmaf(a, x2z0 - z1a, z1a));
            const float zb = fminf(fmaf(ap, x2zL - z1b, z1b), fmaf(a, x2zL - z1b, z1b));
            const int vlo = (int)floorf((fminf(za, zb) - d.bz) * d.ivz) - 1;
            return vlo >> 2;   // floor division by 4 (arithmetic shift)
        };
        // merge path: thread t writes entries [t C, t C + C); ties take x first
        const int C = (KP + COL_ROWS - 1) / COL_ROWS;
        const int e0 = tid * C;
        if (e0 < KP) {
            const int ee = min(e0, K);
            int ilo = max(0, ee - NY), ihi = min(ee, NX);
            while (ilo < ihi) {
                const int mid = (ilo + ihi) >> 1;
                if (XA(mid) <= YA(ee - mid - 1)) ilo = mid + 1;
                else ihi = mid;
            }
            int i = ilo, j = ee - ilo;
            // alpha at which this thread's first segment starts
            float ap = lxy;
            if (ee > 0) ap = fmaxf(i > 0 ? XA(i - 1) : -__builtin_inff(), j > 0 ? YA(j - 1) : -__builtin_inff());
            if (e0 > K) ap = hxy;
            for (int e = e0; e < min(e0 + C, KP); ++e) {
                float a;
                const unsigned off = OFF(i, j);
                if (e >= K) a = hxy;
                else if (i < NX && (j >= NY || XA(i) <= YA(j))) { a = XA(i); ++i; }
                else { a = YA(j); ++j; }
                const int ql = QLO(ap, a);
                la[e] = a;
                lw[e] = ql * (1 << QSH);
                ls[e] = off;
                ap = a;
            }
        }
    }

    // this ray: its length as the per-ray walk forms it (0 for the rays the
    // reference leaves empty), and its height in voxel units along the
    // column's alpha, zeta(alpha) = (z1 + alpha (z2 - z1) - bz)/vz, as one FMA
    // zeta is kept relative to the slab kb the ray enters in, so that it
    // stays small and its rounding does not move the crossings of nearly
    // level rays; kbb = SB kb (bytes) re-bases the window offsets.
    FastWalk<ES, true> fw;
    float len = 0.0f, dist = 0.0f;
    float cz = 0.0f, z0c = 0.0f, icz = 0.0f;
    int kbb = 0;
    {
        const float z1 = d.zmin + (float)row * d.dv;
        float sx, sy, z2;
        ray_source(d, s, c, u, z1, src, sx, sy, z2);
        cz = (z2 - z1) * d.ivz;
        const float kb = floorf(fmaf(lxy, cz, (z1 - d.bz) * d.ivz));
        kbb = SB * (int)kb;
        z0c = fmaf(z1 - d.bz, d.ivz, -kb);
        // alpha per unit of zeta (0 for a level ray, which never changes slab)
        icz = fabsf(cz) > 1e-30f ? 1.0f / fabsf(cz) : 0.0f;
    }
    // CMD 0 (ray sums of the autograd pair): the per-ray walk's own segments,
    // so that the matched adjoint stays its exact transpose.  The lane clips
    // the list to its [alpha_min, alpha_max] and tracks its next z-plane
    // crossing fma(nfz, kz, bz0); zrb is its slab relative to kb, times SB.
    float lo = 0.0f, hi = 0.0f, az = __builtin_inff(), nfz = 0.0f, kz = 0.0f, bz0 = __builtin_inff(), sgzf = 0.0f;
    int zrb = 0, dzb = 0;
    if (!cmiss && row < d.ZP && fw.init(d, s, c, col, row, src)) {
        len = fw.len; dist = fw.dist;
        if constexpr (CMD == 0) {
            lo = fw.ac; hi = fw.amax;
            az = fw.az; nfz = fw.nfz; kz = fw.kz; bz0 = fw.bz0; sgzf = fw.sgzf;
            zrb = ((int)nfz - (sgzf > 0.0f ? 1 : 0)) * SB - kbb;
            dzb = (int)sgzf * SB;
        }
    }
    const bool up = cz >= 0.0f;
    const unsigned nbytes = MN * (unsigned)ZQ4 * 4u * (unsigned)ES;
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(fq, nbytes);
    // this thread's staged quads (the same in every batch): segment k of the
    // batch and quad qi of its window (sk = -1: none), and where they go
    const int nst = CB * NQ;   // quads staged per batch (at most SP * COL_ROWS)
    // Segment-fastest: lanes 8q .. 8q + 7 load quad q of the batch's eight
    // segments.  Consecutive segments are mostly neighbours along the
    // layout's fast axis, whose quads share a 128-byte line, so a wave's load
    // touches ~8-16 lines instead of one per lane (the quads of one segment
    // lie a z-quad plane, M N quads, apart).
    int sk[SP], sq[SP], so[SP];
#pragma unroll
    for (int p = 0; p < SP; ++p) {
        const int sidx = tid + p * COL_ROWS;
        sk[p] = sidx < nst ? sidx % CB : -1;
        sq[p] = sidx / CB;
        so[p] = (sidx % CB) * COL_W + 4 * sq[p];
    }
    // a quad outside the volume reads as zeros: its index is clamped to -1 or
    // ZQ4, whose offsets lie past the buffer's end (the volume is under 2 GiB)
    // Staged quads stay in registers as loaded (FP32 quads, or the 16-bit
    // volume's raw halves; BF16 is widened only at the LDS write), so no
    // instruction waits for them before the commit.
    using SR = typename std::conditional<ES == 2, uint2, float4>::type;
    SR sreg[SP];
    auto issue = [&](int b) {
#pragma unroll
        for (int p = 0; p < SP; ++p) {
            if (sk[p] >= 0) {
                const int e = b * CB + sk[p];
                const int qz = min(max((lw[e] >> QSH) + sq[p], -1), ZQ4);
                const uint32_t off = ls[e] + (unsigned)qz * QS;
                if constexpr (ES == 2) {
                    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
                    sreg[p] = make_uint2(v[0], v[1]);
                } else {
                    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
                    sreg[p] = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                                          __uint_as_float(v[3]));
                }
            }
        }
    };
    auto commit = [&](auto bufc) {
        constexpr int buf = decltype(bufc)::value;
#pragma unroll
        for (int p = 0; p < SP; ++p) {
            if (sk[p] < 0) continue;
            SE *const dst = stg + (buf * CB * COL_W + so[p]);
            if constexpr (std::is_same<SV, __hip_bfloat16>::value) {
                const uint2 r = sreg[p];
                *reinterpret_cast<float4 *>(dst) = make_float4(__uint_as_float(r.x << 16), __uint_as_float(r.x & 0xffff0000u),
                                                               __uint_as_float(r.y << 16), __uint_as_float(r.y & 0xffff0000u));
            } else {
                *reinterpret_cast<SR *>(dst) = sreg[p];
            }
        }
    };
    __syncthreads();   // the list is complete
    const int nb = KP / CB;
    if (nb > 0) { issue(0); commit(std::integral_constant<int, 0>()); }
    // the lower element of a segment's slab pair lies at most 4 NQ - 2 into
    // its window (one voxel of margin each way by construction; clamped)
    const unsigned wmax2 = (unsigned)((4 * NQ - 2) * SB);
    // Kahan-compensated FP32 total of the batches' FP32 sums
    float sum = 0.0f, comp = 0.0f;
    float as = lxy, zs = fmaf(lxy, cz, z0c);
    // CMD 1 (SART corrections): a segment [as, ae] of the column's list lies
    // in one x/y voxel; this ray
    // enters it in slab k = floor(zeta(as)) and, as |dzeta| < 1 per segment
    // (col_walk_nq), crosses at most one z plane in it.  The crossing part,
    // in zeta units clamp(zeta(ae) - (k + 1), 0, 1) for a rising ray or
    // clamp(k - zeta(ae), 0, 1) for a falling one, times alpha per unit of
    // zeta, belongs to the next slab and the rest of the segment to slab k;
    // both slabs are one LDS pair read.  Parts below or above the volume read
    // the window's zeros, so no per-ray clipping is needed: the sum is the
    // per-ray walk's to rounding.
    auto segs = [&](auto upc, const char *sb, const float *ea, const int *ew) {
        constexpr bool UP = decltype(upc)::value;
        float ps = 0.0f;
#pragma unroll
        for (int k = 0; k < CB; ++k) {
            const float ae = ea[k];
            const float ze = fmaf(ae, cz, z0c);
            const float kf = floorf(zs);
            const float wx = (UP ? __builtin_amdgcn_fmed3f(ze - kf - 1.0f, 0.0f, 1.0f)
                                 : __builtin_amdgcn_fmed3f(kf - ze, 0.0f, 1.0f)) * icz;
            const float wy = (ae - as) - wx;
            const int klo = (int)kf - (UP ? 0 : 1);
            const unsigned a = min((unsigned)(klo * SB + ew[k]), wmax2);
            const SE *pr = reinterpret_cast<const SE *>(sb + k * COL_W * SB + a);
            const float f0 = (float)pr[0], f1 = (float)pr[1];
            ps = fmaf(UP ? wy : wx, f0, ps);
            ps = fmaf(UP ? wx : wy, f1, ps);
            as = ae;
            zs = ze;
        }
        return ps;
    };
    const unsigned wmax = (unsigned)((4 * NQ - 1) * SB);
    float asx = lo;
    auto segs_exact = [&](const char *sb, const float *ea, const int *ew) {
        float ps = 0.0f;
#pragma unroll
        for (int k = 0; k < CB; ++k) {
            const float ae = __builtin_amdgcn_fmed3f(ea[k], lo, hi);
            const float m = fminf(az, ae);
            const float d0 = m - asx, d1 = ae - m;
            const bool cr = az < ae;
            // the slab before and after the crossing, as byte offsets in the
            // window (an idle lane's slab may lie outside it: clamped, and its
            // segment lengths are 0)
            const unsigned r0 = min((unsigned)(zrb + ew[k]), wmax);
            zrb += cr ? dzb : 0;
            nfz += cr ? sgzf : 0.0f;
            az = fmaf(nfz, kz, bz0);
            const unsigned r1 = min((unsigned)(zrb + ew[k]), wmax);
            ps = fmaf(d0, (float)*reinterpret_cast<const SE *>(sb + k * COL_W * SB + r0), ps);
            ps = fmaf(d1, (float)*reinterpret_cast<const SE *>(sb + k * COL_W * SB + r1), ps);
            asx = ae;
        }
        return ps;
    };
    // one batch from staging buffer BUF; the next batch's quads are loaded
    // meanwhile and go to the other buffer
    auto batch = [&](auto bufc, int b) {
        constexpr int BUF = decltype(bufc)::value;
        __syncthreads();   // batch b staged; batch b - 1's readers are done with the other buffer
        if (b + 1 < nb) issue(b + 1);
        const char *sb = reinterpret_cast<const char *>(stg + BUF * CB * COL_W);
        float ea[CB];
        int ew[CB];
#pragma unroll
        for (int k = 0; k < CB; k += 4) {
            const float4 a4 = *reinterpret_cast<const float4 *>(la + b * CB + k);
            const int4 w4 = *reinterpret_cast<const int4 *>(lw + b * CB + k);
            ea[k] = a4.x; ea[k + 1] = a4.y; ea[k + 2] = a4.z; ea[k + 3] = a4.w;
            // this ray's slab kb at the window's byte 0
            ew[k] = kbb - w4.x; ew[k + 1] = kbb - w4.y; ew[k + 2] = kbb - w4.z; ew[k + 3] = kbb - w4.w;
        }
        // rising and falling rays (all but one wave per column are uniform)
        float ps;
        if constexpr (CMD == 0) ps = segs_exact(sb, ea, ew);
        else ps = up ? segs(std::true_type(), sb, ea, ew) : segs(std::false_type(), sb, ea, ew);
        const float yk = ps - comp;
        const float tk = sum + yk;
        comp = (tk - sum) - yk;
        sum = tk;
        if (b + 1 < nb) commit(std::integral_constant<int, 1 - BUF>());
    };
    for (int b = 0; b < nb; b += 2) {
        batch(std::integral_constant<int, 0>(), b);
        if (b + 1 < nb) batch(std::integral_constant<int, 1>(), b + 1);
    }
    if (row >= d.ZP) return;
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t pix = (int64_t)row * d.R + col;
    float r = 0.0f;
    if (!(len < d.vz * 1e-6f)) {
        const float ax = (sum - comp) * dist;
        if constexpr (CMD == 0) r = ax;
        else r = ((float)ldf(meas, ((int64_t)(v - mv0) * d.NS + src) * P + pix) - ax) / len;
    }
    St<SO>::st(out, (int64_t)img * P + pix, r);
}

// Data fidelity sum (Ax - b)^2 fused into the ray traversal: Ax never goes to
// memory.  Each 256-ray block writes one FP64 partial (wave64 shuffles, then
// LDS across the 4 waves); the partials are folded in fixed order afterwards.
// (L2_norm_gpu + the objective's Forward_3D_ray_driven_siddon,
// V63/host_tool_functions.cu:28-58 and V63/main_recon.cu:401-407.)
template <typename T, typename SV, typename SM, bool COMP>
__global__ __launch_bounds__(256) void k_residual(DGeo<T> d, const T *__restrict__ ang, const SV *__restrict__ f,
                                                  const SM *__restrict__ meas, int view0, int z0,
                                                  double *__restrict__ part) {
    int col, row;
    const int slot = tile_coords(col, row);
    const int img = slot + z0;
    double e2 = 0.0;
    if (col < d.R && row < d.ZP) {
        const int q = img / d.NS, src = img - q * d.NS;
        const int v = view0 + q;
        const T s = ang[2 * v], c = ang[2 * v + 1];
        T len;
        T sum;
        if constexpr (COMP && std::is_same<T, float>::value) sum = siddon_sum_fast<SV>(d, s, c, col, row, src, f, len);
        else sum = siddon_sum<T, SV, 4, false>(d, s, c, col, row, src, f, len);
        const T ax = (len < d.vz * (T)1e-6) ? (T)0 : sum;
        const int64_t P = (int64_t)d.R * d.ZP;
        const double e = (double)ax - (double)ldf(meas, (int64_t)img * P + (int64_t)row * d.R + col);
        e2 = e * e;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_down(e2, o, 64);
    __shared__ double wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = e2;
    __syncthreads();
    // one slot per workgroup: launches cover images [z0, z0 + gridDim.z)
    if (threadIdx.x == 0)
        part[((int64_t)z0 * gridDim.y + blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] =
            (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

template <typename T, typename SV, typename SM>
static ct_status run_residual(ct_context *ctx, const ct_geometry &g, const void *vol, const void *meas, double *out) {
    ct_status st;
    const T *ang = (const T *)angle_table(ctx, g, sizeof(T) == 8, false, &st);
    if (!ang) return st;
    DGeo<T> d = derive<T>(g);
    const int nimg = g.nviews * d.NS;
    const int gx = (d.R + 15) / 16, gy = (d.ZP + 15) / 16;
    const int64_t nblk = (int64_t)gx * gy * nimg;
    double *part = (double *)slot(ctx, SLOT_TMP_PROJ2, sizeof(double) * nblk, &st);
    if (!part) return st;
    const bool comp = std::is_same<T, float>::value && ctx->math == CT_MATH_FAST;
    {
        ProfScope ps(ctx, CT_K_FORWARD);
        for (int z0 = 0; z0 < nimg; z0 += 65535) {
            const int nz = nimg - z0 < 65535 ? nimg - z0 : 65535;
            dim3 grid(gx, gy, nz);
            if (comp)
                hipLaunchKernelGGL((k_residual<T, SV, SM, true>), grid, dim3(256), 0, ctx->stream, d, ang,
                                   (const SV *)vol, (const SM *)meas, 0, z0, part);
            else
                hipLaunchKernelGGL((k_residual<T, SV, SM, false>), grid, dim3(256), 0, ctx->stream, d, ang,
                                   (const SV *)vol, (const SM *)meas, 0, z0, part);
        }
    }
    st = launch_check(ctx, "residual");
    if (st != CT_OK) return st;
    return reduce_dot(ctx, part, nullptr, nblk, CT_F64, 2, out);
}

template <typename T, typename SV>
static ct_status residual_meas(ct_context *ctx, const ct_geometry &g, const void *vol, const void *meas, ct_dtype mdt,
                               double *out) {
    switch (mdt) {
    case CT_F32: return run_residual<T, SV, float>(ctx, g, vol, meas, out);
    case CT_F16: return run_residual<T, SV, __half>(ctx, g, vol, meas, out);
    case CT_BF16: return run_residual<T, SV, __hip_bfloat16>(ctx, g, vol, meas, out);
    case CT_F64: return run_residual<T, SV, double>(ctx, g, vol, meas, out);
    }
    return fail(ctx, CT_ERR_DTYPE, "residual: bad measurement dtype");
}

ct_status launch_residual(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, const void *meas,
                          ct_dtype mdt, double *out) {
    switch (vdt) {
    case CT_F32: return residual_meas<float, float>(ctx, g, vol, meas, mdt, out);
    case CT_F16: return residual_meas<float, __half>(ctx, g, vol, meas, mdt, out);
    case CT_BF16: return residual_meas<float, __hip_bfloat16>(ctx, g, vol, meas, mdt, out);
    case CT_F64: return residual_meas<double, double>(ctx, g, vol, meas, mdt, out);
    }
    return fail(ctx, CT_ERR_DTYPE, "residual: bad volume dtype");
}

template <typename A> __device__ __forceinline__ void atomic_add(A *p, A v);
template <> __device__ __forceinline__ void atomic_add<float>(float *p, float v) { unsafeAtomicAdd(p, v); }
template <> __device__ __forceinline__ void atomic_add<double>(double *p, double v) { unsafeAtomicAdd(p, v); }

// Matched adjoint: scatter the forward's segments (scatter form of
// V63/pixel_driven_backprj/kernel_IterativeRecon_CBCT.cu:563-1116).  Rays the
// forward zeroes (len < 1e-6*vz) scatter nothing, so the pair is an exact
// transpose.  FAST pairs with the fast forward: the same FastWalk steps, its
// closed-form length, and no scatter for an offset past the grid (the
// forward reads 0 there).  Otherwise a first pass measures the ray length
// of the reference traversal.
template <typename T, typename SP, typename A, bool FAST, bool ZQ = false>
__global__ __launch_bounds__(256) void k_adjoint(DGeo<T> d, const T *__restrict__ ang,
                                                 const SP *__restrict__ proj, A *__restrict__ acc,
                                                 int view0, int z0) {
    int col, row;
    const int slot = tile_coords(col, row);
    if (col >= d.R || row >= d.ZP) return;
    const int img = slot + z0;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = view0 + q;
    const int64_t P = (int64_t)d.R * d.ZP;
    const T y = (T)ldf(proj, (int64_t)img * P + (int64_t)row * d.R + col);
    if (y == (T)0) return;
    const T s = ang[2 * v], c = ang[2 * v + 1];
    if constexpr (FAST) {
        FastWalk<1, ZQ> w;
        if (!w.init(d, s, c, col, row, src)) return;
        if (w.len < d.vz * 1e-6f) return;
        const unsigned nvox = (unsigned)d.M * (unsigned)d.N * (ZQ ? (unsigned)((d.Z + 3) & ~3) : (unsigned)d.Z);
        int guard = d.M + d.N + d.Z + 4;
        do {
#pragma unroll
            for (int u = 0; u < FAST_B; ++u) {
                unsigned id;
                const float dl = w.step(id);
                if (dl != 0.0f && id < nvox) atomic_add<A>(acc + id, (A)(w.dist * dl) * (A)y);
            }
            guard -= FAST_B;
        } while (!w.ended() && guard > 0);
    } else {
        const T len = siddon_traverse(d, s, c, col, row, src, [](int, T) {});
        if (len < d.vz * (T)1e-6) return;
        siddon_traverse(d, s, c, col, row, src, [&](int idx, T seg) { atomic_add<A>(acc + idx, (A)seg * (A)y); });
    }
}

// ------------------------------------------------- gather form of the adjoint
// Aᵀy without atomics, for the fast walk.  The walk's segment in voxel
// (i, j, k) is max(0, min(exit planes, alpha_hi) - max(entry planes,
// alpha_lo)) - it is in that voxel exactly while it is inside all three
// slabs - and every term is separable: the x/y plane models and the x/y parts
// of [alpha_lo, alpha_hi] depend on the detector column, the z ones on the
// row.  So each voxel gathers over the pixels of its detector footprint
// exactly the numbers the walk scatters, in any order and without atomics.
//
// Per-view tables: column  (kx, bx0, ky, by0) (lo_xy, hi_xy, i0x|par, i0y|par)
//                  row     (kz, bz0, lo_z, hi_z) (i0z|par)
//                  pixel   w = dist*y (0 for a ray the forward zeroes)
// A parallel axis is flagged by i0 >= 0 (the slab the ray stays in); -1
// marks a stepping axis.
struct AdjCol { float4 a; float4 b; };

__device__ __forceinline__ void axis_bounds(float v, float b, int n, float x1, float x2, float &k, float &b0, float &lo,
                                            float &hi, int &i0, float iv) {
    // the reference setup's slab interval along this axis
    // (V63/kernel_IterativeRecon_universal.cu:77-104)
    const float id = 1.0f / (x2 - x1);
    const float a0 = (b - x1) * id, a1 = (b + v * (float)n - x1) * id;
    lo = fminf(a0, a1); hi = fmaxf(a0, a1);
    const bool par = v * 1e-6f > fabsf(x1 - x2);
    const PlaneAxis P = plane_axis(v, iv, b, x1, id, 0, par);
    k = P.k; b0 = P.b0;
    i0 = par ? P.i0 : -1;
    if (!par) {
        const float g0 = b0, gn = fmaf((float)n, k, b0);
        lo = hmax(lo, hmin(g0, gn));
        hi = hmin(hi, hmax(g0, gn));
    }
}

// tables for images [0, nimg) of views view0 + q (q = img / NS)
template <typename SP>
__global__ __launch_bounds__(256) void k_adj_prep(DGeo<float> d, const float *__restrict__ ang, int view0,
                                                  const SP *__restrict__ proj, int nimg, float *__restrict__ pixw,
                                                  AdjCol *__restrict__ colT, float4 *__restrict__ rowA,
                                                  int *__restrict__ rowB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, img = blockIdx.z;
    if (i >= d.R) return;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = view0 + q;
    const float s = ang[2 * v], c = ang[2 * v + 1];
    const int64_t P = (int64_t)d.R * d.ZP;
    const float y = (float)ldf(proj, (int64_t)img * P + (int64_t)j * d.R + i);
    float w = 0.0f;
    if (y != 0.0f) {
        FastWalk<1> fw;
        if (fw.init(d, s, c, i, j, src) && !(fw.len < d.vz * 1e-6f)) w = fw.dist * y;
    }
    pixw[(int64_t)img * P + (int64_t)j * d.R + i] = w;
    // the detector pixel and source, as ray_setup forms them
    const float x2x = d.sx * c - d.sy * s, x2y = d.sx * s + d.sy * c;
    if (j == 0 && src == 0) {
        const float u = d.ymin + (float)i * d.du;
        const float x1x = d.dod * c - u * s, x1y = d.dod * s + u * c;
        float kx, bx0, lx, hx, ky, by0, ly, hy;
        int ix0, iy0;
        axis_bounds(d.vx, d.bx, d.M, x1x, x2x, kx, bx0, lx, hx, ix0, d.ivx);
        axis_bounds(d.vy, d.by, d.N, x1y, x2y, ky, by0, ly, hy, iy0, d.ivy);
        AdjCol e;
        e.a = make_float4(kx, bx0, ky, by0);
        e.b = make_float4(hmax(lx, ly), hmin(hx, hy), __int_as_float(ix0), __int_as_float(iy0));
        colT[(int64_t)q * d.R + i] = e;
    }
    if (i == 0) {
        const float x1z = d.zmin + (float)j * d.dv;
        const float x2z = d.tb ? (float)(d.sz0 + (float)src * d.sdz) : d.sz0;
        float kz, bz0, lz, hz;
        int iz0;
        axis_bounds(d.vz, d.bz, d.Z, x1z, x2z, kz, bz0, lz, hz, iz0, d.ivz);
        rowA[(int64_t)img * d.ZP + j] = make_float4(kz, bz0, lz, hz);
        rowB[(int64_t)img * d.ZP + j] = iz0;
    }
}

// slab interval [in, out] of voxel index n along one axis for a ray
__device__ __forceinline__ void slab(float k, float b0, int i0, int n, float &in, float &out) {
    if (i0 >= 0) {   // parallel: inside slab i0 for all alpha, else never
        const bool hit = n == i0;
        in = hit ? -__builtin_inff() : __builtin_inff();
        out = hit ? __builtin_inff() : -__builtin_inff();
    } else {
        const float a0 = fmaf((float)n, k, b0), a1 = fmaf((float)(n + 1), k, b0);
        in = hmin(a0, a1);
        out = hmax(a0, a1);
    }
}

constexpr int AZR = 8;          // z-run per lane
constexpr int ALW = 80;         // staged footprint columns
constexpr int ALP = ALW + 8;    // row pitch (zero columns for the NC overrun)
constexpr int ALH = 48;         // staged footprint rows (48 sources 3 mm apart need 36)
constexpr int ANC = 8;          // bounding-box columns per voxel

template <int NC>
__device__ __forceinline__ float adj_rows(const float *__restrict__ sw, const float4 *__restrict__ sra,
                                          const int *__restrict__ srb, int z0, int z1, int cbase, int iz,
                                          const float *xin, const float *xout) {
    float acc = 0.0f;
    for (int j = z0; j <= z1; ++j) {
        const float4 ra = sra[j];
        const float *prow = sw + j * ALP + cbase;
        float pw[NC];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) pw[cc] = prow[cc];
        float zi, zo;
        slab(ra.x, ra.y, srb[j], iz, zi, zo);
        zi = hmax(zi, ra.z);
        zo = hmin(zo, ra.w);
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            const float seg = hsub_sat(hmin(xout[cc], zo), hmax(xin[cc], zi));   // alpha lengths are below 1
            acc = fmaf(seg, pw[cc], acc);
        }
    }
    return acc;
}

// one lane per (x, y) and its AZR z-voxels; a 16x16 (x, y) tile per
// workgroup; per image the tile's detector footprint of weights is staged in
// LDS.  Each view's per-voxel sum (FP32, ~10 terms) is added into an FP32
// total with Kahan compensation; the volume is written once.
template <typename SO>
__global__ __launch_bounds__(256) void k_adjoint_gather(DGeo<float> d, const float *__restrict__ ang, int view0,
                                                        int nimg, const float *__restrict__ pixw,
                                                        const AdjCol *__restrict__ colT,
                                                        const float4 *__restrict__ rowA, const int *__restrict__ rowB,
                                                        SO *__restrict__ vol) {
    __shared__ float sw[ALH * ALP];
    __shared__ AdjCol sc[ALP];
    __shared__ float4 sra[ALH];
    __shared__ int srb[ALH];
    const int tid = threadIdx.y * 16 + threadIdx.x;
    const int3 tile = xcd_tile();
    const int ix = tile.x * 16 + threadIdx.x, iy = tile.y * 16 + threadIdx.y, iz0 = tile.z * AZR;
    const int nz = min(AZR, d.Z - iz0);
    const bool active = ix < d.M && iy < d.N;
    float cx = 0, cy = 0;
    if (!d.tb) {
        cx = d.bx + d.vx * 0.5f + (float)ix * d.vx;
        cy = d.by + d.vy * 0.5f + (float)iy * d.vy;
    } else {
        cx = d.bx + d.vx * (0.5f + (float)ix);
        cy = d.by + d.vy * (0.5f + (float)iy);
    }
    const float hvx = 0.5f * d.vx, hvy = 0.5f * d.vy, hvz = 0.5f * d.vz;
    const float xl = cx - hvx, xh = cx + hvx, yl = cy - hvy, yh = cy + hvy;
    const int tx0 = tile.x * 16, ty0 = tile.y * 16;
    const float TX0 = d.bx + (float)tx0 * d.vx, TX1 = d.bx + (float)min(tx0 + 16, d.M) * d.vx;
    const float TY0 = d.by + (float)ty0 * d.vy, TY1 = d.by + (float)min(ty0 + 16, d.N) * d.vy;
    const float TZ0 = d.bz + (float)iz0 * d.vz, TZ1 = d.bz + (float)(iz0 + nz) * d.vz;
    // per-voxel sums over the images, Kahan-compensated FP32
    float W[AZR], Wc[AZR];
#pragma unroll
    for (int k = 0; k < AZR; ++k) { W[k] = 0.0f; Wc[k] = 0.0f; }
    const int64_t P = (int64_t)d.R * d.ZP;
    for (int img = 0; img < nimg; ++img) {
        const int q = img / d.NS, src = img - q * d.NS;
        const int v = view0 + q;
        const float s = ang[2 * v], c = ang[2 * v + 1];
        const float szs = d.tb ? (float)(d.sz0 + (float)src * d.sdz) : d.sz0;
        // this lane's voxel column: corner ratios and detector-column box
        float rmin = __builtin_inff(), rmax = -__builtin_inff(), minY = __builtin_inff(), maxY = -__builtin_inff();
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                const float px = ii ? xh : xl, py = jj ? yh : yl;
                const float t = px * c + py * s;
                const float sv = -px * s + py * c;
                const float ratio = (t - d.dod) * frcp(d.dso - t);
                const float u = fmaf(ratio, sv - d.sy, sv);
                minY = fminf(minY, u); maxY = fmaxf(maxY, u);
                rmin = fminf(rmin, ratio); rmax = fmaxf(rmax, ratio);
            }
        const int y0r = shadow_lo((minY - d.ymin) * d.idu);
        const int y1r = shadow_hi((maxY - d.ymin) * d.idu);
        // the tile's footprint
        float Tumin = __builtin_inff(), Tumax = -__builtin_inff(), Trmin = __builtin_inff(), Trmax = -__builtin_inff();
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                const float px = ii ? TX1 : TX0, py = jj ? TY1 : TY0;
                const float t = px * c + py * s;
                const float sv = -px * s + py * c;
                const float ratio = (t - d.dod) * frcp(d.dso - t);
                const float u = fmaf(ratio, sv - d.sy, sv);
                Tumin = fminf(Tumin, u); Tumax = fmaxf(Tumax, u);
                Trmin = fminf(Trmin, ratio); Trmax = fmaxf(Trmax, ratio);
            }
        const int C0 = max((int)floorf(fmaf(Tumin - d.ymin, d.idu, 0.5f)) - 1, 0);
        const int C1 = min(min((int)floorf(fmaf(Tumax - d.ymin, d.idu, 0.5f)) + 1, d.R - 1), C0 + ALW - 1);
        const int ncw = C1 - C0 + 1;
        const float e0 = TZ0 - szs, e1 = TZ1 - szs;
        const float Tw0 = fminf(fmaf(Trmin, e0, TZ0), fmaf(Trmax, e0, TZ0));
        const float Tw1 = fmaxf(fmaf(Trmin, e1, TZ1), fmaf(Trmax, e1, TZ1));
        const int R0 = max((int)floorf(fmaf(Tw0 - d.zmin, d.idv, 0.5f)) - 1, 0);
        const int R1 = min(min((int)floorf(fmaf(Tw1 - d.zmin, d.idv, 0.5f)) + 1, d.ZP - 1), R0 + ALH - 1);
        const int nrow = R1 - R0 + 1;
        if (ncw <= 0 || nrow <= 0) continue;   // uniform
        __syncthreads();
        {
            const float *g = pixw + (int64_t)img * P;
            for (int i = tid; i < nrow * ALP; i += 256) {
                const int r = i / ALP, cl = i - r * ALP;
                sw[i] = cl < ncw ? g[(int64_t)(R0 + r) * d.R + C0 + cl] : 0.0f;
            }
            if (tid < ALP) {
                AdjCol e;
                if (tid < ncw) e = colT[(int64_t)q * d.R + C0 + tid];
                else { e.a = make_float4(0, 0, 0, 0); e.b = make_float4(0, 0, 0, 0); }
                sc[tid] = e;
            }
            if (tid < nrow) {
                sra[tid] = rowA[(int64_t)img * d.ZP + R0 + tid];
                srb[tid] = rowB[(int64_t)img * d.ZP + R0 + tid];
            }
        }
        __syncthreads();
        if (!active) continue;
        if ((y0r < 0 && y1r < 0) || (y0r > d.R - 1 && y1r > d.R - 1)) continue;
        const int y0 = max(y0r, C0), y1 = min(min(y1r, C1), y0 + ANC - 1);
        const int ncol = y1 - y0 + 1;
        if (ncol <= 0) continue;
        const int cbase = y0 - C0;
        int NC;
        if (__builtin_amdgcn_ballot_w64(ncol > 6)) NC = 8;
        else if (__builtin_amdgcn_ballot_w64(ncol > 4)) NC = 6;
        else if (__builtin_amdgcn_ballot_w64(ncol > 3)) NC = 4;
        else NC = 3;
        // x/y slab interval of this voxel column for each footprint column
        float xin[ANC], xout[ANC];
#pragma unroll
        for (int cc = 0; cc < ANC; ++cc) {
            const AdjCol e = sc[min(cbase + cc, ALP - 1)];
            float ai, ao, bi, bo;
            slab(e.a.x, e.a.y, __float_as_int(e.b.z), ix, ai, ao);
            slab(e.a.z, e.a.w, __float_as_int(e.b.w), iy, bi, bo);
            const bool in = cc < ncol;
            xin[cc] = in ? hmax(hmax(ai, bi), e.b.x) : __builtin_inff();
            xout[cc] = in ? hmin(hmin(ao, bo), e.b.y) : -__builtin_inff();
        }
#pragma unroll
        for (int k = 0; k < AZR; ++k) {
            if (k >= nz) break;
            const int iz = iz0 + k;
            const float cz = !d.tb ? d.bz + d.vz * 0.5f + (float)iz * d.vz : d.bz + d.vz * (0.5f + (float)iz);
            const float zl = cz - hvz, zh = cz + hvz;
            const float el = zl - szs, eh = zh - szs;
            const float wl = fmaf(el >= 0.0f ? rmin : rmax, el, zl);
            const float wh = fmaf(eh >= 0.0f ? rmax : rmin, eh, zh);
            int z0 = shadow_lo((wl - d.zmin) * d.idv);
            int z1 = shadow_hi((wh - d.zmin) * d.idv);
            if ((z0 < 0 && z1 < 0) || (z0 > d.ZP - 1 && z1 > d.ZP - 1)) continue;
            z0 = max(z0, R0) - R0;
            z1 = min(z1, R1) - R0;
            float a;
            if (NC == 8) a = adj_rows<8>(sw, sra, srb, z0, z1, cbase, iz, xin, xout);
            else if (NC == 6) a = adj_rows<6>(sw, sra, srb, z0, z1, cbase, iz, xin, xout);
            else if (NC == 4) a = adj_rows<4>(sw, sra, srb, z0, z1, cbase, iz, xin, xout);
            else a = adj_rows<3>(sw, sra, srb, z0, z1, cbase, iz, xin, xout);
            const float y = a - Wc[k], t = W[k] + y;
            Wc[k] = (t - W[k]) - y;
            W[k] = t;
        }
    }
    if (!active) return;
#pragma unroll
    for (int k = 0; k < AZR; ++k) {
        if (k >= nz) break;
        St<SO>::st(vol, ((int64_t)(iz0 + k) * d.N + iy) * d.M + ix, W[k] - Wc[k]);
    }
}

// Host bound for k_adjoint_gather: every tile footprint fits ALW x ALH and
// every voxel box ANC columns (the bound of bp_lds_fits without the
// reconstruction cylinder: the adjoint covers every voxel).
static bool adj_gather_fits(const DGeo<float> &d) {
    if (d.style != CT_CONE) return false; // the column/row tables assume one source per view
    const double vx = d.vx, vy = d.vy, vz = d.vz, dso = d.dso, dod = d.dod;
    const double xs[2] = {(double)d.bx, (double)d.bx + d.M * vx}, ys[2] = {(double)d.by, (double)d.by + d.N * vy};
    double rho = 0;
    for (double x : xs)
        for (double y : ys) rho = std::max(rho, std::sqrt(x * x + y * y));
    if (!(dso - rho > 0)) return false;
    const double r_lo = (-rho - dod) / (dso + rho), r_hi = (rho - dod) / (dso - rho);
    const double rmax = std::max(std::fabs(r_lo), std::fabs(r_hi));
    const double dr = (dso - dod) / ((dso - rho) * (dso - rho));
    double zs_lo = d.sz0, zs_hi = d.sz0 + (double)(d.NS - 1) * d.sdz;
    if (zs_lo > zs_hi) std::swap(zs_lo, zs_hi);
    const double z_lo = d.bz, z_hi = (double)d.bz + d.Z * vz;
    const double zrel = std::max({std::fabs(z_lo - zs_lo), std::fabs(z_lo - zs_hi), std::fabs(z_hi - zs_lo),
                                  std::fabs(z_hi - zs_hi)});
    const double lat = rho + std::fabs((double)d.sy);
    auto cols = [&](double D) { return ((1.0 + rmax) * D + dr * D * lat) / (double)d.du; };
    auto rows = [&](double D, double H) { return ((1.0 + rmax) * H + dr * D * zrel) / (double)d.dv; };
    const double DT = std::hypot(16 * vx, 16 * vy), DV = std::hypot(vx, vy);
    return cols(DT) + 4.0 <= ALW && rows(DT, AZR * vz) + 4.0 <= ALH && cols(DV) + 2.0 <= ANC;
}

// Host conditions of k_project_col; returns the staged window NQ (quads per
// segment) or 0.  The list and two staging buffers fit the workgroup's LDS,
// the z-quad volume is under 2 GiB, and no list segment holds two z
// crossings: along every ray |dz|/vz < max(|dx|/vx, |dy|/vy), checked with
// |dz| at its largest (rows against the farthest source height) and the
// in-plane length at its smallest (|dso - dod|), with margin.  The window
// covers the z range of a workgroup's rows at any alpha inside the volume:
// for cone beam the rows fan out from the source, so their spread at a point
// is (1 - alpha) times the spread on the detector, and 1 - alpha (the
// point's share of the source-to-pixel distance) is at most
// (r_xy + |S|) / |dso - dod| with r_xy the volume's largest in-plane radius;
// fan and parallel rays keep their row height.  One voxel of the segment's
// own z travel, one of rounding margin on each side and the quad alignment
// are added.
static int col_walk_nq(const DGeo<float> &d) {
    if ((uint64_t)d.M * d.N * (uint64_t)((d.Z + 3) & ~3) * 4 >= (1ull << 31)) return 0;
    if ((d.ZP + COL_ROWS - 1) / COL_ROWS > 65535) return 0;
    double dzmax = 0.0;
    const double inplane = std::fabs((double)d.dso - (double)d.dod);
    const int rows = std::min(COL_ROWS, d.ZP) - 1;
    double span;
    if (d.style == CT_CONE) {
        const double s0 = d.sz0, s1 = d.tb ? d.sz0 + (double)(d.NS - 1) * d.sdz : d.sz0;
        dzmax = std::max(std::max(std::fabs(d.zmin - s0), std::fabs(d.zmax - s0)),
                         std::max(std::fabs(d.zmin - s1), std::fabs(d.zmax - s1)));
        const double X0 = d.bx, X1 = d.bx + (double)d.M * d.vx, Y0 = d.by, Y1 = d.by + (double)d.N * d.vy;
        const double rxy = std::sqrt(std::max(X0 * X0, X1 * X1) + std::max(Y0 * Y0, Y1 * Y1));
        const double rho = (rxy + std::hypot((double)d.dso, (double)d.sy)) / inplane;
        span = rho * rows * (double)d.dv / (double)d.vz + 1.0;
    } else {
        span = rows * (double)d.dv / (double)d.vz;
    }
    const double vxy = std::max((double)d.vx, (double)d.vy);
    if (!(dzmax / (double)d.vz < 0.9 * inplane / (std::sqrt(2.0) * vxy))) return 0;
    const double nq = std::ceil((span + 8.0) / 4.0);
    if (!(nq >= 1.0) || nq > COL_NQ) return 0;
    // (the padding of col_dyn_lds keeps a workgroup under 33 KiB)
    if (col_lds_bytes(d.M, d.N) > 65536) return 0;
    return (int)nq;
}

// Dynamic LDS of a launch: the list, padded so that a workgroup holds at
// least col_lds_floor bytes in all, which caps the workgroups per CU.
// Measured per view at 512^3 / 1024^2 (a one-view launch is 4096
// workgroups): FP32 volumes 0.457 ms at four per CU against 0.464 at five and
// 0.485 at six; FP16 0.331 ms at four against 0.376 at seven; BF16 (windows
// staged as FP32) 0.378 ms at six against 0.436 at four.
template <typename SV> constexpr size_t col_lds_floor() {
    return std::is_same<SV, __hip_bfloat16>::value ? 0 : 32 * 1024 + 512;   // > 160 KiB / 5
}
template <typename SV> static size_t col_dyn_lds(int M, int N) {
    const size_t stat = sizeof(typename StageOf<SV>::type) * 2 * col_batch<SV>() * COL_W + 16;
    const size_t want = col_lds_floor<SV>();
    const size_t list = col_dyn_bytes(M, N);
    return stat + list >= want ? list : want - stat;
}
template <typename SV, typename SO, typename SM, int CMD, bool TR = false>
static void launch_col(ct_context *ctx, const DGeo<float> &d, int nq, const float *ang, const void *volq, void *out,
                       const void *meas, int first, int stride, int mv0, int nimg) {
    for (int z0 = 0; z0 < nimg; z0 += 65535) {
        const int nz = nimg - z0 < 65535 ? nimg - z0 : 65535;
        dim3 grid(d.R, (d.ZP + COL_ROWS - 1) / COL_ROWS, nz);
        hipLaunchKernelGGL((k_project_col<SV, SO, SM, CMD, TR>), grid, dim3(COL_ROWS), col_dyn_lds<SV>(d.M, d.N),
                           ctx->stream, d, ang, (const SV *)volq, (SO *)out, (const SM *)meas, first, stride, mv0,
                           z0, nq);
    }
}

// The column walk stages, per segment, one z-quad of each (x, y) on the
// column's path; a 128-byte line holds eight x-neighbours, so paths that run
// along x use whole lines and paths along y one quad per line.  Measured at
// 512^3 / 1024^2 per view: 0.39 vs 0.46 ms (per-ray walk) along x, even near
// 40 degrees, 0.67 vs 0.44 ms along y.  Views whose central ray is within
// ~41 degrees of the x axis (|cos| >= 0.75) take the column walk; by_run
// splits a launch's views into runs of one choice.
static bool col_walk_view(const ct_geometry &g, int v) { return std::fabs(view_cos(g, v)) >= 0.75f; }
template <typename F> static void by_run(const ct_geometry &g, bool allow, int first, int stride, int nv, F &&launch) {
    int q0 = 0;
    bool cur = allow && nv > 0 && col_walk_view(g, first);
    for (int q = 1; q < nv; ++q) {
        const bool cw = allow && col_walk_view(g, first + q * stride);
        if (cw != cur) { launch(q0, q - q0, cur); q0 = q; cur = cw; }
    }
    if (nv > 0) launch(q0, nv - q0, cur);
}

// ---------------------------------------------------------------- launchers
template <typename T, typename SV, typename SO, typename SM, int CMD>
static ct_status run_forward(ct_context *ctx, const ct_geometry &g, const void *vol, void *out, const void *meas,
                             int first, int stride, int nv, int mv0) {
    ct_status st;
    const T *ang = (const T *)angle_table(ctx, g, sizeof(T) == 8, false, &st);
    if (!ang) return st;
    DGeo<T> d = derive<T>(g);
    const int nimg = nv * d.NS;
    dim3 block(256);
    // compensated FP32 ray sums in fast mode; the exact mode keeps the
    // reference's plain accumulation
    const bool comp = std::is_same<T, float>::value && ctx->math == CT_MATH_FAST;
    if constexpr (std::is_same<T, float>::value) {
        // four or more images: copy the volume once into the z-quad layout
        // and walk by detector column (k_project_col) where the geometry
        // admits it; otherwise one walk per ray for ray sums and each ray
        // split in two for corrections
        const uint64_t qbytes = (uint64_t)d.M * d.N * (uint64_t)((d.Z + 3) & ~3) * sizeof(SV);
        if (comp && nimg >= 4 && qbytes < (1ull << 32)) {
            void *volq = slot(ctx, SLOT_ZQVOL, qbytes, &st);
            if (!volq) return st;
            ProfScope ps(ctx, CT_K_FORWARD);   // the copy is part of the call's cost
            st = launch_zq_pack(ctx, vol, dtype_of<SV>(), volq, d.M, d.N, d.Z);
            if (st != CT_OK) return st;
            const int nq = col_walk_nq(d);
            const int64_t P = (int64_t)d.R * d.ZP;
            by_run(g, nq > 0, first, stride, nv, [&](int q0, int n, bool cw) {
                SO *o = (SO *)out + (int64_t)q0 * d.NS * P;
                const int f0 = first + q0 * stride, ni = n * d.NS;
                if (cw) {
                    launch_col<SV, SO, SM, CMD>(ctx, d, nq, ang, volq, o, meas, f0, stride, mv0, ni);
                    return;
                }
                constexpr int S = CMD == 1 ? CORR_SPLIT : 1;
                constexpr int tw = S <= 2 ? 16 : 8, th = S == 1 ? 16 : 8;
                for (int z0 = 0; z0 < ni; z0 += 65535) {
                    const int nz = ni - z0 < 65535 ? ni - z0 : 65535;
                    dim3 grid((d.R + tw - 1) / tw, (d.ZP + th - 1) / th, nz);
                    hipLaunchKernelGGL((k_project_zq<SV, SO, SM, CMD, S>), grid, dim3(256), 0, ctx->stream, d, ang,
                                       (const SV *)volq, o, (const SM *)meas, f0, stride, mv0, z0);
                }
            });
            return launch_check(ctx, "forward");
        }
    }
    ProfScope ps(ctx, CT_K_FORWARD);
    for (int z0 = 0; z0 < nimg; z0 += 65535) {
        int nz = nimg - z0 < 65535 ? nimg - z0 : 65535;
        dim3 grid((d.R + 15) / 16, (d.ZP + 15) / 16, nz);
        if (comp)
            hipLaunchKernelGGL((k_forward<T, SV, SO, SM, CMD, true>), grid, block, 0, ctx->stream, d, ang,
                               (const SV *)vol, (SO *)out, (const SM *)meas, first, stride, mv0, z0);
        else
            hipLaunchKernelGGL((k_forward<T, SV, SO, SM, CMD, false>), grid, block, 0, ctx->stream, d, ang,
                               (const SV *)vol, (SO *)out, (const SM *)meas, first, stride, mv0, z0);
    }
    return launch_check(ctx, "forward");
}

template <typename T, typename SV, typename SO, int CMD>
static ct_status fwd_meas(ct_context *ctx, const ct_geometry &g, const void *vol, void *out, const void *meas,
                          ct_dtype mdt, int first, int stride, int nv, int mv0) {
    if constexpr (CMD == 0) return run_forward<T, SV, SO, SO, CMD>(ctx, g, vol, out, meas, first, stride, nv, mv0);
    switch (mdt) {
    case CT_F32: return run_forward<T, SV, SO, float, CMD>(ctx, g, vol, out, meas, first, stride, nv, mv0);
    case CT_F16: return run_forward<T, SV, SO, __half, CMD>(ctx, g, vol, out, meas, first, stride, nv, mv0);
    case CT_BF16: return run_forward<T, SV, SO, __hip_bfloat16, CMD>(ctx, g, vol, out, meas, first, stride, nv, mv0);
    case CT_F64: return run_forward<T, SV, SO, double, CMD>(ctx, g, vol, out, meas, first, stride, nv, mv0);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported measurement dtype");
}

template <typename T, typename SV, int CMD>
static ct_status fwd_out(ct_context *ctx, const ct_geometry &g, const void *vol, void *out, ct_dtype odt,
                         const void *meas, ct_dtype mdt, int first, int stride, int nv, int mv0) {
    switch (odt) {
    case CT_F32: return fwd_meas<T, SV, float, CMD>(ctx, g, vol, out, meas, mdt, first, stride, nv, mv0);
    case CT_F16: return fwd_meas<T, SV, __half, CMD>(ctx, g, vol, out, meas, mdt, first, stride, nv, mv0);
    case CT_BF16: return fwd_meas<T, SV, __hip_bfloat16, CMD>(ctx, g, vol, out, meas, mdt, first, stride, nv, mv0);
    case CT_F64: return fwd_meas<T, SV, double, CMD>(ctx, g, vol, out, meas, mdt, first, stride, nv, mv0);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported output dtype");
}

// first/stride select the batch's views; for CMD 1 meas is indexed by view - mv0
// (the sweep passes the whole stack with mv0 = 0).
ct_status launch_forward_mv(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, void *out,
                            ct_dtype odt, const void *meas, ct_dtype mdt, int32_t first, int32_t stride,
                            int32_t nv, int cmd, int32_t mv0) {
    if (nv <= 0) return CT_OK;
    if (vdt == CT_F64) {
        if (cmd == 0) return fwd_out<double, double, 0>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0);
        return fwd_out<double, double, 1>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0);
    }
    switch (vdt) {
    case CT_F32:
        return cmd == 0 ? fwd_out<float, float, 0>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0)
                        : fwd_out<float, float, 1>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0);
    case CT_F16:
        return cmd == 0 ? fwd_out<float, __half, 0>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0)
                        : fwd_out<float, __half, 1>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0);
    case CT_BF16:
        return cmd == 0 ? fwd_out<float, __hip_bfloat16, 0>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0)
                        : fwd_out<float, __hip_bfloat16, 1>(ctx, g, vol, out, odt, meas, mdt, first, stride, nv, mv0);
    default: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported volume dtype");
}

// ---------------------------------------------------------------------------
// Allocation-free forward projection (ct_forward_noalloc).  Its only scratch
// is the per-view (sin, cos) table the kernels read, which ct_forward keeps
// in a context buffer allocated on first use.  Here the table lives in the
// file-scope array `scratch` (256 KiB: up to 16384 views, in FP64 or FP32),
// one copy per GPU, held until the process ends.  Its device address
// is looked up once per device at context creation (noalloc_init) and only
// that address is passed to kernels and copies.  Calls that overlap on one
// GPU (two streams or two contexts) race on the table; a call first waits for
// its own stream, so back-to-back calls on one stream are ordered.
constexpr int NOALLOC_MAX_VIEWS = 16384;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_VIEWS];

static std::mutex g_noalloc_mtx;
static void *g_noalloc_ptr[64] = {};

ct_status noalloc_init(ct_context *ctx) {
    if (ctx->device < 0 || ctx->device >= 64) return fail(ctx, CT_ERR_DEVICE, "device index beyond 63");
    ctx->noalloc_host.assign(sizeof(scratch), 0);   // the host side of the table, allocated here once
    std::lock_guard<std::mutex> l(g_noalloc_mtx);
    if (g_noalloc_ptr[ctx->device]) return CT_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        hipGetLastError();
        return fail(ctx, CT_ERR_DEVICE, std::string("scratch symbol lookup failed: ") + hipGetErrorString(e));
    }
    g_noalloc_ptr[ctx->device] = p;
    return CT_OK;
}

// The per-ray forward walk of run_forward (no z-quad copy) on a caller-given
// angle table.
template <typename T, typename SV, typename SO>
static ct_status run_forward_table(ct_context *ctx, const ct_geometry &g, const T *ang, const void *vol, void *out,
                                   int first, int nv) {
    DGeo<T> d = derive<T>(g);
    const int nimg = nv * d.NS;
    const bool comp = std::is_same<T, float>::value && ctx->math == CT_MATH_FAST;
    for (int z0 = 0; z0 < nimg; z0 += 65535) {
        const int nz = nimg - z0 < 65535 ? nimg - z0 : 65535;
        dim3 grid((d.R + 15) / 16, (d.ZP + 15) / 16, nz);
        if (comp)
            hipLaunchKernelGGL((k_forward<T, SV, SO, float, 0, true>), grid, dim3(256), 0, ctx->stream, d, ang,
                               (const SV *)vol, (SO *)out, (const float *)nullptr, first, 1, first, z0);
        else
            hipLaunchKernelGGL((k_forward<T, SV, SO, float, 0, false>), grid, dim3(256), 0, ctx->stream, d, ang,
                               (const SV *)vol, (SO *)out, (const float *)nullptr, first, 1, first, z0);
    }
    return launch_check(ctx, "forward (allocation-free)");
}

template <typename T, typename SV>
static ct_status noalloc_out(ct_context *ctx, const ct_geometry &g, const T *ang, const void *vol, void *out,
                             ct_dtype odt, int first, int nv) {
    switch (odt) {
    case CT_F32: return run_forward_table<T, SV, float>(ctx, g, ang, vol, out, first, nv);
    case CT_F16: return run_forward_table<T, SV, __half>(ctx, g, ang, vol, out, first, nv);
    case CT_BF16: return run_forward_table<T, SV, __hip_bfloat16>(ctx, g, ang, vol, out, first, nv);
    case CT_F64: return run_forward_table<T, SV, double>(ctx, g, ang, vol, out, first, nv);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported output dtype");
}

ct_status launch_forward_noalloc(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, void *out,
                                 ct_dtype odt, int32_t view0, int32_t nv) {
    if (g.nviews > NOALLOC_MAX_VIEWS)
        return fail(ctx, CT_ERR_INVALID_ARG, "allocation-free forward: nviews " + std::to_string(g.nviews) +
                                                 " exceeds its limit of " + std::to_string(NOALLOC_MAX_VIEWS));
    if (nv <= 0) return CT_OK;
    void *tab = nullptr;
    {
        std::lock_guard<std::mutex> l(g_noalloc_mtx);
        if (ctx->device >= 0 && ctx->device < 64) tab = g_noalloc_ptr[ctx->device];
    }
    if (!tab) return fail(ctx, CT_ERR_DEVICE, "allocation-free forward: scratch not set up on this device");
    // the previous call on this stream may still read the table
    ct_status st = hip_check(ctx, hipStreamSynchronize(ctx->stream), "stream synchronize");
    if (st != CT_OK) return st;
    const bool dbl = vdt == CT_F64;
    const size_t bytes = (size_t)g.nviews * 2 * (dbl ? 8 : 4);
    // sized for NOALLOC_MAX_VIEWS at context creation (noalloc_init)
    if (ctx->noalloc_host.size() < bytes) return fail(ctx, CT_ERR_INVALID_ARG, "allocation-free forward: host table");
    fill_angle_table(g, dbl, false, ctx->noalloc_host.data());
    st = hip_check(ctx, hipMemcpyAsync(tab, ctx->noalloc_host.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                   "angle table upload");
    if (st != CT_OK) return st;
    switch (vdt) {
    case CT_F32: return noalloc_out<float, float>(ctx, g, (const float *)tab, vol, out, odt, view0, nv);
    case CT_F16: return noalloc_out<float, __half>(ctx, g, (const float *)tab, vol, out, odt, view0, nv);
    case CT_BF16: return noalloc_out<float, __hip_bfloat16>(ctx, g, (const float *)tab, vol, out, odt, view0, nv);
    case CT_F64: return noalloc_out<double, double>(ctx, g, (const double *)tab, vol, out, odt, view0, nv);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported volume dtype");
}

template <typename SV, typename SM>
static ct_status run_correction_zq(ct_context *ctx, const ct_geometry &g, const void *volq, float *corr,
                                   const void *meas, int first, int stride, int nv, bool tr) {
    ct_status st;
    const float *ang = (const float *)angle_table(ctx, g, false, false, &st);
    if (!ang) return st;
    DGeo<float> d = derive<float>(g);
    ProfScope ps(ctx, CT_K_FORWARD);
    const int nq = col_walk_nq(d);
    const int64_t P = (int64_t)d.R * d.ZP;
    if (tr) {
        // transposed layout (sweep_view_transposed held for the caller's
        // views): every view takes the column walk
        if (nq <= 0) return fail(ctx, CT_ERR_INVALID_ARG, "transposed correction without the column walk");
        launch_col<SV, float, SM, 1, true>(ctx, d, nq, ang, volq, corr, meas, first, stride, 0, nv * d.NS);
        return launch_check(ctx, "correction");
    }
    by_run(g, nq > 0, first, stride, nv, [&](int q0, int n, bool cw) {
        float *o = corr + (int64_t)q0 * d.NS * P;
        const int f0 = first + q0 * stride, ni = n * d.NS;
        if (cw) {
            launch_col<SV, float, SM, 1>(ctx, d, nq, ang, volq, o, meas, f0, stride, 0, ni);
            return;
        }
        for (int z0 = 0; z0 < ni; z0 += 65535) {
            const int nz = ni - z0 < 65535 ? ni - z0 : 65535;
            constexpr int tw = CORR_SPLIT <= 2 ? 16 : 8, th = CORR_SPLIT == 1 ? 16 : 8;
            dim3 grid((d.R + tw - 1) / tw, (d.ZP + th - 1) / th, nz);
            hipLaunchKernelGGL((k_project_zq<SV, float, SM, 1, CORR_SPLIT>), grid, dim3(256), 0, ctx->stream, d, ang,
                               (const SV *)volq, o, (const SM *)meas, f0, stride, 0, z0);
        }
    });
    return launch_check(ctx, "correction");
}

template <typename SV>
static ct_status corr_zq_meas(ct_context *ctx, const ct_geometry &g, const void *volq, float *corr, const void *meas,
                              ct_dtype mdt, int first, int stride, int nv, bool tr) {
    switch (mdt) {
    case CT_F32: return run_correction_zq<SV, float>(ctx, g, volq, corr, meas, first, stride, nv, tr);
    case CT_F16: return run_correction_zq<SV, __half>(ctx, g, volq, corr, meas, first, stride, nv, tr);
    case CT_BF16: return run_correction_zq<SV, __hip_bfloat16>(ctx, g, volq, corr, meas, first, stride, nv, tr);
    case CT_F64: return run_correction_zq<SV, double>(ctx, g, volq, corr, meas, first, stride, nv, tr);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported measurement dtype");
}

// The sweep's SART correction on a z-quad-layout volume (views first,
// first + stride, ...; the measurement stack is indexed by view).
ct_status launch_correction_zq(ct_context *ctx, const ct_geometry &g, const void *volq, ct_dtype vdt, float *corr,
                               const void *meas, ct_dtype mdt, int32_t first, int32_t stride, int32_t nv, bool tr) {
    if (nv <= 0) return CT_OK;
    switch (vdt) {
    case CT_F32: return corr_zq_meas<float>(ctx, g, volq, corr, meas, mdt, first, stride, nv, tr);
    case CT_F16: return corr_zq_meas<__half>(ctx, g, volq, corr, meas, mdt, first, stride, nv, tr);
    case CT_BF16: return corr_zq_meas<__hip_bfloat16>(ctx, g, volq, corr, meas, mdt, first, stride, nv, tr);
    default: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "z-quad correction: unsupported volume dtype");
}

// The sweep walks views whose rays run along y on the transposed z-quad
// layout, so that every view takes the column walk with line-sharing staging.
bool sweep_view_transposed(const ct_geometry &g, int v) {
    return !col_walk_view(g, v) && col_walk_nq(derive<float>(g)) > 0;
}

ct_status launch_forward(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, void *out,
                         ct_dtype odt, const void *meas, ct_dtype mdt, int32_t first, int32_t stride,
                         int32_t nv, int cmd) {
    return launch_forward_mv(ctx, g, vol, vdt, out, odt, meas, mdt, first, stride, nv, cmd, first);
}

template <typename T, typename SP, typename A>
static ct_status run_adjoint(ct_context *ctx, const ct_geometry &g, const void *proj, A *acc, int view0, int nv) {
    ct_status st;
    const T *ang = (const T *)angle_table(ctx, g, sizeof(T) == 8, false, &st);
    if (!ang) return st;
    DGeo<T> d = derive<T>(g);
    const int nimg = nv * d.NS;
    dim3 block(256);
    ProfScope ps(ctx, CT_K_ADJOINT);
    for (int z0 = 0; z0 < nimg; z0 += 65535) {
        int nz = nimg - z0 < 65535 ? nimg - z0 : 65535;
        dim3 grid((d.R + 15) / 16, (d.ZP + 15) / 16, nz);
        // the fast forward's walk in fast mode, so the pair stays matched; the
        // accumulator is then in the z-quad layout (neighbouring rays' atomics
        // share cache lines)
        if (std::is_same<T, float>::value && ctx->math == CT_MATH_FAST)
            hipLaunchKernelGGL((k_adjoint<T, SP, A, std::is_same<T, float>::value, true>), grid, block, 0,
                               ctx->stream, d, ang, (const SP *)proj, acc, view0, z0);
        else
            hipLaunchKernelGGL((k_adjoint<T, SP, A, false>), grid, block, 0, ctx->stream, d, ang, (const SP *)proj,
                               acc, view0, z0);
    }
    return launch_check(ctx, "adjoint");
}

template <typename SO>
static void launch_gather(ct_context *ctx, const DGeo<float> &d, const float *ang, int view0, int nimg,
                          const float *pixw, const AdjCol *colT, const float4 *rowA, const int *rowB, void *vol) {
    dim3 grid((d.M + 15) / 16, (d.N + 15) / 16, (d.Z + AZR - 1) / AZR);
    hipLaunchKernelGGL((k_adjoint_gather<SO>), grid, dim3(16, 16, 1), 0, ctx->stream, d, ang, view0, nimg, pixw, colT,
                       rowA, rowB, (SO *)vol);
}

static ct_status run_adjoint_gather(ct_context *ctx, const ct_geometry &g, const DGeo<float> &d, const void *proj,
                                    ct_dtype pdt, void *vol, ct_dtype vdt, int view0, int nv) {
    ct_status st;
    const float *ang = (const float *)angle_table(ctx, g, false, false, &st);
    if (!ang) return st;
    const int nimg = nv * d.NS;
    const size_t P = (size_t)d.R * d.ZP;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pw_b = up(sizeof(float) * P * nimg), col_b = up(sizeof(AdjCol) * (size_t)d.R * nv);
    const size_t ra_b = up(sizeof(float4) * (size_t)d.ZP * nimg), rb_b = up(sizeof(int) * (size_t)d.ZP * nimg);
    char *tab = (char *)slot(ctx, SLOT_TMP_VOL, pw_b + col_b + ra_b + rb_b, &st);
    if (!tab) return st;
    float *pixw = (float *)tab;
    AdjCol *colT = (AdjCol *)(tab + pw_b);
    float4 *rowA = (float4 *)(tab + pw_b + col_b);
    int *rowB = (int *)(tab + pw_b + col_b + ra_b);
    ProfScope ps(ctx, CT_K_ADJOINT);
    dim3 pg((d.R + 255) / 256, d.ZP, nimg);
    switch (pdt) {
    case CT_F32: hipLaunchKernelGGL((k_adj_prep<float>), pg, dim3(256), 0, ctx->stream, d, ang, view0, (const float *)proj, nimg, pixw, colT, rowA, rowB); break;
    case CT_F16: hipLaunchKernelGGL((k_adj_prep<__half>), pg, dim3(256), 0, ctx->stream, d, ang, view0, (const __half *)proj, nimg, pixw, colT, rowA, rowB); break;
    case CT_BF16: hipLaunchKernelGGL((k_adj_prep<__hip_bfloat16>), pg, dim3(256), 0, ctx->stream, d, ang, view0, (const __hip_bfloat16 *)proj, nimg, pixw, colT, rowA, rowB); break;
    default: return fail(ctx, CT_ERR_DTYPE, "unsupported projection dtype");
    }
    switch (vdt) {
    case CT_F32: launch_gather<float>(ctx, d, ang, view0, nimg, pixw, colT, rowA, rowB, vol); break;
    case CT_F16: launch_gather<__half>(ctx, d, ang, view0, nimg, pixw, colT, rowA, rowB, vol); break;
    case CT_BF16: launch_gather<__hip_bfloat16>(ctx, d, ang, view0, nimg, pixw, colT, rowA, rowB, vol); break;
    default: return fail(ctx, CT_ERR_DTYPE, "unsupported volume dtype");
    }
    return launch_check(ctx, "adjoint");
}

ct_status launch_adjoint(ct_context *ctx, const ct_geometry &g, const void *proj, ct_dtype pdt, void *vol,
                         ct_dtype vdt, int32_t view0, int32_t nv) {
    const int64_t V = (int64_t)g.nx * g.ny * g.nz;
    if (vdt == CT_F64) {
        if (pdt != CT_F64) return fail(ctx, CT_ERR_DTYPE, "F64 volume needs F64 projections");
        ct_status st = launch_fill(ctx, vol, CT_F64, V, 0.0);
        if (st != CT_OK) return st;
        return run_adjoint<double, double, double>(ctx, g, proj, (double *)vol, view0, nv);
    }
    if (pdt == CT_F64) return fail(ctx, CT_ERR_DTYPE, "F64 projections need an F64 volume");
    // FP32 segments scattered into an FP64 accumulator (f64 atomics): the
    // ~10^2 contributions per voxel are summed without FP32 rounding, then
    // rounded once to the output type.  Fast mode accumulates in the z-quad
    // layout and unpacks while converting.
    const bool zq = ctx->math == CT_MATH_FAST;
    if (zq) {
        // gather form: the walk's segments summed per voxel, no atomics
        const DGeo<float> d = derive<float>(g);
        const int nimg = nv * d.NS;
        if (nimg <= 65535 && adj_gather_fits(d)) return run_adjoint_gather(ctx, g, d, proj, pdt, vol, vdt, view0, nv);
    }
    const int64_t Va = zq ? (int64_t)g.nx * g.ny * ((g.nz + 3) & ~3) : V;
    ct_status st;
    double *acc = (double *)slot(ctx, SLOT_TMP_VOL, sizeof(double) * Va, &st);
    if (!acc) return st;
    st = launch_fill(ctx, acc, CT_F64, Va, 0.0);
    if (st != CT_OK) return st;
    switch (pdt) {
    case CT_F32: st = run_adjoint<float, float, double>(ctx, g, proj, acc, view0, nv); break;
    case CT_F16: st = run_adjoint<float, __half, double>(ctx, g, proj, acc, view0, nv); break;
    case CT_BF16: st = run_adjoint<float, __hip_bfloat16, double>(ctx, g, proj, acc, view0, nv); break;
    default: return fail(ctx, CT_ERR_DTYPE, "unsupported projection dtype");
    }
    if (st != CT_OK) return st;
    if (zq) return launch_zq_unpack_f64(ctx, acc, vol, vdt, g.nx, g.ny, g.nz);
    return launch_convert(ctx, acc, CT_F64, vol, vdt, V);
}

} // namespace ct
