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
#include <type_traits>

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
    const T x2x = d.sx * c - d.sy * s;
    const T x2y = d.sx * s + d.sy * c;
    const T x2z = d.tb ? (T)(d.sz0 + (T)src * d.sdz) : d.sz0;
    const T u = d.ymin + (T)col * d.du;
    const T x1x = d.dod * c - u * s;
    const T x1y = d.dod * s + u * c;
    const T x1z = d.zmin + (T)row * d.dv;
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
    T amax, ac, ax, ay, az, dist, x1x, x1y, x1z, ixd, iyd, izd;
    int nx, ny, nz, vi, vj, vk, sgx, sgy, sgz;
    bool hit;
};

template <typename T>
__device__ __forceinline__ RayState<T> ray_setup(const DGeo<T> &d, T s, T c, int col, int row, int src) {
    RayState<T> r;
    const int M = d.M, N = d.N, Z = d.Z;
    const T x2x = d.sx * c - d.sy * s;
    const T x2y = d.sx * s + d.sy * c;
    const T x2z = d.tb ? (T)(d.sz0 + (T)src * d.sdz) : d.sz0;
    const T u = d.ymin + (T)col * d.du;
    const T x1x = d.dod * c - u * s;
    const T x1y = d.dod * s + u * c;
    const T x1z = d.zmin + (T)row * d.dv;
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
template <int ES, bool ZQ = false> struct FastWalk {
    float kx, ky, kz, bx0, by0, bz0, sgxf, sgyf, sgzf, nfx, nfy, nfz, ax, ay, az, ac, amax, dist, len;
    int dix, diy, diz, dizb;
    unsigned idx;
    // returns false for an empty ray (the forward yields 0 for it)
    __device__ __forceinline__ bool init(const DGeo<float> &d, float s, float c, int col, int row, int src) {
        RayState<float> r = ray_setup(d, s, c, col, row, src);
        const unsigned M = d.M, N = d.N, Z = d.Z, MN = M * N;
        if (!(r.hit && (unsigned)r.vi < M && (unsigned)r.vj < N && (unsigned)r.vk < Z)) return false;
        kx = d.vx * r.ixd; ky = d.vy * r.iyd; kz = d.vz * r.izd;
        bx0 = (d.bx - r.x1x) * r.ixd; by0 = (d.by - r.x1y) * r.iyd; bz0 = (d.bz - r.x1z) * r.izd;
        sgxf = (float)r.sgx; sgyf = (float)r.sgy; sgzf = (float)r.sgz;
        nfx = (float)r.nx; nfy = (float)r.ny; nfz = (float)r.nz;
        if constexpr (ZQ) {
            dix = r.sgx * 4 * ES; diy = r.sgy * (int)M * 4 * ES;
            diz = r.sgz * ES; dizb = r.sgz * (4 * (int)MN - 3) * ES;
            idx = (((unsigned)(r.vk >> 2) * MN + (unsigned)r.vj * M + (unsigned)r.vi) * 4u + ((unsigned)r.vk & 3u)) *
                  (unsigned)ES;
        } else {
            dix = r.sgx * ES; diy = r.sgy * (int)M * ES; diz = r.sgz * (int)MN * ES; dizb = diz;
            idx = ((unsigned)r.vk * MN + (unsigned)r.vj * M + (unsigned)r.vi) * (unsigned)ES;
        }
        ax = r.ax; ay = r.ay; az = r.az; ac = r.ac; dist = r.dist;
        // end the walk at the grid's exit planes as this walk computes them,
        // so no step past the grid carries a nonzero length
        amax = r.amax;
        if (r.ax != __builtin_inff()) {
            const int n = r.sgx > 0 ? (int)M : 0;
            amax = hmin(amax, n == r.nx ? r.ax : fmaf((float)n, kx, bx0));
        }
        if (r.ay != __builtin_inff()) {
            const int n = r.sgy > 0 ? (int)N : 0;
            amax = hmin(amax, n == r.ny ? r.ay : fmaf((float)n, ky, by0));
        }
        if (r.az != __builtin_inff()) {
            const int n = r.sgz > 0 ? (int)Z : 0;
            amax = hmin(amax, n == r.nz ? r.az : fmaf((float)n, kz, bz0));
        }
        len = dist * (amax - ac);
        return true;
    }
    // one step: the segment's alpha length (times dist = its length in mm)
    // and the offset of the voxel it lies in.  alpha_min <= every pending
    // crossing, so once ac reaches alpha_max every later step has length 0;
    // the caller tests for the end of the ray once per batch (ended()).
    __device__ __forceinline__ float step(unsigned &id) {
        const float m = hmin(hmin3(ax, ay, az), amax);
        const float dl = m - ac;
        id = idx;
        const bool sx = ax == m, sy = ay == m, sz = az == m;
        // ZQ: the z plane about to be crossed bounds a quad
        const bool qb = ZQ && __builtin_amdgcn_fractf(nfz * 0.25f) == 0.0f;
        const float nx1 = nfx + sgxf, ny1 = nfy + sgyf, nz1 = nfz + sgzf;
        const float ax1 = fmaf(nx1, kx, bx0), ay1 = fmaf(ny1, ky, by0), az1 = fmaf(nz1, kz, bz0);
        nfx = sx ? nx1 : nfx; ax = sx ? ax1 : ax;
        nfy = sy ? ny1 : nfy; ay = sy ? ay1 : ay;
        nfz = sz ? nz1 : nfz; az = sz ? az1 : az;
        idx += (unsigned)((sx ? dix : 0) + (sy ? diy : 0) + (sz ? (qb ? dizb : diz) : 0));
        ac = m;
        return dl;
    }
    __device__ __forceinline__ bool ended() const { return !(amax - ac > loop_eps<float>()); }
};

constexpr int FAST_B = 16;  // steps per batch: their gathers are in flight together

// Fast-mode ray sum: batches of FAST_B steps whose loads are issued together;
// each batch is summed in FP32 and the batch sums accumulate in FP64.
// Volumes under 4 GiB are read with range-checked buffer loads on 32-bit
// byte offsets; larger ones use 64-bit addresses with an explicit range test.
template <typename SV, bool BUF, bool ZQ = false>
__device__ __forceinline__ float siddon_walk_fast(const DGeo<float> &d, float s, float c, int col, int row, int src,
                                                  const SV *__restrict__ f, float &len_out) {
    constexpr int ES = BUF ? (int)sizeof(SV) : 1;
    FastWalk<ES, ZQ> w;
    len_out = 0.0f;
    if (!w.init(d, s, c, col, row, src)) return 0.0f;
    // elements of the array (the z-quad layout pads z to a multiple of 4)
    const unsigned nvox = (unsigned)d.M * (unsigned)d.N * (ZQ ? (unsigned)((d.Z + 3) & ~3) : (unsigned)d.Z);
    // (for the 64-bit path the resource is built but never used)
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(f, BUF ? nvox * (unsigned)sizeof(SV) : 0u);
    // sum of alpha-length * value; times dist at the end
    double sum = 0.0;
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
        sum += (double)ps;
    } while (!w.ended() && guard > 0);
    len_out = w.len;
    return (float)(sum * (double)w.dist);
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
template <typename SV, typename SM>
__global__ __launch_bounds__(256) void k_correction_zq(DGeo<float> d, const float *__restrict__ ang,
                                                       const SV *__restrict__ fq, float *__restrict__ out,
                                                       const SM *__restrict__ meas, int first, int stride, int mv0,
                                                       int z0) {
    int col, row;
    const int slot = tile_coords(col, row);
    if (col >= d.R || row >= d.ZP) return;
    const int img = slot + z0;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = first + q * stride;
    const float s = ang[2 * v], c = ang[2 * v + 1];
    float len;
    const float sum = siddon_walk_fast<SV, true, true>(d, s, c, col, row, src, fq, len);
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t pix = (int64_t)row * d.R + col;
    float r = 0.0f;
    if (!(len < d.vz * 1e-6f)) {
        const float b = (float)ldf(meas, ((int64_t)(v - mv0) * d.NS + src) * P + pix);
        r = (b - sum) / len;
    }
    out[(int64_t)img * P + pix] = r;
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

template <typename SV, typename SM>
static ct_status run_correction_zq(ct_context *ctx, const ct_geometry &g, const void *volq, float *corr,
                                   const void *meas, int first, int stride, int nv) {
    ct_status st;
    const float *ang = (const float *)angle_table(ctx, g, false, false, &st);
    if (!ang) return st;
    DGeo<float> d = derive<float>(g);
    const int nimg = nv * d.NS;
    ProfScope ps(ctx, CT_K_FORWARD);
    for (int z0 = 0; z0 < nimg; z0 += 65535) {
        const int nz = nimg - z0 < 65535 ? nimg - z0 : 65535;
        dim3 grid((d.R + 15) / 16, (d.ZP + 15) / 16, nz);
        hipLaunchKernelGGL((k_correction_zq<SV, SM>), grid, dim3(256), 0, ctx->stream, d, ang, (const SV *)volq, corr,
                           (const SM *)meas, first, stride, 0, z0);
    }
    return launch_check(ctx, "correction");
}

template <typename SV>
static ct_status corr_zq_meas(ct_context *ctx, const ct_geometry &g, const void *volq, float *corr, const void *meas,
                              ct_dtype mdt, int first, int stride, int nv) {
    switch (mdt) {
    case CT_F32: return run_correction_zq<SV, float>(ctx, g, volq, corr, meas, first, stride, nv);
    case CT_F16: return run_correction_zq<SV, __half>(ctx, g, volq, corr, meas, first, stride, nv);
    case CT_BF16: return run_correction_zq<SV, __hip_bfloat16>(ctx, g, volq, corr, meas, first, stride, nv);
    case CT_F64: return run_correction_zq<SV, double>(ctx, g, volq, corr, meas, first, stride, nv);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported measurement dtype");
}

// The sweep's SART correction on a z-quad-layout volume (views first,
// first + stride, ...; the measurement stack is indexed by view).
ct_status launch_correction_zq(ct_context *ctx, const ct_geometry &g, const void *volq, ct_dtype vdt, float *corr,
                               const void *meas, ct_dtype mdt, int32_t first, int32_t stride, int32_t nv) {
    if (nv <= 0) return CT_OK;
    switch (vdt) {
    case CT_F32: return corr_zq_meas<float>(ctx, g, volq, corr, meas, mdt, first, stride, nv);
    case CT_F16: return corr_zq_meas<__half>(ctx, g, volq, corr, meas, mdt, first, stride, nv);
    case CT_BF16: return corr_zq_meas<__hip_bfloat16>(ctx, g, volq, corr, meas, mdt, first, stride, nv);
    default: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "z-quad correction: unsupported volume dtype");
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
