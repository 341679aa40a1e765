// voxel_bp.hip - voxel-driven SART backprojection for gfx950.
//
// One voxel per lane, a 64x4 (x,y) tile per workgroup, one z plane per
// workgroup row of the grid.  Each voxel loops over the views of a subset
// and the sources of a view, accumulating W = sum l*c and L = sum l in the
// arithmetic type, and updates the volume once per subset.
//
// The per-(voxel, pixel) chord is the one of backprj_ray_driven_3d_kernel
// (V63/kernel_IterativeRecon_universal.cu:571-810; multi-source with source
// culling: TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:546-763).  The
// reference evaluates 6 divisions and one sqrt per bounding-box pixel; the
// x/y slab crossings depend only on the pixel column and the z crossings
// only on the row, so they are evaluated once per column (cached in
// registers) and once per row.  max/min are exact and the sums keep the
// reference's row-major order, so the result is bit-identical to the
// per-pixel form.
#include "ct_internal.hpp"
#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ct {

template <typename T> __device__ __forceinline__ T vfloor(T x);
template <> __device__ __forceinline__ float vfloor<float>(float x) { return floorf(x); }
template <> __device__ __forceinline__ double vfloor<double>(double x) { return floor(x); }
template <typename T> __device__ __forceinline__ T vmin(T a, T b);
template <> __device__ __forceinline__ float vmin<float>(float a, float b) { return fminf(a, b); }
template <> __device__ __forceinline__ double vmin<double>(double a, double b) { return fmin(a, b); }
template <typename T> __device__ __forceinline__ T vmax(T a, T b);
template <> __device__ __forceinline__ float vmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double vmax<double>(double a, double b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T vsqrt(T a);
template <> __device__ __forceinline__ float vsqrt<float>(float a) { return sqrtf(a); }
template <> __device__ __forceinline__ double vsqrt<double>(double a) { return sqrt(a); }

constexpr int MAXC = 8;   // bounding-box columns cached in registers

template <typename T>
__device__ __forceinline__ void voxel_wl(const DGeo<T> &d, T cx, T cy, T cz, T s, T c, T szs,
                                         const T *__restrict__ corr, T &W, T &L) {
    if (d.tb) {
        // source culling, TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:618-627
        const T tc = cx * c + cy * s;
        const T upper = (cz + d.vz - d.zmin) * (d.dso - d.dod) / (tc - d.dod) + d.zmin;
        const T lower = (cz - d.vz - d.zmax) * (d.dso - d.dod) / (tc - d.dod) + d.zmax;
        if (szs < lower || szs > upper) return;
    }
    // corner projections (:639-678): the ratio depends on (x,y) only
    const T hvx = (T)0.5 * d.vx, hvy = (T)0.5 * d.vy, hvz = (T)0.5 * d.vz;
    const T px0 = cx - hvx, px1 = cx + d.vx - hvx;
    const T py0 = cy - hvy, py1 = cy + d.vy - hvy;
    const T pz0 = cz - hvz, pz1 = cz + d.vz - hvz;
    T minY = __builtin_inf(), maxY = -__builtin_inf(), minZ = __builtin_inf(), maxZ = -__builtin_inf();
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int ii = 0; ii < 2; ++ii) {
            const T px = ii ? px1 : px0, py = jj ? py1 : py0;
            const T t = px * c + py * s;
            const T sv = -px * s + py * c;
            const T ratio = (t - d.dod) / (d.dso - t);
            // vertex on the detector per CT_style (:656-670): fan and parallel
            // beam keep the voxel's height, parallel beam also its s
            const T u = d.style == CT_PARALLEL ? sv : ratio * (sv - d.sy) + sv;
            const T w0 = d.style == CT_CONE ? ratio * (pz0 - szs) + pz0 : cz;
            const T w1 = d.style == CT_CONE ? ratio * (pz1 - szs) + pz1 : cz;
            minY = vmin(minY, u); maxY = vmax(maxY, u);
            minZ = vmin(minZ, vmin(w0, w1)); maxZ = vmax(maxZ, vmax(w0, w1));
        }
    int y0 = (int)vfloor((minY - d.ymin) * d.idu + (T)0.5);
    int y1 = (int)vfloor((maxY - d.ymin) * d.idu + (T)0.5);
    int z0 = (int)vfloor((minZ - d.zmin) * d.idv + (T)0.5);
    int z1 = (int)vfloor((maxZ - d.zmin) * d.idv + (T)0.5);
    if ((y0 < 0 && y1 < 0) || (y0 > d.R - 1 && y1 > d.R - 1) || (z0 < 0 && z1 < 0) ||
        (z0 > d.ZP - 1 && z1 > d.ZP - 1))
        return;
    if (y0 <= 0) y0 = 0;
    if (y1 >= d.R - 1) y1 = d.R - 1;
    if (z0 <= 0) z0 = 0;
    if (z1 >= d.ZP - 1) z1 = d.ZP - 1;

    // source per CT_style (:709-724); parallel beam: per column (:738-742)
    const T sxw = d.sx * c - d.sy * s;
    const T syw = d.sx * s + d.sy * c;
    const T szw = d.style == CT_CONE ? szs : cz;
    const T xl = cx - hvx, xh = cx + hvx, yl = cy - hvy, yh = cy + hvy, zl = cz - hvz, zh = cz + hvz;
    const T thr = d.vx * (T)1e-6;
    const int ncol = y1 - y0 + 1;
    if (ncol <= MAXC) {
        T lo[MAXC], hi[MAXC], a2[MAXC];
#pragma unroll
        for (int cc = 0; cc < MAXC; ++cc) {
            if (cc < ncol) {
                const T u = d.ymin + (T)(y0 + cc) * d.du;
                const T qx = d.dod * c - u * s;
                const T qy = d.dod * s + u * c;
                const T sxc = d.style == CT_PARALLEL ? d.sx * c - u * s : sxw;
                const T syc = d.style == CT_PARALLEL ? d.sx * s + u * c : syw;
                const T ex = sxc - qx, ey = syc - qy;
                const T ax0 = (xl - qx) / ex, ax1 = (xh - qx) / ex;
                const T ay0 = (yl - qy) / ey, ay1 = (yh - qy) / ey;
                lo[cc] = vmax(vmin(ax0, ax1), vmin(ay0, ay1));
                hi[cc] = vmin(vmax(ax0, ax1), vmax(ay0, ay1));
                a2[cc] = ex * ex + ey * ey;
            }
        }
        for (int j = z0; j <= z1; ++j) {
            const T qz = d.zmin + (T)j * d.dv;
            const T ez = szw - qz;
            const T az0 = (zl - qz) / ez, az1 = (zh - qz) / ez;
            const T loz = vmin(az0, az1), hiz = vmax(az0, az1);
            const T c2 = ez * ez;
            const T *crow = corr + (int64_t)j * d.R + y0;
#pragma unroll
            for (int cc = 0; cc < MAXC; ++cc) {
                if (cc < ncol) {
                    const T amin = vmax(lo[cc], loz);
                    const T amax = vmin(hi[cc], hiz);
                    if (amax - amin > (T)0) {
                        const T tmp = vsqrt(a2[cc] + c2) * (amax - amin);
                        if (tmp > thr) {
                            W = W + tmp * crow[cc];
                            L = L + tmp;
                        }
                    }
                }
            }
        }
    } else {
        // wide bounding boxes (extreme magnification): per-pixel form
        for (int j = z0; j <= z1; ++j) {
            const T qz = d.zmin + (T)j * d.dv;
            const T ez = szw - qz;
            const T az0 = (zl - qz) / ez, az1 = (zh - qz) / ez;
            const T loz = vmin(az0, az1), hiz = vmax(az0, az1);
            for (int i = y0; i <= y1; ++i) {
                const T u = d.ymin + (T)i * d.du;
                const T qx = d.dod * c - u * s;
                const T qy = d.dod * s + u * c;
                const T sxc = d.style == CT_PARALLEL ? d.sx * c - u * s : sxw;
                const T syc = d.style == CT_PARALLEL ? d.sx * s + u * c : syw;
                const T ex = sxc - qx, ey = syc - qy;
                const T ax0 = (xl - qx) / ex, ax1 = (xh - qx) / ex;
                const T ay0 = (yl - qy) / ey, ay1 = (yh - qy) / ey;
                const T amin = vmax(vmax(vmin(ax0, ax1), vmin(ay0, ay1)), loz);
                const T amax = vmin(vmin(vmax(ax0, ax1), vmax(ay0, ay1)), hiz);
                if (amax - amin > (T)0) {
                    const T tmp = vsqrt(ex * ex + ey * ey + ez * ez) * (amax - amin);
                    if (tmp > thr) {
                        W = W + tmp * corr[(int64_t)j * d.R + i];
                        L = L + tmp;
                    }
                }
            }
        }
    }
}

// FP32 fast form of voxel_wl: the same chord per (voxel, pixel) with the
// hardware reciprocal in place of IEEE division, the hardware square root and
// a branch-free pixel accumulation.  Every coordinate (pixel position, slab
// plane, source offset) keeps the reference's multiply/add sequence: a ray
// that lies exactly in a voxel boundary plane (zero divisor) then resolves
// exactly as in the reference (x/0 = inf, 0/0 = NaN), and the remaining
// ~1 ulp differences only move chords shorter than the 1e-6*vx cut.

__device__ __forceinline__ void voxel_wl_fast(const DGeo<float> &d, float cx, float cy, float cz, float s, float c,
                                              float szs, const float *__restrict__ corr, float &W, float &L) {
    if (d.tb) {
        const float tc = cx * c + cy * s;
        const float rk = (d.dso - d.dod) * frcp(tc - d.dod);
        const float upper = (cz + d.vz - d.zmin) * rk + d.zmin;
        const float lower = (cz - d.vz - d.zmax) * rk + d.zmax;
        if (szs < lower || szs > upper) return;
    }
    const float hvx = 0.5f * d.vx, hvy = 0.5f * d.vy, hvz = 0.5f * d.vz;
    const float xl = cx - hvx, xh = cx + hvx, yl = cy - hvy, yh = cy + hvy, zl = cz - hvz, zh = cz + hvz;
    float minY = __builtin_inff(), maxY = -__builtin_inff(), minZ = __builtin_inff(), maxZ = -__builtin_inff();
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int ii = 0; ii < 2; ++ii) {
            const float px = ii ? xh : xl, py = jj ? yh : yl;
            const float t = px * c + py * s;
            const float sv = -px * s + py * c;
            const float ratio = (t - d.dod) * frcp(d.dso - t);
            const float u = d.style == CT_PARALLEL ? sv : fmaf(ratio, sv - d.sy, sv);
            const float w0 = d.style == CT_CONE ? fmaf(ratio, zl - szs, zl) : cz;
            const float w1 = d.style == CT_CONE ? fmaf(ratio, zh - szs, zh) : cz;
            minY = fminf(minY, u); maxY = fmaxf(maxY, u);
            minZ = fminf(minZ, fminf(w0, w1)); maxZ = fmaxf(maxZ, fmaxf(w0, w1));
        }
    int y0 = (int)floorf(fmaf(minY - d.ymin, d.idu, 0.5f));
    int y1 = (int)floorf(fmaf(maxY - d.ymin, d.idu, 0.5f));
    int z0 = (int)floorf(fmaf(minZ - d.zmin, d.idv, 0.5f));
    int z1 = (int)floorf(fmaf(maxZ - d.zmin, d.idv, 0.5f));
    if ((y0 < 0 && y1 < 0) || (y0 > d.R - 1 && y1 > d.R - 1) || (z0 < 0 && z1 < 0) ||
        (z0 > d.ZP - 1 && z1 > d.ZP - 1))
        return;
    y0 = max(y0, 0); y1 = min(y1, d.R - 1);
    z0 = max(z0, 0); z1 = min(z1, d.ZP - 1);
    const float sxw = d.sx * c - d.sy * s;
    const float syw = d.sx * s + d.sy * c;
    if (d.style != CT_CONE) szs = cz;
    const float thr = d.vx * 1e-6f;
    const int ncol = y1 - y0 + 1;
    if (ncol > MAXC) {
        // wide bounding boxes: per-pixel form
        for (int j = z0; j <= z1; ++j) {
            const float qz = d.zmin + (float)j * d.dv;
            const float ez = szs - qz;
            const float rz = frcp(ez);
            const float az0 = (zl - qz) * rz, az1 = (zh - qz) * rz;
            const float loz = fminf(az0, az1), hiz = fmaxf(az0, az1);
            for (int i = y0; i <= y1; ++i) {
                const float u = d.ymin + (float)i * d.du;
                const float qx = d.dod * c - u * s, qy = d.dod * s + u * c;
                const float ex = (d.style == CT_PARALLEL ? d.sx * c - u * s : sxw) - qx;
                const float ey = (d.style == CT_PARALLEL ? d.sx * s + u * c : syw) - qy;
                const float rx = frcp(ex), ry = frcp(ey);
                const float ax0 = (xl - qx) * rx, ax1 = (xh - qx) * rx;
                const float ay0 = (yl - qy) * ry, ay1 = (yh - qy) * ry;
                const float amin = fmaxf(fmaxf(fminf(ax0, ax1), fminf(ay0, ay1)), loz);
                const float amax = fminf(fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)), hiz);
                const float len = fsqrt(fmaf(ex, ex, fmaf(ey, ey, ez * ez))) * (amax - amin);
                if (amax > amin && len > thr) {
                    W = fmaf(len, corr[(int64_t)j * d.R + i], W);
                    L += len;
                }
            }
        }
        return;
    }
    float lo[MAXC], hi[MAXC], a2[MAXC];
#pragma unroll
    for (int cc = 0; cc < MAXC; ++cc) {
        if (cc < ncol) {
            const float u = d.ymin + (float)(y0 + cc) * d.du;
            const float qx = d.dod * c - u * s, qy = d.dod * s + u * c;
            const float ex = (d.style == CT_PARALLEL ? d.sx * c - u * s : sxw) - qx;
            const float ey = (d.style == CT_PARALLEL ? d.sx * s + u * c : syw) - qy;
            const float rx = frcp(ex), ry = frcp(ey);
            const float ax0 = (xl - qx) * rx, ax1 = (xh - qx) * rx;
            const float ay0 = (yl - qy) * ry, ay1 = (yh - qy) * ry;
            lo[cc] = fmaxf(fminf(ax0, ax1), fminf(ay0, ay1));
            hi[cc] = fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1));
            a2[cc] = fmaf(ex, ex, ey * ey);
        }
    }
    for (int j = z0; j <= z1; ++j) {
        const float qz = d.zmin + (float)j * d.dv;
        const float ez = szs - qz;
        const float rz = frcp(ez);
        const float az0 = (zl - qz) * rz, az1 = (zh - qz) * rz;
        const float loz = fminf(az0, az1), hiz = fmaxf(az0, az1);
        const float c2 = ez * ez;
        const float *crow = corr + (int64_t)j * d.R + y0;
#pragma unroll
        for (int cc = 0; cc < MAXC; ++cc) {
            if (cc < ncol) {
                const float span = fminf(hi[cc], hiz) - fmaxf(lo[cc], loz);
                const float len = fsqrt(a2[cc] + c2) * span;
                const bool ok = (span > 0.0f) & (len > thr);
                const float cv = crow[cc];
                W = ok ? fmaf(len, cv, W) : W;
                L = ok ? L + len : L;
            }
        }
    }
}

// Per-view tables of the LDS-staged fast backprojector, built in one pass
// over the subset's correction images: per pixel (c*|S-P|, |S-P|), per column
// (qx, qy, 1/ex, 1/ey), per row and source (qa, qb, 1/ez).  The pixel,
// slab-plane and source coordinates use the reference's arithmetic (see
// voxel_wl_fast).
// The row entry orders the ray's two z-plane crossings of a voxel with lower
// face zl: its alpha interval is [(zl - qa)/ez, (zl - qb)/ez], with
// (qa, qb) = (qz, qz - vz) for a rising ray (1/ez > 0) and (qz - vz, qz) for
// a falling one, so the update needs no min/max per row.  A horizontal ray
// (1/ez = +-inf) crosses a voxel exactly when zl < qz < zl + vz; at a tie the
// reference's min/max of (+-inf, NaN) leaves it out.  The infinite products
// then give +-inf, and NaN (taken by min/max as "no bound") exactly where zl
// equals qa or qb, so qa and qb are moved one float outwards (qz down, qz - vz
// up): every tie then lands on the empty side.
// PK (the packed-half engine of reduced-precision volumes): the pixel entry
// holds the column pair (i, i + 1) as half pairs, (c*|S-P|, c*|S-P|) and
// (|S-P|, |S-P|) rounded to nearest (zero past the last column), so that the
// engine stages a footprint row by plain 8-byte copies.  Finite values are
// clamped to the FP16 range first: a correction beyond it (a ray that grazes
// the volume has a tiny length to divide by) would otherwise turn the voxel's
// sums into infinities or NaN.
template <bool PK>
__global__ __launch_bounds__(256) void k_bp_prep(DGeo<float> d, const float *__restrict__ ang,
                                                 const float *__restrict__ corr, int first, int stride,
                                                 float2 *__restrict__ pix, float4 *__restrict__ colT,
                                                 float4 *__restrict__ rowT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, img = blockIdx.z;
    if (i >= d.R) return;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = first + q * stride;
    const float s = ang[2 * v], c = ang[2 * v + 1];
    const float szs = d.tb ? (float)(d.sz0 + (float)src * d.sdz) : d.sz0;
    const float sxw = d.sx * c - d.sy * s;
    const float syw = d.sx * s + d.sy * c;
    const float qz = d.zmin + (float)j * d.dv;
    const float ez = szs - qz;
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t p = (int64_t)img * P + (int64_t)j * d.R + i;
    // (c*|S-P|, |S-P|) of column i2 of this row; its detector point (qx, qy)
    // and 1/ex, 1/ey for the column table
    auto pixel = [&](int i2, float4 &ct) {
        const float u = d.ymin + (float)i2 * d.du;
        const float qx = d.dod * c - u * s;
        const float qy = d.dod * s + u * c;
        const float ex = sxw - qx, ey = syw - qy;
        const float dist = sqrtf(ex * ex + ey * ey + ez * ez);
        ct = make_float4(qx, qy, 1.0f / ex, 1.0f / ey);
        return make_float2(corr[p + (i2 - i)] * dist, dist);
    };
    float4 ct, cn;
    const float2 a = pixel(i, ct);
    if constexpr (PK) {
        const float2 b = i + 1 < d.R ? pixel(i + 1, cn) : make_float2(0.0f, 0.0f);
        auto h = [](float v) {   // (a NaN stays NaN)
            const float cl = fminf(fmaxf(v, -65504.0f), 65504.0f);
            return (unsigned)__builtin_bit_cast(unsigned short, (_Float16)(v == v ? cl : v));
        };
        const unsigned wx = h(a.x) | (h(b.x) << 16);
        const unsigned wy = h(a.y) | (h(b.y) << 16);
        pix[p] = make_float2(__uint_as_float(wx), __uint_as_float(wy));
    } else {
        pix[p] = a;
    }
    if (j == 0 && src == 0) colT[(int64_t)q * d.R + i] = ct;
    if (i == 0) {
        const float kz = 1.0f / ez, qd = qz - d.vz;
        float qa = kz > 0.0f ? qz : qd, qb = kz > 0.0f ? qd : qz;
        if (kz == __builtin_inff()) {
            qa = nextafterf(qz, -__builtin_inff());
            qb = nextafterf(qd, __builtin_inff());
        } else if (kz == -__builtin_inff()) {
            qa = nextafterf(qd, __builtin_inff());
            qb = nextafterf(qz, -__builtin_inff());
        }
        rowT[(int64_t)img * d.ZP + j] = make_float4(qa, qb, kz, 0.0f);
    }
}

constexpr int BPX = 16, BPY = 16;
// z-run per lane: 16 voxels for single-image launches, 8 when per-voxel sums
// are carried across images (they live in registers)
template <bool ONE> constexpr int zrun() { return ONE ? 16 : 8; }
constexpr int LW = 64;          // staged footprint columns
constexpr int LP = LW + 8;      // LDS row pitch: 8 zero columns for the NC overrun
constexpr int LH = 48;          // staged footprint rows
constexpr int NCMAX = 8;

// Packed-half chord engine for reduced-precision storage.  Two detector
// columns share one 32-bit register (lo half, hi half).  Alpha values are
// taken in voxel-local units, l = (alpha - alpha0)*kl with
// alpha0 = (t_centre - DOD)/(DSO - DOD) (alpha is linear in the depth t along
// every ray) and kl = (DSO - DOD)/|voxel diagonal|, so the chord interval of
// a ray through the voxel lies within |l| <= 1/2 and FP16 resolves it to
// ~5e-4 of the voxel diagonal; values far outside only take part in the
// min/max and may round or saturate freely.  Products are accumulated in
// FP32 by v_dot2_f32_f16; W and L come out in local units (times kl), so
// W/L is unchanged and both are scaled back before the update.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pk_rtz(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(a, b));
}
// min(a, b.hi) and max(a, b.lo) per half (one half of b broadcast)
__device__ __forceinline__ unsigned pk_min_h(unsigned a, unsigned b) {
    unsigned r;
    asm("v_pk_min_f16 %0, %1, %2 op_sel:[0,1] op_sel_hi:[1,1]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ unsigned pk_max_l(unsigned a, unsigned b) {
    unsigned r;
    asm("v_pk_max_f16 %0, %1, %2 op_sel:[0,0] op_sel_hi:[1,0]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// max(a - b, 0) per half, for a - b < 1 (output clamp to [0, 1]); local
// spans are at most (DSO - DOD)/|S - P| <= 1
__device__ __forceinline__ unsigned pk_span(unsigned a, unsigned b) {
    unsigned r;
    asm("v_pk_add_f16 %0, %1, %2 neg_lo:[0,1] neg_hi:[0,1] clamp" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float dot2(unsigned a, unsigned b, float c) {
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, a), __builtin_bit_cast(half2_t, b), c, false);
}

// rows of the packed-half engine: per row (qa, qb, kl/ez) from LDS, per
// column pair the staged (c*|S-P|, |S-P|) half pairs
template <int NP>
__device__ __forceinline__ void bp_rows_h(const uint2 *__restrict__ sp, const float4 *__restrict__ sr, int z0, int z1,
                                          int cbase, float zl, float nA, const unsigned *lo2,
                                          const unsigned *hi2, float &wk, float &lk) {
    const uint2 *prow = sp + (__umul24((uint32_t)z0, LP) + cbase);
    for (const float4 *rp = sr + z0, *const re = sr + z1; rp <= re; ++rp, prow += LP) {
        const float4 rt = *rp;
        uint2 pv[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) pv[p] = prow[2 * p];
        // the voxel's local alpha interval on this row's ray, (lower, upper)
        // in (lo, hi) half (a horizontal ray: +-inf, or NaN for no bound)
        const unsigned az = pk_rtz(fmaf(zl - rt.x, rt.z, nA), fmaf(zl - rt.y, rt.z, nA));
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const unsigned span = pk_span(pk_min_h(hi2[p], az), pk_max_l(lo2[p], az));
            wk = dot2(span, pv[p].x, wk);
            lk = dot2(span, pv[p].y, lk);
        }
    }
}

// LDS-staged fast backprojector.  A workgroup owns a 16x16 (x, y) tile of
// ZR (16 or 8) voxels in z; each lane owns the ZR voxels of one (x, y).  Per image the
// workgroup stages its detector footprint - the projection of the tile's box,
// widened by one pixel each way - from the per-view tables into LDS: pixel
// values (c*|S-P|, |S-P|) in rows of LP entries (columns past the footprint
// are zero), the column table (qx, qy, 1/ex, 1/ey) and the ordered row table
// (qa, qb, 1/ez) of k_bp_prep.  Each voxel's bounding box lies inside the
// footprint, so every
// pixel read of the chord sum is an unconditional LDS read.  The per-lane
// column count is raised to a wave-uniform NC in {4, 5, 6, 8}; columns past a
// lane's own bounding box get an empty alpha interval and add exactly zero.
// Per pixel: span = max(min(hi_col, hi_row) - max(lo_col, lo_row), 0),
// W += span*(c*|S-P|), L += span*|S-P|.  Chords shorter than 1e-6*vx are not
// cut per pixel (they move W and L by less than 1e-6*vx each); the update
// threshold on L stays.  Footprints larger than the LDS window take the
// per-voxel table path for that image.

template <int NC>
__device__ __forceinline__ void bp_rows(const float2 *__restrict__ sp, const float4 *__restrict__ sr, int z0, int z1,
                                        int cbase, float zl, const float *lo, const float *hi, float &wk,
                                        float &lk) {
    const float2 *prow = sp + (__umul24((uint32_t)z0, LP) + cbase);
    for (const float4 *rp = sr + z0, *const re = sr + z1; rp <= re; ++rp, prow += LP) {
        const float4 rt = *rp;
        float2 pv[NC];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) pv[cc] = prow[cc];
        // the voxel's alpha interval on this row's ray, from the ordered
        // plane offsets in the reference's product form (a horizontal ray:
        // +-inf, or NaN, which min/max take as no bound)
        const float loz = (zl - rt.x) * rt.z, hiz = (zl - rt.y) * rt.z;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            // alpha spans are below 1 (a chord is a fraction of |S - P|)
            const float span = hsub_sat(hmin(hi[cc], hiz), hmax(lo[cc], loz));
            wk = fmaf(span, pv[cc].x, wk);
            lk = fmaf(span, pv[cc].y, lk);
        }
    }
}

// Stochastic rounding of an updated voxel to 16-bit storage.  A SART step
// adds ~1e-4 of a voxel's value, a fraction of an FP16 ulp; rounding to
// nearest loses it with a bias that accumulates over iterations.  Stochastic
// rounding keeps it in expectation: the value rounds up with probability
// equal to its fraction of an ulp.  The random bits come from the voxel's own
// sums W and L (sums of ~10^2 products that change with every view) through
// sr_bits, which keeps runs reproducible.
//  * FP16: |v| + u ulp16(v), u in [0, 1) from the low 16 bits, truncated
//    towards zero by the packed round-to-zero conversion: the magnitude
//    rounds up exactly when u >= 1 - fraction, probability `fraction` to
//    within 2^-16, so the stored value is the exact one in expectation
//    (also at powers of two, where truncation lands on the grid of v's own
//    binade or the next).  ulp16(v) = 2^(e - 10) from v's exponent e, 2^-24
//    in the subnormal range (|v| < 2^-14); the exponent is capped at 2^15,
//    so infinities stay infinite and a NaN stays NaN; finite values beyond
//    the FP16 range store 65504.  A value already on the FP16 grid never
//    moves.
//  * BF16: the 16 dropped bits get the dither, then the value is truncated
//    (exact in expectation); a NaN keeps a quiet-NaN pattern instead of
//    carrying into the sign bit.
template <typename SV, typename I> __device__ __forceinline__ void st_sr(SV *p, I i, float v, uint32_t r) {
    St<SV>::st(p, i, v);
}
template <typename I> __device__ __forceinline__ void st_sr(__half *p, I i, float v, uint32_t r) {
    const uint32_t b = __float_as_uint(v);
    // sign(v) 2^e, e clamped to [-14, 15]
    const float e2 = __uint_as_float(min(max(b & 0x7F800000u, 0x38800000u), 0x47000000u) | (b & 0x80000000u));
    const auto h2 = __builtin_amdgcn_cvt_pkrtz(fmaf((float)(r & 0xFFFFu) * 0x1p-26f, e2, v), 0.0f);
    p[i] = __builtin_bit_cast(__half, h2[0]);
}
template <typename I> __device__ __forceinline__ void st_sr(__hip_bfloat16 *p, I i, float v, uint32_t r) {
    const uint32_t u = __float_as_uint(v) + (r & 0xFFFFu);
    const unsigned short h = v == v ? (unsigned short)(u >> 16) : (unsigned short)0x7FC0u;
    p[i] = __builtin_bit_cast(__hip_bfloat16, h);
}
// The dither bits: the low significand bits of (W c) XOR (L c) for
// c = 1/sqrt(2) rounded to float.  The packed-half engine's W and L are short
// sums of exact FP16 products whose low bits are mostly zero, and under a
// uniform correction W is nearly a fixed multiple of L, so their own low bits
// are neither uniform nor independent; each product with c carries ~48
// significant bits into its rounding, which makes the low bits of both vary
// at every ulp.  Two multiplies and an XOR.  Over 20 SART-FISTA-TV
// iterations at 512^3 / 1024 views this leaves FP16 storage 4.7x closer to
// the FP32 path than a hash of the voxel index and view does, with no larger
// mean drift; (W c + L) alone correlates with the rounded fraction and drifts
// 10x more.
__device__ __forceinline__ uint32_t sr_bits(float W, float L) {
    return __float_as_uint(W * 0x1.6a09e6p-1f) ^ __float_as_uint(L * 0x1.6a09e6p-1f);
}

// The update thresholds on L as one float comparison: !(L < vx 1e-6) for
// cone beam (V63/kernel_IterativeRecon_universal.cu:799) and L > vx 1e-3 in
// double for multi-source (TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:772),
// which for a float L is L > (vx 1e-3 rounded down to float).  `scale`
// carries L's units (the packed-half engine's kl).  Both are one comparison
// !(L < lth): for multi-source, L > t is L >= t' with t' the next float
// above t (L is a sum of clamped spans times finite lengths, never NaN).
struct BpThresh {
    float lth;
    __device__ __forceinline__ BpThresh(const DGeo<float> &d, float scale) {
        if (!d.tb) lth = d.vx * 1e-6f * scale;
        else {
            const double t = (double)d.vx * 1e-3;
            float f = (float)t;
            if ((double)f > t) f = __uint_as_float(__float_as_uint(f) - 1u);   // next float towards 0 (f > 0)
            lth = __uint_as_float(__float_as_uint(f * scale) + 1u);            // next float above (f * scale > 0)
        }
    }
    __device__ __forceinline__ bool pass(float L) const { return !(L < lth); }
};

// SART update of one voxel from its accumulated W = sum l*c and L = sum l
// (V63/kernel_IterativeRecon_universal.cu:799-806; TB threshold
// TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:772) in fast mode: W / L by
// the hardware reciprocal (1 ulp), 16-bit volumes stored with stochastic
// rounding.  W and L may carry a common scale (their ratio is unchanged);
// wscale takes W back to length units for mode 0 (x = W).
template <typename SV, typename I>
__device__ __forceinline__ void bp_update(SV *__restrict__ f, I idx, float W, float L, float beta, int mode,
                                          const BpThresh &th, float wscale) {
    if (!th.pass(L)) return;
    if (mode == 0) St<SV>::st(f, idx, W * wscale);
    else st_sr(f, idx, fmaf(beta * W, frcp(L), ldf(f, idx)), sr_bits(W, L));
}
// the same with the voxel's old value loaded ahead (raw storage form; mode 1)
template <typename SV, typename I>
__device__ __forceinline__ void bp_update_x(SV *__restrict__ f, I idx, typename RawOf<SV>::type xr, float W, float L,
                                            float beta, int mode, const BpThresh &th, float wscale) {
    if (!th.pass(L)) return;
    if (mode == 0) St<SV>::st(f, idx, W * wscale);
    else st_sr(f, idx, fmaf(beta * W, frcp(L), from_raw<float, SV>(xr)), sr_bits(W, L));
}

// ONE: the launch backprojects a single image (one view, one source), so each
// voxel is updated as soon as its sums are formed and no per-voxel sums are
// carried across images.  H: the packed-half chord engine (reduced-precision
// storage paths).  ZQ: the volume is in the sweep's z-quad layout
// [z/4][y][x][z%4].
// TR: the transposed z-quad layout [z/4][x][y][z%4] of the sweep's
// y-directed views; the lanes then run along y (threadIdx.x -> y) so that a
// wave's accesses stay contiguous.
// (element index; the LDS-staged kernel runs on volumes of under 2^31
// elements, bp_lds_fits)
template <bool ZQ, bool TR = false>
__device__ __forceinline__ uint32_t vox_index(const DGeo<float> &d, int ix, int iy, int iz) {
    const uint32_t M = d.M, N = d.N;
    if constexpr (ZQ && TR) return ((((uint32_t)iz >> 2) * M + (uint32_t)ix) * N + (uint32_t)iy) * 4u + ((uint32_t)iz & 3u);
    else if constexpr (ZQ) return ((((uint32_t)iz >> 2) * N + (uint32_t)iy) * M + (uint32_t)ix) * 4u + ((uint32_t)iz & 3u);
    else return ((uint32_t)iz * N + (uint32_t)iy) * M + (uint32_t)ix;
}
// vox_index split into the part set by the z index (uniform over a
// workgroup's z-run step) and the part set by (x, y) (per lane):
// vox_index = vox_plane(iz) + vox_lane(ix, iy).  The voxel's address is then
// a scalar base plus a 32-bit lane offset, with no per-voxel vector address
// arithmetic.
template <bool ZQ>
__device__ __forceinline__ uint32_t vox_plane(const DGeo<float> &d, int iz) {
    const uint32_t MN = (uint32_t)d.M * (uint32_t)d.N;
    if constexpr (ZQ) return ((uint32_t)iz >> 2) * MN * 4u + ((uint32_t)iz & 3u);
    else return (uint32_t)iz * MN;
}
template <bool ZQ, bool TR = false>
__device__ __forceinline__ uint32_t vox_lane(const DGeo<float> &d, int ix, int iy) {
    const uint32_t M = d.M, N = d.N;
    if constexpr (ZQ && TR) return ((uint32_t)ix * N + (uint32_t)iy) * 4u;
    else if constexpr (ZQ) return ((uint32_t)iy * M + (uint32_t)ix) * 4u;
    else return (uint32_t)iy * M + (uint32_t)ix;
}

template <typename SV, bool ONE, bool H, bool ZQ, bool TR = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ONE ? 5 : 4))) void k_voxel_bp_lds(DGeo<float> d, const float *__restrict__ ang,
                                                      SV *__restrict__ f, const float2 *__restrict__ pixT,
                                                      const float4 *__restrict__ colT,
                                                      const float4 *__restrict__ rowT, int first, int stride,
                                                      int nv, float beta, int mode) {
    __shared__ float2 sp[LH * LP];   // (c*|S-P|, |S-P|); H: column-pair halves
    __shared__ float4 sc[LP];
    __shared__ float4 sr[LH];        // (qa, qb, 1/ez); H: (qa, qb, kl/ez)
    uint2 *const sph = reinterpret_cast<uint2 *>(sp);
    const int tid = threadIdx.y * BPX + threadIdx.x;
    const int3 tile = xcd_tile();   // each XCD: a contiguous z range of tiles
    const int ix = tile.x * BPX + (TR ? threadIdx.y : threadIdx.x);
    const int iy = tile.y * BPY + (TR ? threadIdx.x : threadIdx.y);
    constexpr int ZR = zrun<ONE>();
    const int iz0 = tile.z * ZR;
    const int nz = min(ZR, d.Z - iz0);
    bool active = ix < d.M && iy < d.N;
    if (active) {   // cylinder mask
        const float hxm = (float)d.M * 0.5f - 0.5f, hym = (float)d.N * 0.5f - 0.5f;
        const float ma = (float)ix - hxm - (float)d.moff;
        const float mb = (float)iy - hym;
        active = !(ma * d.vx * ma * d.vx + mb * d.vy * mb * d.vy >= hxm * d.vx * hym * d.vy);
    }
    if (!__syncthreads_or(active)) return;   // the whole tile is outside the cylinder
    float cx, cy;
    if (!d.tb) {
        cx = d.bx + d.vx * 0.5f + (float)ix * d.vx;
        cy = d.by + d.vy * 0.5f + (float)iy * d.vy;
    } else {
        cx = d.bx + d.vx * (0.5f + (float)ix);
        cy = d.by + d.vy * (0.5f + (float)iy);
    }
    const float hvx = 0.5f * d.vx, hvy = 0.5f * d.vy, hvz = 0.5f * d.vz;
    const float xl = cx - hvx, xh = cx + hvx, yl = cy - hvy, yh = cy + hvy;
    // tile box (clipped to the volume) for the footprint
    const int tx0 = tile.x * BPX, ty0 = tile.y * BPY;
    const float TX0 = d.bx + (float)tx0 * d.vx, TX1 = d.bx + (float)min(tx0 + BPX, d.M) * d.vx;
    const float TY0 = d.by + (float)ty0 * d.vy, TY1 = d.by + (float)min(ty0 + BPY, d.N) * d.vy;
    const float TZ0 = d.bz + (float)iz0 * d.vz, TZ1 = d.bz + (float)(iz0 + nz) * d.vz;
    // packed-half engine: alpha scale kl and its inverse
    const float vdiag = sqrtf(d.vx * d.vx + d.vy * d.vy + d.vz * d.vz);
    const float ivd = frcp(vdiag);
    const float kl = (d.dso - d.dod) * ivd, ikl = vdiag * frcp(d.dso - d.dod);
    // the packed-half engine's sums come out times kl: W/L is unchanged, the
    // threshold on L and mode 0's W carry the factor
    const BpThresh th(d, H ? kl : 1.0f);
    const float wsc = H ? ikl : 1.0f;
    float W[ONE ? 1 : ZR], L[ONE ? 1 : ZR];
#pragma unroll
    for (int k = 0; k < (ONE ? 1 : ZR); ++k) { W[k] = 0.0f; L[k] = 0.0f; }
    const int64_t P = (int64_t)d.R * d.ZP;
    const uint32_t lane_b = vox_lane<ZQ, TR>(d, ix, iy) * (uint32_t)sizeof(SV);
    for (int q = 0; q < nv; ++q) {
        const int v = first + q * stride;
        const float s = ang[2 * v], c = ang[2 * v + 1];
        // this lane's corner ratios and column bounding box (:639-678)
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
        const bool yoff = (y0r < 0 && y1r < 0) || (y0r > d.R - 1 && y1r > d.R - 1);
        // the tile's footprint: columns from its four (x, y) corners, rows
        // from its z faces at the extreme ratios
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
        // the host admits this kernel only when every footprint fits the LDS
        // window (bp_lds_fits); the clamps keep LDS indexing in range regardless
        const int C1 = min(min((int)floorf(fmaf(Tumax - d.ymin, d.idu, 0.5f)) + 1, d.R - 1), C0 + LW - 1);
        const int ncw = C1 - C0 + 1;
        for (int src = 0; src < d.NS; ++src) {
            const int img = q * d.NS + src;
            const float szs = d.tb ? (float)(d.sz0 + (float)src * d.sdz) : d.sz0;
            const float e0 = TZ0 - szs, e1 = TZ1 - szs;
            const float Tw0 = fminf(fmaf(Trmin, e0, TZ0), fmaf(Trmax, e0, TZ0));
            const float Tw1 = fmaxf(fmaf(Trmin, e1, TZ1), fmaf(Trmax, e1, TZ1));
            const int R0 = max((int)floorf(fmaf(Tw0 - d.zmin, d.idv, 0.5f)) - 1, 0);
            const int R1 = min(min((int)floorf(fmaf(Tw1 - d.zmin, d.idv, 0.5f)) + 1, d.ZP - 1), R0 + LH - 1);
            const int nrow = R1 - R0 + 1;
            if (ncw <= 0 || nrow <= 0) continue;   // footprint off the detector (uniform)
            const float2 *pimg = pixT + (int64_t)img * P;
            const float4 *rimg = rowT + (int64_t)img * d.ZP;
            const float4 *cimg = colT + (int64_t)q * d.R;
            // TB source culling, TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:618-627
            float rk = 0.0f;
            if (d.tb) rk = (d.dso - d.dod) * frcp(cx * c + cy * s - d.dod);
            __syncthreads();   // the previous image's readers are done
            // 4 rows x 64 columns per pass; columns past the footprint (up to
            // the row pitch) are zero.  (H: the entries are already column
            // pairs of halves.)
            const int cl = tid & 63;
            if constexpr (H) {
                // every pass's load, and the column and row entries, issued
                // before the first LDS write: one memory round trip per image
                // instead of one per pass (FP16 update 0.350 -> 0.336 ms per
                // view at 512^3; the FP32 variant, at its register limit,
                // spills with it and keeps one pass at a time)
                static_assert(LH % 4 == 0, "staging passes");
                constexpr int RP = LH / 4;
                const int r0 = tid >> 6;
                const float2 *g = pimg + (int64_t)(R0 + r0) * d.R + C0 + cl;
                float2 pv[RP];
#pragma unroll
                for (int i = 0; i < RP; ++i)
                    pv[i] = r0 + 4 * i < nrow && cl < ncw ? g[(int64_t)(4 * i) * d.R] : make_float2(0.0f, 0.0f);
                float4 ct = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rt = ct;
                if (tid < ncw) ct = cimg[C0 + tid];
                if (tid < nrow) rt = rimg[R0 + tid];
#pragma unroll
                for (int i = 0; i < RP; ++i) {
                    const int r = r0 + 4 * i;
                    if (r < nrow) {
                        sp[r * LP + cl] = pv[i];
                        if (cl < LP - LW) sp[r * LP + LW + cl] = make_float2(0.0f, 0.0f);
                    }
                }
                if (tid < LP) sc[tid] = ct;
                if (tid < nrow) sr[tid] = make_float4(rt.x, rt.y, kl * rt.z, 0.0f);
            } else {
                for (int r = tid >> 6; r < nrow; r += 4) {
                    const float2 *g = pimg + (int64_t)(R0 + r) * d.R + C0;
                    sp[r * LP + cl] = cl < ncw ? g[cl] : make_float2(0.0f, 0.0f);
                    if (cl < LP - LW) sp[r * LP + LW + cl] = make_float2(0.0f, 0.0f);
                }
                if (tid < LP) sc[tid] = tid < ncw ? cimg[C0 + tid] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (tid < nrow) sr[tid] = rimg[R0 + tid];
            }
            __syncthreads();
            if (!active || yoff) continue;
            const int y0 = max(y0r, C0), y1 = min(min(y1r, C1), y0 + NCMAX - 1);
            const int ncol = y1 - y0 + 1;
            if (ncol <= 0) continue;
            const int cbase = y0 - C0;
            // wave-uniform column count
            int NC;
            if (__builtin_amdgcn_ballot_w64(ncol > 6)) NC = 8;
            else if (__builtin_amdgcn_ballot_w64(ncol > 5)) NC = 6;
            else if (__builtin_amdgcn_ballot_w64(ncol > 4)) NC = 5;
            else if (__builtin_amdgcn_ballot_w64(ncol > 3)) NC = 4;
            else NC = 3;
            float lo[NCMAX], hi[NCMAX];
#pragma unroll
            for (int cc = 0; cc < NCMAX; ++cc) {
                if (cc >= NC) { lo[cc] = __builtin_inff(); hi[cc] = -__builtin_inff(); continue; }
                const float4 ct = sc[min(cbase + cc, LP - 1)];
                const float ax0 = (xl - ct.x) * ct.z, ax1 = (xh - ct.x) * ct.z;
                const float ay0 = (yl - ct.y) * ct.w, ay1 = (yh - ct.y) * ct.w;
                const bool in = cc < ncol;
                lo[cc] = in ? hmax(hmin(ax0, ax1), hmin(ay0, ay1)) : __builtin_inff();
                hi[cc] = in ? hmin(hmax(ax0, ax1), hmax(ay0, ay1)) : -__builtin_inff();
            }
            // H: the column intervals in voxel-local alpha units, paired
            const float nA = (d.dod - (cx * c + cy * s)) * ivd;
            unsigned lo2[NCMAX / 2], hi2[NCMAX / 2];
            if constexpr (H) {
#pragma unroll
                for (int p = 0; p < NCMAX / 2; ++p) {
                    if (2 * p >= NC) break;
                    lo2[p] = pk_rtz(fmaf(lo[2 * p], kl, nA), fmaf(lo[2 * p + 1], kl, nA));
                    hi2[p] = pk_rtz(fmaf(hi[2 * p], kl, nA), fmaf(hi[2 * p + 1], kl, nA));
                }
            }
            // The voxels' z faces: the planes from a float plane counter (one
            // FMA each), each plane evaluated once for the two voxels it
            // bounds.  Its shadow height w = z + r (z - z_source) is monotone
            // in the corner ratio r, so the extremes over the eight corners
            // come from the extreme ratios: min(w(rmin), w(rmax)) is the lowest
            // shadow of a lower face, the max the highest of an upper face.
            // Shadow rows, counted from the staged window's first row R0:
            // shadow_lo / shadow_hi with their 0.01-row margins folded into
            // the offset.  The row range only has to cover the rows whose
            // rays cross the voxel (any other row adds exactly zero), which
            // the margins keep through the different rounding.
            const float zrow = -d.zmin * d.idv - (float)R0;
            const float zrow_lo = zrow - 0.01f, zrow_hi = zrow + 0.01f;
            float bz = d.bz;
            asm volatile("" : "+v"(bz));   // a lane register: one scalar operand per FMA
            float zk = (float)iz0;
            float zh = fmaf(zk, d.vz, bz);
            int z0n;   // shadow of the next voxel's lower face
            {
                const float e = zh - szs;
                z0n = (int)ceilf(fmaf(hmin(fmaf(rmin, e, zh), fmaf(rmax, e, zh)), d.idv, zrow_lo));
            }
#pragma unroll 1
            for (int k = 0; k < nz; ++k) {
                const float zl = zh;
                zk += 1.0f;
                zh = fmaf(zk, d.vz, bz);
                const float e = zh - szs;
                const float wa = fmaf(rmin, e, zh), wb = fmaf(rmax, e, zh);
                int z0 = z0n;
                int z1 = (int)floorf(fmaf(hmax(wa, wb), d.idv, zrow_hi));
                z0n = (int)ceilf(fmaf(hmin(wa, wb), d.idv, zrow_lo));
                // single image: the voxel's old value is loaded here, its
                // latency hidden behind the chord sums (mode 0 leaves it
                // unset and unused: an initialising move would wait for the
                // previous voxel's load)
                // the voxel: scalar plane base plus the lane's byte offset
                SV *const fv = reinterpret_cast<SV *>(reinterpret_cast<char *>(f + vox_plane<ZQ>(d, iz0 + k)) + lane_b);
                typename RawOf<SV>::type xr;
                if constexpr (ONE) {
                    if (mode != 0) xr = ld_raw(fv, 0);
                }
                if (d.tb) {
                    const float cz = d.bz + d.vz * (0.5f + (float)(iz0 + k));
                    const float upper = (cz + d.vz - d.zmin) * rk + d.zmin;
                    const float lower = (cz - d.vz - d.zmax) * rk + d.zmax;
                    if (szs < lower || szs > upper) continue;
                }
                // rows of the shadow inside the staged window; a shadow off
                // the detector leaves the range empty (no rows, L = 0, below
                // the update threshold)
                z0 = max(z0, 0);
                z1 = min(z1, nrow - 1);
                float wk = ONE ? 0.0f : W[k], lk = ONE ? 0.0f : L[k];
                if constexpr (H) {
                    if (NC == 8) bp_rows_h<4>(sph, sr, z0, z1, cbase, zl, nA, lo2, hi2, wk, lk);
                    else if (NC >= 5) bp_rows_h<3>(sph, sr, z0, z1, cbase, zl, nA, lo2, hi2, wk, lk);
                    else bp_rows_h<2>(sph, sr, z0, z1, cbase, zl, nA, lo2, hi2, wk, lk);
                } else {
                    if (NC == 8) bp_rows<8>(sp, sr, z0, z1, cbase, zl, lo, hi, wk, lk);
                    else if (NC == 6) bp_rows<6>(sp, sr, z0, z1, cbase, zl, lo, hi, wk, lk);
                    else if (NC == 5) bp_rows<5>(sp, sr, z0, z1, cbase, zl, lo, hi, wk, lk);
                    else if (NC == 4) bp_rows<4>(sp, sr, z0, z1, cbase, zl, lo, hi, wk, lk);
                    else bp_rows<3>(sp, sr, z0, z1, cbase, zl, lo, hi, wk, lk);
                }
                if constexpr (ONE) bp_update_x(fv, 0, xr, wk, lk, beta, mode, th, wsc);
                else { W[k] = wk; L[k] = lk; }
            }
        }
    }
    if constexpr (!ONE) {
        if (!active) return;
#pragma unroll
        for (int k = 0; k < ZR; ++k) {
            if (k >= nz) break;
            bp_update(f, vox_index<ZQ, TR>(d, ix, iy, iz0 + k), W[k], L[k], beta, mode, th, wsc);
        }
    }
}

template <typename T, typename SV, bool FAST>
__global__ __launch_bounds__(256) void k_voxel_bp(DGeo<T> d, const T *__restrict__ ang, SV *__restrict__ f,
                                                  const T *__restrict__ corr, int first, int stride, int nv,
                                                  T beta, int mode) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y * 4 + threadIdx.y;
    const int iz = blockIdx.z;
    if (ix >= d.M || iy >= d.N) return;
    T cx, cy, cz;
    if (!d.tb) {   // :592-594
        cx = d.bx + d.vx * (T)0.5 + (T)ix * d.vx;
        cy = d.by + d.vy * (T)0.5 + (T)iy * d.vy;
        cz = d.bz + d.vz * (T)0.5 + (T)iz * d.vz;
    } else {       // TB :566-568
        cx = d.bx + d.vx * ((T)0.5 + (T)ix);
        cy = d.by + d.vy * ((T)0.5 + (T)iy);
        cz = d.bz + d.vz * ((T)0.5 + (T)iz);
    }
    T W = 0, L = 0;
    // cylinder mask (:626-628)
    const T hx = (T)d.M * (T)0.5 - (T)0.5, hy = (T)d.N * (T)0.5 - (T)0.5;
    const T ma = (T)ix - hx - (T)d.moff;
    const T mb = (T)iy - hy;
    const bool outside = ma * d.vx * ma * d.vx + mb * d.vy * mb * d.vy >= hx * d.vx * hy * d.vy;
    if (!outside) {
        const int64_t P = (int64_t)d.R * d.ZP;
        for (int q = 0; q < nv; ++q) {
            const int v = first + q * stride;
            const T s = ang[2 * v], c = ang[2 * v + 1];
            for (int src = 0; src < d.NS; ++src) {
                const T szs = d.tb ? (T)(d.sz0 + (T)src * d.sdz) : d.sz0;
                const T *cimg = corr + ((int64_t)q * d.NS + src) * P;
                if constexpr (FAST) {
                    if (d.tb) {
                        float w = 0, l = 0;
                        voxel_wl_fast(d, cx, cy, cz, s, c, szs, cimg, w, l);
                        W += w; L += l;
                    } else {
                        voxel_wl_fast(d, cx, cy, cz, s, c, szs, cimg, W, L);
                    }
                } else if (d.tb) {
                    T w = 0, l = 0;
                    voxel_wl(d, cx, cy, cz, s, c, szs, cimg, w, l);
                    W += w; L += l;
                } else {
                    voxel_wl(d, cx, cy, cz, s, c, szs, cimg, W, L);
                }
            }
        }
    }
    const int64_t idx = ((int64_t)iz * d.N + iy) * d.M + ix;
    bool upd;
    if (!d.tb) upd = !(L < d.vx * (T)1e-6);              // :799
    else upd = (double)L > (double)d.vx * 1e-3;          // TB :772
    if (!upd) return;
    if (mode == 0) St<SV>::st(f, idx, W);
    else {
        const T x = (T)ldf(f, idx);
        // fast mode stores 16-bit updates with stochastic rounding, as the
        // LDS-staged and planar kernels do (exact mode rounds to nearest)
        if constexpr (FAST) st_sr(f, idx, x + beta * W / L, sr_bits((float)W, (float)L));
        else St<SV>::st(f, idx, x + beta * W / L);
    }
}

// Fan- and parallel-beam update with one detector row per slice (the
// reference's setup for CT_style 1/2: Z_prj = ZETA, volumn_z =
// Detector_pixel_x).  A voxel's rays all run in its own slice (the source
// sits at the voxel's height, :715-724), so its chord through each footprint
// column is the x/y slab interval times |S - P| in the plane, the same for
// every slice of the (x, y) column, and only the correction row changes with
// z.  A lane owns a column of PZR voxels: per view it forms the column range
// (the reference's corner projection without magnification in z) and the
// chords once, then each voxel adds chord x correction of its own row.  The
// sums and the threshold are those of voxel_wl_fast with the z slab
// unconstrained (the row lies at the voxel's height).
constexpr int PZR = 16;
template <typename SV>
__global__ __launch_bounds__(256) void k_voxel_bp_planar(DGeo<float> d, const float *__restrict__ ang,
                                                         SV *__restrict__ f, const float *__restrict__ corr, int first,
                                                         int stride, int nv, float beta, int mode) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y * 4 + threadIdx.y;
    const int iz0 = blockIdx.z * PZR;
    const BpThresh th(d, 1.0f);
    if (ix >= d.M || iy >= d.N) return;
    const int nz = min(PZR, d.Z - iz0);
    // cylinder mask (:626-628)
    const float hx = (float)d.M * 0.5f - 0.5f, hy = (float)d.N * 0.5f - 0.5f;
    const float ma = (float)ix - hx - (float)d.moff, mb = (float)iy - hy;
    if (ma * d.vx * ma * d.vx + mb * d.vy * mb * d.vy >= hx * d.vx * hy * d.vy) return;
    const float cx = d.bx + d.vx * 0.5f + (float)ix * d.vx;
    const float cy = d.by + d.vy * 0.5f + (float)iy * d.vy;
    const float hvx = 0.5f * d.vx, hvy = 0.5f * d.vy;
    const float xl = cx - hvx, xh = cx + hvx, yl = cy - hvy, yh = cy + hvy;
    const float thr = d.vx * 1e-6f;
    const int64_t P = (int64_t)d.R * d.ZP;
    float W[PZR];
#pragma unroll
    for (int k = 0; k < PZR; ++k) W[k] = 0.0f;
    float L = 0.0f;
    for (int q = 0; q < nv; ++q) {
        const int v = first + q * stride;
        const float s = ang[2 * v], c = ang[2 * v + 1];
        float minY = __builtin_inff(), maxY = -__builtin_inff();
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
                const float px = ii ? xh : xl, py = jj ? yh : yl;
                const float t = px * c + py * s;
                const float sv = -px * s + py * c;
                const float u = d.style == CT_PARALLEL ? sv : fmaf((t - d.dod) * frcp(d.dso - t), sv - d.sy, sv);
                minY = fminf(minY, u); maxY = fmaxf(maxY, u);
            }
        int y0 = (int)floorf(fmaf(minY - d.ymin, d.idu, 0.5f));
        int y1 = (int)floorf(fmaf(maxY - d.ymin, d.idu, 0.5f));
        if ((y0 < 0 && y1 < 0) || (y0 > d.R - 1 && y1 > d.R - 1)) continue;
        y0 = max(y0, 0); y1 = min(y1, d.R - 1);
        const float sxw = d.sx * c - d.sy * s, syw = d.sx * s + d.sy * c;
        const float *cq = corr + (int64_t)q * P + (int64_t)iz0 * d.R;
        for (int i = y0; i <= y1; ++i) {
            const float u = d.ymin + (float)i * d.du;
            const float qx = d.dod * c - u * s, qy = d.dod * s + u * c;
            const float ex = (d.style == CT_PARALLEL ? d.sx * c - u * s : sxw) - qx;
            const float ey = (d.style == CT_PARALLEL ? d.sx * s + u * c : syw) - qy;
            const float rx = frcp(ex), ry = frcp(ey);
            const float ax0 = (xl - qx) * rx, ax1 = (xh - qx) * rx;
            const float ay0 = (yl - qy) * ry, ay1 = (yh - qy) * ry;
            const float span = fminf(fmaxf(ax0, ax1), fmaxf(ay0, ay1)) - fmaxf(fminf(ax0, ax1), fminf(ay0, ay1));
            const float len = fsqrt(fmaf(ex, ex, ey * ey)) * span;
            if (!(span > 0.0f && len > thr)) continue;
            L += len;
#pragma unroll
            for (int k = 0; k < PZR; ++k)
                if (k < nz) W[k] = fmaf(len, cq[(int64_t)k * d.R + i], W[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < PZR; ++k)
        if (k < nz) bp_update(f, ((int64_t)(iz0 + k) * d.N + iy) * d.M + ix, W[k], L, beta, mode, th, 1.0f);
}

// Fan or parallel beam with the detector rows at the slice centres.
static bool planar_eligible(const ct_context *ctx, const DGeo<float> &d) {
    return ctx->math == CT_MATH_FAST && d.style != CT_CONE && d.ZP == d.Z &&
           std::fabs((double)d.dv - (double)d.vz) <= 1e-6 * (double)d.vz && d.NS == 1;
}

// Host-side bound for k_voxel_bp_lds: every tile footprint must fit the
// LDS window (LW columns, LH rows) and every voxel's bounding box NCMAX
// columns.  With rho the largest distance from the axis of a point of a tile
// that runs, the corner ratio r = (t - DOD)/(DSO - t) lies in [r(-rho), r(rho)]
// and changes by at most (DSO - DOD)/(DSO - rho)^2 per unit depth.  A box of
// (x, y) diagonal D and height H then projects to at most
// (1 + r_max)*D + dr*D*(rho + |off|) in u and
// (1 + r_max)*H + dr*D*max|z - z_source| in v; the floor(x + 0.5) bounds add
// one pixel and the footprint pad two more.
static bool bp_lds_fits(const DGeo<float> &d, int zr) {
    const double vx = d.vx, vy = d.vy, vz = d.vz, dso = d.dso, dod = d.dod;
    const double xs[2] = {(double)d.bx, (double)d.bx + d.M * vx}, ys[2] = {(double)d.by, (double)d.by + d.N * vy};
    double rho = 0;
    for (double x : xs)
        for (double y : ys) rho = std::max(rho, std::sqrt(x * x + y * y));
    // a tile runs only if it holds a voxel of the reconstruction cylinder, so
    // its points lie within the cylinder radius plus the tile diagonal
    const double DT = std::hypot(BPX * vx, BPY * vy), DV = std::hypot(vx, vy);
    const double hx = d.M * 0.5 - 0.5, hy = d.N * 0.5 - 0.5;
    rho = std::min(rho, std::sqrt(std::max(hx * vx * hy * vy, 0.0)) + DT);
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
    return cols(DT) + 4.0 <= LW && rows(DT, zr * vz) + 4.0 <= LH && cols(DV) + 2.0 <= NCMAX;
}

// per-view tables, then the LDS-staged update (fast FP32 arithmetic)
template <typename SV, bool ZQ, bool TR = false>
static ct_status run_bp_lds(ct_context *ctx, const DGeo<float> &d, const float *ang, void *vol, const void *corr,
                            int first, int stride, int nv, double beta, int mode) {
    ct_status st;
    const int nimg = nv * d.NS;
    const size_t P = (size_t)d.R * d.ZP;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pix_b = up(sizeof(float2) * P * nimg), col_b = up(sizeof(float4) * (size_t)d.R * nv);
    const size_t row_b = up(sizeof(float4) * (size_t)d.ZP * nimg);
    char *tab = (char *)slot(ctx, SLOT_BPTAB, pix_b + col_b + row_b, &st);
    if (!tab) return st;
    float2 *pix = (float2 *)tab;
    float4 *colT = (float4 *)(tab + pix_b);
    float4 *rowT = (float4 *)(tab + pix_b + col_b);
    ProfScope ps(ctx, CT_K_VOXEL_BP);
    // reduced-precision storage takes the packed-half chord engine
    constexpr bool H = !std::is_same<SV, float>::value;
    hipLaunchKernelGGL(k_bp_prep<H>, dim3((d.R + 255) / 256, d.ZP, nimg), dim3(256), 0, ctx->stream, d, ang,
                       (const float *)corr, first, stride, pix, colT, rowT);
    const int zr = nimg == 1 ? zrun<true>() : zrun<false>();
    dim3 lgrid((d.M + BPX - 1) / BPX, (d.N + BPY - 1) / BPY, (d.Z + zr - 1) / zr);
    if (nimg == 1)
        hipLaunchKernelGGL((k_voxel_bp_lds<SV, true, H, ZQ, TR>), lgrid, dim3(BPX, BPY, 1), 0, ctx->stream, d, ang,
                           (SV *)vol, (const float2 *)pix, (const float4 *)colT, (const float4 *)rowT, first, stride,
                           nv, (float)beta, mode);
    else
        hipLaunchKernelGGL((k_voxel_bp_lds<SV, false, H, ZQ, TR>), lgrid, dim3(BPX, BPY, 1), 0, ctx->stream, d, ang,
                           (SV *)vol, (const float2 *)pix, (const float4 *)colT, (const float4 *)rowT, first, stride,
                           nv, (float)beta, mode);
    return launch_check(ctx, "voxel backprojection");
}

static bool lds_eligible(const ct_context *ctx, const DGeo<float> &d, int nimg) {
    if (d.style != CT_CONE) return false; // the footprint tables are the cone-beam projection
    // the per-lane voxel offset (vox_lane, in bytes) must fit 32 bits: one
    // z-quad slab of up to 4-byte voxels
    if ((int64_t)d.M * d.N * 4 * 4 > (int64_t)UINT32_MAX) return false;
    return ctx->math == CT_MATH_FAST && nimg <= 65535 && bp_lds_fits(d, nimg == 1 ? zrun<true>() : zrun<false>());
}

template <typename T, typename SV>
static ct_status run_bp(ct_context *ctx, const ct_geometry &g, void *vol, const void *corr, int first, int stride,
                        int nv, double beta, int mode) {
    ct_status st;
    const T *ang = (const T *)angle_table(ctx, g, sizeof(T) == 8, false, &st);
    if (!ang) return st;
    DGeo<T> d = derive<T>(g);
    dim3 block(64, 4, 1);
    dim3 grid((d.M + 63) / 64, (d.N + 3) / 4, d.Z);
    if constexpr (std::is_same<T, float>::value) {
        if (lds_eligible(ctx, d, nv * d.NS))
            return run_bp_lds<SV, false>(ctx, d, ang, vol, corr, first, stride, nv, beta, mode);
        if (planar_eligible(ctx, d)) {
            ProfScope ps(ctx, CT_K_VOXEL_BP);
            dim3 pgrid((d.M + 63) / 64, (d.N + 3) / 4, (d.Z + PZR - 1) / PZR);
            hipLaunchKernelGGL((k_voxel_bp_planar<SV>), pgrid, block, 0, ctx->stream, d, ang, (SV *)vol,
                               (const float *)corr, first, stride, nv, (float)beta, mode);
            return launch_check(ctx, "voxel backprojection");
        }
        if (ctx->math == CT_MATH_FAST) {
            ProfScope ps(ctx, CT_K_VOXEL_BP);
            hipLaunchKernelGGL((k_voxel_bp<T, SV, true>), grid, block, 0, ctx->stream, d, ang, (SV *)vol,
                               (const T *)corr, first, stride, nv, (T)beta, mode);
            return launch_check(ctx, "voxel backprojection");
        }
    }
    ProfScope ps(ctx, CT_K_VOXEL_BP);
    hipLaunchKernelGGL((k_voxel_bp<T, SV, false>), grid, block, 0, ctx->stream, d, ang, (SV *)vol, (const T *)corr,
                       first, stride, nv, (T)beta, mode);
    return launch_check(ctx, "voxel backprojection");
}

// Whether the sweep can run in the z-quad layout for subsets of up to
// `subset` views: fast FP32 arithmetic, reduced or FP32 storage under 4 GiB
// (the correction's 32-bit buffer offsets) and an LDS-staged update.
bool sweep_zq_eligible(const ct_context *ctx, const ct_geometry &g, ct_dtype vdt, int subset) {
    if (vdt == CT_F64 || ctx->math != CT_MATH_FAST) return false;
    const DGeo<float> d = derive<float>(g);
    const uint64_t bytes = (uint64_t)d.M * d.N * (uint64_t)((d.Z + 3) & ~3) * dtype_size(vdt);
    if (bytes >= (1ull << 32)) return false;
    for (int nv : {1, subset})
        if (!lds_eligible(ctx, d, nv * d.NS)) return false;
    return true;
}

// SART update of a z-quad-layout volume (sweep_zq_eligible holds)
ct_status launch_voxel_bp_zq(ct_context *ctx, const ct_geometry &g, void *volq, ct_dtype vdt, const void *corr,
                             int32_t first, int32_t stride, int32_t nv, double beta, bool tr) {
    if (nv <= 0) return CT_OK;
    ct_status st;
    const float *ang = (const float *)angle_table(ctx, g, false, false, &st);
    if (!ang) return st;
    const DGeo<float> d = derive<float>(g);
    switch (vdt) {
    case CT_F32:
        return tr ? run_bp_lds<float, true, true>(ctx, d, ang, volq, corr, first, stride, nv, beta, 1)
                  : run_bp_lds<float, true>(ctx, d, ang, volq, corr, first, stride, nv, beta, 1);
    case CT_F16:
        return tr ? run_bp_lds<__half, true, true>(ctx, d, ang, volq, corr, first, stride, nv, beta, 1)
                  : run_bp_lds<__half, true>(ctx, d, ang, volq, corr, first, stride, nv, beta, 1);
    case CT_BF16:
        return tr ? run_bp_lds<__hip_bfloat16, true, true>(ctx, d, ang, volq, corr, first, stride, nv, beta, 1)
                  : run_bp_lds<__hip_bfloat16, true>(ctx, d, ang, volq, corr, first, stride, nv, beta, 1);
    default: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "z-quad update: unsupported volume dtype");
}

ct_status launch_voxel_bp(ct_context *ctx, const ct_geometry &g, void *vol, ct_dtype vdt, const void *corr,
                          int32_t first, int32_t stride, int32_t nv, double beta, int mode) {
    if (nv <= 0) return CT_OK;
    switch (vdt) {
    case CT_F32: return run_bp<float, float>(ctx, g, vol, corr, first, stride, nv, beta, mode);
    case CT_F16: return run_bp<float, __half>(ctx, g, vol, corr, first, stride, nv, beta, mode);
    case CT_BF16: return run_bp<float, __hip_bfloat16>(ctx, g, vol, corr, first, stride, nv, beta, mode);
    case CT_F64: return run_bp<double, double>(ctx, g, vol, corr, first, stride, nv, beta, mode);
    }
    return fail(ctx, CT_ERR_DTYPE, "unsupported volume dtype");
}

} // namespace ct
