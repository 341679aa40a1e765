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
            const T u = ratio * (sv - d.sy) + sv;
            const T w0 = ratio * (pz0 - szs) + pz0;
            const T w1 = ratio * (pz1 - szs) + pz1;
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

    const T sxw = d.sx * c - d.sy * s;
    const T syw = d.sx * s + d.sy * c;
    const T szw = szs;
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
                const T ex = sxw - qx, ey = syw - qy;
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
                const T ex = sxw - qx, ey = syw - qy;
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
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

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
            const float u = fmaf(ratio, sv - d.sy, sv);
            const float w0 = fmaf(ratio, zl - szs, zl);
            const float w1 = fmaf(ratio, zh - szs, zh);
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
                const float ex = sxw - qx, ey = syw - qy;
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
            const float ex = sxw - qx, ey = syw - qy;
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
        St<SV>::st(f, idx, x + beta * W / L);
    }
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
    ProfScope ps(ctx, CT_K_VOXEL_BP);
    if constexpr (std::is_same<T, float>::value) {
        if (ctx->math == CT_MATH_FAST) {
            hipLaunchKernelGGL((k_voxel_bp<T, SV, true>), grid, block, 0, ctx->stream, d, ang, (SV *)vol,
                               (const T *)corr, first, stride, nv, (T)beta, mode);
            return launch_check(ctx, "voxel backprojection");
        }
    }
    hipLaunchKernelGGL((k_voxel_bp<T, SV, false>), grid, block, 0, ctx->stream, d, ang, (SV *)vol, (const T *)corr,
                       first, stride, nv, (T)beta, mode);
    return launch_check(ctx, "voxel backprojection");
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
