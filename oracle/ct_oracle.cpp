/*
 * ct_oracle.cpp - CPU restatement of the reference's CT reconstruction kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the reported
 * CPU baseline: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product (gpu_amd/libctcore.so) never links,
 * loads or calls it.
 *
 * Every function restates one reference routine in plain C++ (templated on
 * float / double), keeping the reference's operation order so that the FP32
 * instantiation reproduces the reference arithmetic (IEEE, no contraction:
 * built with -ffp-contract=off).  Paths are relative to the reference root:
 *   V63 = CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823
 *   TB  = TBCT/TBCT_1024V_1024P_v4.2_20171012
 *
 * Pinning (checked in tests/test_oracle.py): Siddon vs an independent FP64
 * brute-force "sort every plane crossing" traversal and vs analytic ray-box
 * chords; voxel backprojection vs its closed forms (c = const -> dx = beta*c,
 * c = 0 -> no change) and vs the Siddon lengths (sum l per ray); SART vs the
 * inverse-crime fixed point; matched adjoint by the dot-product identity;
 * FISTA vs the textbook recurrence; FGP D/D^T by the dot-product identity,
 * constant input, unit-ball dual bound and lambda->0; norms vs numpy closed
 * forms; TV_eps gradient vs FP64 central differences; FFT/ramp vs numpy.fft;
 * FDK weight vs its formula and FDK orientation on an asymmetric phantom.
 */
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <complex>
#include <algorithm>
#include <limits>
#include <type_traits>
#include <omp.h>

#include "../include/ct_core.h"

namespace {

template <typename T> struct Geo {
    int M, N, Z, R, ZP, NS, moff, nviews, style;
    T vx, vy, vz, ivx, ivy, ivz, bx, by, bz;
    T du, dv, ymin, zmin, zmax, off, dso, dod, sx, sy, sz0, sdz;
    T us_rate, ang0, shift;
    bool tb;
};

/* Derived constants as V63/main_recon.cu:61-111 computes them (in T). */
template <typename T> Geo<T> derive(const ct_geometry *g) {
    Geo<T> d;
    d.M = g->nx; d.N = g->ny; d.Z = g->nz; d.R = g->ncols; d.ZP = g->nrows;
    d.NS = g->nsources < 1 ? 1 : g->nsources; d.moff = g->x_offset; d.nviews = g->nviews;
    d.tb = d.NS > 1; d.style = g->style;
    d.vx = (T)g->vx; d.vy = (T)g->vy; d.vz = (T)g->vz;
    d.ivx = (T)(1.0 / (double)d.vx); d.ivy = (T)(1.0 / (double)d.vy); d.ivz = (T)(1.0 / (double)d.vz);
    d.bx = -d.vx * ((T)d.M / (T)2 + (T)d.moff);
    d.by = -d.vy * ((T)d.N / (T)2);
    d.bz = -d.vz * ((T)d.Z / (T)2);
    d.du = (T)g->du; d.dv = (T)g->dv; d.off = (T)g->det_offset;
    d.ymin = -d.du * ((T)d.R / (T)2 - (T)0.5) + d.off;
    d.zmin = -d.dv * ((T)d.ZP / (T)2 - (T)0.5);
    d.zmax = +d.dv * ((T)d.ZP / (T)2 - (T)0.5);
    d.dso = (T)g->dso; d.dod = (T)g->dod;
    d.sx = d.dso; d.sy = d.off;
    d.sz0 = (T)g->src_z0; d.sdz = (T)g->src_dz;
    d.us_rate = (T)g->angle_step_deg; d.ang0 = (T)g->angle0_deg; d.shift = (T)g->angle_shift_deg;
    return d;
}

template <typename T> T pi_c();
template <> float pi_c<float>() { return 3.141592653589793f; }
template <> double pi_c<double>() { return 3.141592653589793; }

/* View angle, V63/host_IterativeRecon_CBCT.c:34-41 (iterative, -theta) and
 * V63/host_FDK.cpp:35-40 (FDK, +theta). */
template <typename T> void view_angle(const Geo<T> &d, int j, bool fdk, T &s, T &c) {
    T th = (T)((T)j * d.us_rate + d.ang0);
    th = (th + d.shift) * pi_c<T>() / (T)180;
    if (!fdk) th = -th;
    s = std::sin(th); c = std::cos(th);
}

template <typename T> inline T source_z(const Geo<T> &d, int src) {
    return d.tb ? (T)(d.sz0 + (T)src * d.sdz) : d.sz0;
}

/* Incremental Siddon traversal of the ray detector pixel (col,row) -> source,
 * forward_ray_driven_3d_kernel_correction, V63/kernel_IterativeRecon_universal.cu:14-551
 * (TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:1-545 adds the source axis).
 * visit(voxel_index, segment_length) is called per segment in traversal order;
 * returns the summed length. */
template <typename T, typename F>
T siddon_ray(const Geo<T> &d, T s, T c, int col, int row, int src, F &&visit) {
    const int M = d.M, N = d.N, Z = d.Z;
    const T inf = std::numeric_limits<T>::infinity();
    T u = d.ymin + (T)col * d.du;
    T x1x = d.dod * c - u * s;
    T x1y = d.dod * s + u * c;
    T x1z = d.zmin + (T)row * d.dv;
    // source per CT_style, V63/kernel_IterativeRecon_universal.cu:22-39
    T x2x, x2y, x2z;
    if (d.style == 2) {
        x2x = d.sx * c - u * s;
        x2y = d.sx * s + u * c;
    } else {
        x2x = d.sx * c - d.sy * s;
        x2y = d.sx * s + d.sy * c;
    }
    x2z = d.style == 0 ? source_z(d, src) : x1z;
    T ixd = (T)1 / (x2x - x1x);
    T iyd = (T)1 / (x2y - x1y);
    T izd = (T)1 / (x2z - x1z);

    T amin, amax;
    amin = (d.bx - x1x) * ixd;
    amax = (d.bx + d.vx * (T)M - x1x) * ixd;
    T axmin = std::fmin(amin, amax), axmax = std::fmax(amin, amax);
    amin = (d.by - x1y) * iyd;
    amax = (d.by + d.vy * (T)N - x1y) * iyd;
    T aymin = std::fmin(amin, amax), aymax = std::fmax(amin, amax);
    amin = (d.bz - x1z) * izd;
    amax = (d.bz + d.vz * (T)Z - x1z) * izd;
    T azmin = std::fmin(amin, amax), azmax = std::fmax(amin, amax);
    amin = std::fmax(std::fmax(axmin, aymin), std::fmax(aymin, azmin));
    amax = std::fmin(std::fmin(axmax, aymax), std::fmin(aymax, azmax));

    T len = 0;
    if (amax <= amin) return len;

    int imin, imax, jmin, jmax, kmin, kmax;
    if (x1x < x2x) {
        imin = (amin == axmin) ? 1 : (int)std::floor((amin * (x2x - x1x) + x1x - d.bx) * d.ivx) + 1;
        imax = (amax == axmax) ? M : (int)std::floor((amax * (x2x - x1x) + x1x - d.bx) * d.ivx);
    } else {
        imax = (amin == axmin) ? M - 1 : (int)std::floor((amin * (x2x - x1x) + x1x - d.bx) * d.ivx);
        imin = (amax == axmax) ? 0 : (int)std::floor((amax * (x2x - x1x) + x1x - d.bx) * d.ivx) + 1;
    }
    if (x1y < x2y) {
        jmin = (amin == aymin) ? 1 : (int)std::floor((amin * (x2y - x1y) + x1y - d.by) * d.ivy) + 1;
        jmax = (amax == aymax) ? N : (int)std::floor((amax * (x2y - x1y) + x1y - d.by) * d.ivy);
    } else {
        jmax = (amin == aymin) ? N - 1 : (int)std::floor((amin * (x2y - x1y) + x1y - d.by) * d.ivy);
        jmin = (amax == aymax) ? 0 : (int)std::floor((amax * (x2y - x1y) + x1y - d.by) * d.ivy) + 1;
    }
    const bool zpar = std::fabs(x1z - x2z) < d.vz * (T)1e-6;
    if (zpar) {
        kmin = (int)std::floor((x1z - d.bz) * d.ivz) + 1;
        kmax = (int)std::floor((x1z - d.bz) * d.ivz);
    } else if (x1z < x2z) {
        kmin = (amin == azmin) ? 1 : (int)std::floor((amin * (x2z - x1z) + x1z - d.bz) * d.ivz) + 1;
        kmax = (amax == azmax) ? Z : (int)std::floor((amax * (x2z - x1z) + x1z - d.bz) * d.ivz);
    } else {
        kmax = (amin == azmin) ? Z - 1 : (int)std::floor((amin * (x2z - x1z) + x1z - d.bz) * d.ivz);
        kmin = (amax == azmax) ? 0 : (int)std::floor((amax * (x2z - x1z) + x1z - d.bz) * d.ivz) + 1;
    }

    /* initial plane indices and next crossings (:218-266) */
    int i = 0, j = 0, k = 0;
    T ax, ay, az;
    if (std::fabs(x1x - x2x) < d.vx * (T)1e-6) { ax = inf; i = imin - 1; }
    else if (x1x < x2x) { ax = (d.vx * (T)imin + d.bx - x1x) * ixd; i = imin - 1; }
    else { ax = (d.vx * (T)imax + d.bx - x1x) * ixd; i = imax + 1; }
    if (std::fabs(x1y - x2y) < d.vy * (T)1e-6) { ay = inf; j = jmin - 1; }
    else if (x1y < x2y) { ay = (d.vy * (T)jmin + d.by - x1y) * iyd; j = jmin - 1; }
    else { ay = (d.vy * (T)jmax + d.by - x1y) * iyd; j = jmax + 1; }
    if (zpar) { az = inf; k = kmin - 1; }
    else if (x1z < x2z) { az = (d.vz * (T)kmin + d.bz - x1z) * izd; k = kmin - 1; }
    else { az = (d.vz * (T)kmax + d.bz - x1z) * izd; k = kmax + 1; }
    /* (:271-284) */
    int vi = (x1x < x2x) ? imin - 1 : imax;
    int vj = (x1y < x2y) ? jmin - 1 : jmax;
    int vk = (x1z < x2z) ? kmin - 1 : kmax;

    T dist = std::sqrt((x2x - x1x) * (x2x - x1x) + (x2y - x1y) * (x2y - x1y) + (x2z - x1z) * (x2z - x1z));
    T ac = amin;
    const int sgx = (x1x < x2x) ? 1 : ((x1x > x2x) ? -1 : 0);
    const int sgy = (x1y < x2y) ? 1 : ((x1y > x2y) ? -1 : 0);
    const int sgz = (x1z < x2z) ? 1 : ((x1z > x2z) ? -1 : 0);
    const T lim = std::is_same<T, float>::value ? (T)1e-16f : (T)1e-16;

    /* step loop (:297-551): the seven branches of the reference collapse to
     * "step every axis whose next crossing equals the minimum" */
    /* a valid ray crosses at most M+N+Z+1 voxels; the cap only guarantees
     * termination for non-finite input */
    int guard = M + N + Z + 4;
    while (amax - ac > lim && guard-- > 0) {
        if (vi > M - 1 || vi < 0 || vj > N - 1 || vj < 0 || vk > Z - 1 || vk < 0) break;
        bool sx_ = (ax < ay) && (ax < az);
        bool sy_ = (ay < ax) && (ay < az);
        bool sz_ = (az < ax) && (az < ay);
        bool sxy = (ax == ay) && (ax < az);
        bool sxz = (ax == az) && (ax < ay);
        bool syz = (ay == az) && (ay < ax);
        bool sxyz = (ax == az) && (ax == ay);
        T anext;
        bool stepx = false, stepy = false, stepz = false;
        if (sx_) { anext = ax; stepx = true; }
        else if (sy_) { anext = ay; stepy = true; }
        else if (sz_) { anext = az; stepz = true; }
        else if (sxy) { anext = ax; stepx = stepy = true; }
        else if (sxz) { anext = ax; stepx = stepz = true; }
        else if (syz) { anext = ay; stepy = stepz = true; }
        else if (sxyz) { anext = ax; stepx = stepy = stepz = true; }
        else { break; /* unordered (NaN) crossings: the reference would spin; stop */ }
        T seg = dist * (anext - ac);
        len += seg;
        visit((int64_t)vk * M * N + (int64_t)vj * M + vi, seg);
        ac = anext;
        if (stepx) {
            int nxt = 0;
            if (sgx > 0) { i++; vi++; nxt = i + 1; }
            if (sgx < 0) { i--; vi--; nxt = i - 1; }
            ax = (d.vx * (T)nxt + d.bx - x1x) * ixd;
        }
        if (stepy) {
            int nxt = 0;
            if (sgy > 0) { j++; vj++; nxt = j + 1; }
            else if (sgy < 0) { j--; vj--; nxt = j - 1; }
            ay = (d.vy * (T)nxt + d.by - x1y) * iyd;
        }
        if (stepz) {
            int nxt = 0;
            if (sgz > 0) { k++; vk++; nxt = k + 1; }
            else if (sgz < 0) { k--; vk--; nxt = k - 1; }
            az = (d.vz * (T)nxt + d.bz - x1z) * izd;
        }
    }
    return len;
}

/* forward / correction for views [v0, v0+nv): V63/kernel_IterativeRecon_universal.cu:554-563 */
template <typename T>
void forward_impl(const Geo<T> &d, const T *f, T *out, const T *meas, int v0, int nv, int cmd) {
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t rays = (int64_t)nv * d.NS * d.ZP;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t rr = 0; rr < rays; ++rr) {
        int row = (int)(rr % d.ZP);
        int src = (int)((rr / d.ZP) % d.NS);
        int vv = (int)(rr / ((int64_t)d.ZP * d.NS));
        T s, c;
        view_angle(d, v0 + vv, false, s, c);
        T *o = out + ((int64_t)vv * d.NS + src) * P + (int64_t)row * d.R;
        const T *b = meas ? meas + ((int64_t)vv * d.NS + src) * P + (int64_t)row * d.R : nullptr;
        for (int col = 0; col < d.R; ++col) {
            T sum = 0;
            auto vis = [&](int64_t idx, T seg) { sum += seg * f[idx]; };
            T len = siddon_ray(d, s, c, col, row, src, vis);
            T r;
            if (len < d.vz * (T)1e-6) r = 0;
            else if (cmd == 0) r = sum;
            else r = (b[col] - sum) / len;
            o[col] = r;
        }
    }
}

/* Voxel-driven (W, L) for one voxel and one view/source,
 * V63/kernel_IterativeRecon_universal.cu:592-797 (CBCT) and
 * TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:566-752 (multi-source, with
 * the source culling of :618-627). */
template <typename T>
void voxel_wl(const Geo<T> &d, int ix, int iy, int iz, T s, T c, int src, const T *corr, T &W, T &L) {
    T cx, cy, cz;
    if (!d.tb) {
        cx = d.bx + d.vx * (T)0.5 + (T)ix * d.vx;
        cy = d.by + d.vy * (T)0.5 + (T)iy * d.vy;
        cz = d.bz + d.vz * (T)0.5 + (T)iz * d.vz;
    } else {
        cx = d.bx + d.vx * ((T)0.5 + (T)ix);
        cy = d.by + d.vy * ((T)0.5 + (T)iy);
        cz = d.bz + d.vz * ((T)0.5 + (T)iz);
    }
    const T szs = source_z(d, src);
    if (d.tb) {
        T tc = cx * c + cy * s;
        T upper = (cz + d.vz - d.zmin) * (d.dso - d.dod) / (tc - d.dod) + d.zmin;
        T lower = (cz - d.vz - d.zmax) * (d.dso - d.dod) / (tc - d.dod) + d.zmax;
        if (szs < lower || szs > upper) return;
    }
    const T inf = std::numeric_limits<T>::infinity();
    T minY = inf, minZ = inf, maxY = -inf, maxZ = -inf;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                T px = cx + (T)i * d.vx - (T)0.5 * d.vx;
                T py = cy + (T)j * d.vy - (T)0.5 * d.vy;
                T pz = cz + (T)k * d.vz - (T)0.5 * d.vz;
                T t = px * c + py * s;
                T sv = -px * s + py * c;
                // vertex on the detector per CT_style, :656-670
                T ud, wd;
                if (d.style == 0) {
                    ud = (t - d.dod) / (d.dso - t) * (sv - d.sy) + sv;
                    wd = (t - d.dod) / (d.dso - t) * (pz - szs) + pz;
                } else if (d.style == 1) {
                    ud = (t - d.dod) / (d.dso - t) * (sv - d.sy) + sv;
                    wd = cz;
                } else {
                    ud = sv;
                    wd = cz;
                }
                minY = std::fmin(minY, ud); maxY = std::fmax(maxY, ud);
                minZ = std::fmin(minZ, wd); maxZ = std::fmax(maxZ, wd);
            }
    const T idu = (T)1 / d.du;
    const T idv = (T)1 / d.dv;
    int y0 = (int)std::floor((minY - d.ymin) * idu + (T)0.5);
    int y1 = (int)std::floor((maxY - d.ymin) * idu + (T)0.5);
    int z0 = (int)std::floor((minZ - d.zmin) * idv + (T)0.5);
    int z1 = (int)std::floor((maxZ - d.zmin) * idv + (T)0.5);
    if ((y0 < 0 && y1 < 0) || (y0 > d.R - 1 && y1 > d.R - 1) || (z0 < 0 && z1 < 0) ||
        (z0 > d.ZP - 1 && z1 > d.ZP - 1))
        return;
    if (y0 <= 0) y0 = 0;
    if (y1 >= d.R - 1) y1 = d.R - 1;
    if (z0 <= 0) z0 = 0;
    if (z1 >= d.ZP - 1) z1 = d.ZP - 1;
    // source per CT_style, :709-724 (CT_PARALLEL: per pixel, :738-742)
    T sx = d.sx * c - d.sy * s;
    T sy = d.sx * s + d.sy * c;
    T sz = d.style == 0 ? szs : cz;
    for (int j = z0; j <= z1; j++)
        for (int i = y0; i <= y1; i++) {
            T u = d.ymin + (T)i * d.du;
            T qx = d.dod * c - u * s;
            T qy = d.dod * s + u * c;
            T qz = d.zmin + (T)j * d.dv;
            if (d.style == 2) {
                sx = d.sx * c - u * s;
                sy = d.sx * s + u * c;
            }
            T ax0 = ((cx - (T)0.5 * d.vx) - qx) / (sx - qx);
            T ax1 = ((cx + (T)0.5 * d.vx) - qx) / (sx - qx);
            T ay0 = ((cy - (T)0.5 * d.vy) - qy) / (sy - qy);
            T ay1 = ((cy + (T)0.5 * d.vy) - qy) / (sy - qy);
            T az0 = ((cz - (T)0.5 * d.vz) - qz) / (sz - qz);
            T az1 = ((cz + (T)0.5 * d.vz) - qz) / (sz - qz);
            T tx = std::fmin(ax0, ax1), ty = std::fmin(ay0, ay1), tz = std::fmin(az0, az1);
            T amin = std::fmax(std::fmax(tx, ty), std::fmax(ty, tz));
            tx = std::fmax(ax0, ax1); ty = std::fmax(ay0, ay1); tz = std::fmax(az0, az1);
            T amax = std::fmin(std::fmin(tx, ty), std::fmin(ty, tz));
            if (amax - amin > 0) {
                T dist = std::sqrt((sx - qx) * (sx - qx) + (sy - qy) * (sy - qy) + (sz - qz) * (sz - qz));
                T tmp = dist * (amax - amin);
                if (tmp > d.vx * (T)1e-6) {
                    W = W + tmp * corr[(int64_t)j * d.R + i];
                    L = L + tmp;
                }
            }
        }
}

/* Cylinder mask, V63/kernel_IterativeRecon_universal.cu:626-628 */
template <typename T> bool outside_fov(const Geo<T> &d, int ix, int iy) {
    T hx = (T)d.M * (T)0.5 - (T)0.5, hy = (T)d.N * (T)0.5 - (T)0.5;
    T a = (T)ix - hx - (T)d.moff;
    T b = (T)iy - hy;
    return a * d.vx * a * d.vx + b * d.vy * b * d.vy >= hx * d.vx * hy * d.vy;
}

/* Backprojection summed over views [v0,v0+nv) (and sources), then the update
 * of V63/kernel_IterativeRecon_universal.cu:799-807 (CBCT) or
 * TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:765-775 (multi-source). */
template <typename T>
void backproject_impl(const Geo<T> &d, T *f, const T *corr, int v0, int nv, T beta, int mode) {
    const int64_t P = (int64_t)d.R * d.ZP;
    std::vector<T> sn(nv), cs(nv);
    for (int v = 0; v < nv; ++v) view_angle(d, v0 + v, false, sn[v], cs[v]);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int iz = 0; iz < d.Z; ++iz)
        for (int iy = 0; iy < d.N; ++iy)
            for (int ix = 0; ix < d.M; ++ix) {
                T W = 0, L = 0;
                if (!outside_fov(d, ix, iy)) {
                    for (int v = 0; v < nv; ++v)
                        for (int s = 0; s < d.NS; ++s) {
                            if (d.tb) {
                                T w = 0, l = 0;
                                voxel_wl(d, ix, iy, iz, sn[v], cs[v], s, corr + ((int64_t)v * d.NS + s) * P, w, l);
                                W += w; L += l;
                            } else {
                                voxel_wl(d, ix, iy, iz, sn[v], cs[v], s, corr + ((int64_t)v * d.NS + s) * P, W, L);
                            }
                        }
                }
                int64_t idx = ((int64_t)iz * d.N + iy) * d.M + ix;
                if (!d.tb) {
                    if (L < d.vx * (T)1e-6) continue;
                    if (mode == 0) f[idx] = W;
                    else f[idx] += beta * W / L;
                } else {
                    if ((double)L > (double)d.vx * 1e-3) {
                        if (mode == 0) f[idx] = W;
                        else f[idx] += beta * W / L;
                    }
                }
            }
}

/* Subsets of a sweep (the build's choice; the reference updates view by view):
 * ascending: consecutive chunks of `subset` views; interleaved: with
 * n_sub = ceil(nviews/subset) subsets, subset s holds views s, s+n_sub, ... */
std::vector<std::vector<int>> subsets_of(int nviews, int subset, int order) {
    std::vector<std::vector<int>> out;
    if (subset < 1) subset = 1;
    if (order == CT_ORDER_INTERLEAVED && subset > 1) {
        int nsub = (nviews + subset - 1) / subset;
        for (int s = 0; s < nsub; ++s) {
            std::vector<int> g;
            for (int j = s; j < nviews; j += nsub) g.push_back(j);
            out.push_back(g);
        }
    } else {
        for (int p0 = 0; p0 < nviews; p0 += subset) {
            std::vector<int> g;
            for (int j = p0; j < std::min(nviews, p0 + subset); ++j) g.push_back(j);
            out.push_back(g);
        }
    }
    return out;
}

/* SART sweep, V63/host_IterativeRecon_CBCT.c:31-52 generalised to subsets. */
template <typename T>
void sart_sweep_impl(const Geo<T> &d, T *f, const T *meas, T beta, int subset, int order) {
    const int64_t P = (int64_t)d.R * d.ZP * d.NS;
    std::vector<std::vector<int>> subs = subsets_of(d.nviews, subset, order);
    std::vector<T> corr((size_t)P * (subset < 1 ? 1 : subset));
    for (const std::vector<int> &sv : subs) {
        const int ns = (int)sv.size();
        std::vector<int> ord = sv;
        const int p0 = 0;
        /* corrections of the subset's views from the current volume */
        for (int q = 0; q < ns; ++q)
            forward_impl(d, f, corr.data() + (int64_t)q * P, meas + (int64_t)ord[p0 + q] * P, ord[p0 + q], 1, 1);
        /* backprojection summed over the subset, one update */
        if (ns == 1) {
            backproject_impl(d, f, corr.data(), ord[p0], 1, beta, 1);
        } else {
            /* views of a subset need not be contiguous: accumulate by view */
            const int64_t V = (int64_t)d.M * d.N * d.Z;
            std::vector<T> sn(ns), cs(ns);
            for (int q = 0; q < ns; ++q) view_angle(d, ord[p0 + q], false, sn[q], cs[q]);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
            for (int iz = 0; iz < d.Z; ++iz)
                for (int iy = 0; iy < d.N; ++iy)
                    for (int ix = 0; ix < d.M; ++ix) {
                        T W = 0, L = 0;
                        if (!outside_fov(d, ix, iy))
                            for (int q = 0; q < ns; ++q)
                                for (int s = 0; s < d.NS; ++s) {
                                    const T *cc = corr.data() + (int64_t)q * P + (int64_t)s * d.R * d.ZP;
                                    if (d.tb) {
                                        T w = 0, l = 0;
                                        voxel_wl(d, ix, iy, iz, sn[q], cs[q], s, cc, w, l);
                                        W += w; L += l;
                                    } else {
                                        voxel_wl(d, ix, iy, iz, sn[q], cs[q], s, cc, W, L);
                                    }
                                }
                        int64_t idx = ((int64_t)iz * d.N + iy) * d.M + ix;
                        (void)V;
                        if (!d.tb) {
                            if (L < d.vx * (T)1e-6) continue;
                            f[idx] += beta * W / L;
                        } else if ((double)L > (double)d.vx * 1e-3) {
                            f[idx] += beta * W / L;
                        }
                    }
        }
    }
}

/* Matched adjoint: scatter of the forward's Siddon segments
 * (V63/pixel_driven_backprj/kernel_IterativeRecon_CBCT.cu:563-1116), with the
 * forward's ray-length cut (V63/kernel_IterativeRecon_universal.cu:554) so that
 * <A x, y> = <x, A^T y> exactly up to summation order.  Deterministic: each
 * thread scatters a fixed slice of views into a private volume; the private
 * volumes are summed in thread order. */
template <typename T>
void adjoint_impl(const Geo<T> &d, const T *proj, T *out, int v0, int nv) {
    const int64_t V = (int64_t)d.M * d.N * d.Z;
    const int64_t P = (int64_t)d.R * d.ZP;
    int nth = omp_get_max_threads();
    if (nth > nv * d.NS) nth = nv * d.NS;
    if (nth < 1) nth = 1;
    std::vector<std::vector<T>> priv(nth);
#pragma omp parallel num_threads(nth)
    {
        int t = omp_get_thread_num();
        priv[t].assign(V, (T)0);
        T *acc = priv[t].data();
        int64_t nimg = (int64_t)nv * d.NS;
        int64_t lo = nimg * t / nth, hi = nimg * (t + 1) / nth;
        for (int64_t im = lo; im < hi; ++im) {
            int vv = (int)(im / d.NS), src = (int)(im % d.NS);
            T s, c;
            view_angle(d, v0 + vv, false, s, c);
            const T *y = proj + im * P;
            for (int row = 0; row < d.ZP; ++row)
                for (int col = 0; col < d.R; ++col) {
                    T yv = y[(int64_t)row * d.R + col];
                    auto none = [](int64_t, T) {};
                    T len = siddon_ray(d, s, c, col, row, src, none);
                    if (len < d.vz * (T)1e-6) continue;
                    auto sc = [&](int64_t idx, T seg) { acc[idx] += seg * yv; };
                    siddon_ray(d, s, c, col, row, src, sc);
                }
        }
    }
#pragma omp parallel for
    for (int64_t i = 0; i < V; ++i) {
        T a = 0;
        for (int t = 0; t < nth; ++t) a += priv[t][i];
        out[i] = a;
    }
}

/* ---- FGP (V63/kernel_FGP_Denoise_GPUx7.cu + host_FGP_Denoise_GPUx7.c) ---- */
/* Dual fields are volume-sized: r[v] is the reference's staggered d_r[i+1]
 * (the x-face between v and v+e_x); r[v] = 0 on the last x plane. */
template <typename T>
void fgp_L(int M, int N, int Z, const T *b, const T *r, const T *s, const T *t, T *L, T lambda) {
    const int64_t MN = (int64_t)M * N;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int64_t v = k * MN + (int64_t)j * M + i;
                T rm = i > 0 ? r[v - 1] : (T)0;
                T sm = j > 0 ? s[v - M] : (T)0;
                T tm = k > 0 ? t[v - MN] : (T)0;
                T tl = (r[v] - rm) + (s[v] - sm) + (t[v] - tm);    /* :14 */
                L[v] = b[v] - lambda * tl;                         /* :17 */
            }
}

/* duals (optional): the projected dual fields (o, p, q) after the last
 * iteration, for the unit-ball check of the tests */
template <typename T>
void fgp_impl(int M, int N, int Z, const T *b, T *out, T lambda, int niter, T *duals = nullptr) {
    const int64_t V = (int64_t)M * N * Z, MN = (int64_t)M * N;
    if (lambda == 0) { std::memcpy(out, b, sizeof(T) * V); return; }
    std::vector<T> r(V, 0), s(V, 0), t(V, 0), o(V, 0), p(V, 0), q(V, 0), L(V, 0);
    T tk1 = 1, tk;
    const T a12 = (T)1 / (lambda * (T)12);   /* host_FGP_Denoise_GPUx7.c:76 */
    for (int it = 0; it < niter; ++it) {
        tk = ((T)1 + std::sqrt((T)1 + (T)4 * tk1 * tk1)) / (T)2;   /* :71 */
        fgp_L(M, N, Z, b, r.data(), s.data(), t.data(), L.data(), lambda);
        const T g = (tk1 - (T)1) / tk;
#pragma omp parallel for collapse(2)
        for (int k = 0; k < Z; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < M; ++i) {
                    int64_t v = k * MN + (int64_t)j * M + i;
                    T tr = r[v], ts = s[v], tt = t[v];               /* kernel :52-54 */
                    if (i < M - 1) tr = tr + a12 * (L[v] - L[v + 1]);
                    if (j < N - 1) ts = ts + a12 * (L[v] - L[v + M]);
                    if (k < Z - 1) tt = tt + a12 * (L[v] - L[v + MN]);
                    T nrm = std::sqrt(tr * tr + ts * ts + tt * tt);  /* :67 */
                    T den = std::fmax((T)1, nrm);
                    if (i < M - 1) { T on = tr / den; r[v] = on + g * (on - o[v]); o[v] = on; }
                    if (j < N - 1) { T pn = ts / den; s[v] = pn + g * (pn - p[v]); p[v] = pn; }
                    if (k < Z - 1) { T qn = tt / den; t[v] = qn + g * (qn - q[v]); q[v] = qn; }
                }
        tk1 = tk;
    }
    fgp_L(M, N, Z, b, o.data(), p.data(), q.data(), out, lambda);     /* :83 */
    if (duals) {
        std::memcpy(duals, o.data(), sizeof(T) * V);
        std::memcpy(duals + V, p.data(), sizeof(T) * V);
        std::memcpy(duals + 2 * V, q.data(), sizeof(T) * V);
    }
}

/* D: x -> (x_v - x_{v+e}) inside, 0 on the last plane; D^T: p -> p_v - p_{v-e}. */
template <typename T>
void fgp_D(int M, int N, int Z, const T *x, T *r, T *s, T *t) {
    const int64_t MN = (int64_t)M * N;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int64_t v = k * MN + (int64_t)j * M + i;
                r[v] = i < M - 1 ? x[v] - x[v + 1] : (T)0;
                s[v] = j < N - 1 ? x[v] - x[v + M] : (T)0;
                t[v] = k < Z - 1 ? x[v] - x[v + MN] : (T)0;
            }
}
template <typename T>
void fgp_Dt(int M, int N, int Z, const T *r, const T *s, const T *t, T *out) {
    const int64_t MN = (int64_t)M * N;
#pragma omp parallel for collapse(2)
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int64_t v = k * MN + (int64_t)j * M + i;
                T rv = i < M - 1 ? r[v] : (T)0, sv = j < N - 1 ? s[v] : (T)0, tv = k < Z - 1 ? t[v] : (T)0;
                T rm = i > 0 ? r[v - 1] : (T)0;
                T sm = j > 0 ? s[v - M] : (T)0;
                T tm = k > 0 ? t[v - MN] : (T)0;
                out[v] = (rv - rm) + (sv - sm) + (tv - tm);
            }
}

/* TV norm, V63/kernel_tool_functions.cu:39-65, summed in FP64. */
template <typename T> double tv_norm_impl(int M, int N, int Z, const T *f) {
    const int64_t MN = (int64_t)M * N;
    double acc = 0;
#pragma omp parallel for collapse(2) reduction(+ : acc)
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int64_t v = k * MN + (int64_t)j * M + i;
                double c = f[v];
                bool hx = i < M - 1, hy = j < N - 1, hz = k < Z - 1;
                double dx = hx ? c - (double)f[v + 1] : 0.0;
                double dy = hy ? c - (double)f[v + M] : 0.0;
                double dz = hz ? c - (double)f[v + MN] : 0.0;
                acc += std::sqrt(dx * dx + dy * dy + dz * dz);
            }
    return acc;
}

/* TV_eps with backward differences that exist (definition of
 * V63/kernel_IterativeRecon_universal.cu:912-949 extended to the boundary). */
template <typename T> double tv_eps_impl(int M, int N, int Z, const T *f, double eps) {
    const int64_t MN = (int64_t)M * N;
    double acc = 0;
#pragma omp parallel for collapse(2) reduction(+ : acc)
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int64_t v = k * MN + (int64_t)j * M + i;
                double c = f[v], q = eps;
                if (i > 0) { double e = c - (double)f[v - 1]; q += e * e; }
                if (j > 0) { double e = c - (double)f[v - M]; q += e * e; }
                if (k > 0) { double e = c - (double)f[v - MN]; q += e * e; }
                acc += std::sqrt(q);
            }
    return acc;
}
template <typename T> void tv_eps_grad_impl(int M, int N, int Z, const T *f, double eps, T *g) {
    const int64_t MN = (int64_t)M * N;
    auto phi = [&](int i, int j, int k) {
        int64_t v = k * MN + (int64_t)j * M + i;
        double c = f[v], q = eps;
        if (i > 0) { double e = c - (double)f[v - 1]; q += e * e; }
        if (j > 0) { double e = c - (double)f[v - M]; q += e * e; }
        if (k > 0) { double e = c - (double)f[v - MN]; q += e * e; }
        return std::sqrt(q);
    };
#pragma omp parallel for collapse(2)
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                int64_t v = k * MN + (int64_t)j * M + i;
                double c = f[v];
                double num = 0;
                if (i > 0) num += c - (double)f[v - 1];
                if (j > 0) num += c - (double)f[v - M];
                if (k > 0) num += c - (double)f[v - MN];
                double gv = num / phi(i, j, k);
                if (i < M - 1) gv -= ((double)f[v + 1] - c) / phi(i + 1, j, k);
                if (j < N - 1) gv -= ((double)f[v + M] - c) / phi(i, j + 1, k);
                if (k < Z - 1) gv -= ((double)f[v + MN] - c) / phi(i, j, k + 1);
                g[v] = (T)gv;
            }
}

/* ---- FDK (V63/host_FDK.cpp, V63/FDK/kernel_FDK.cu, V63/FDK/FFT_1d_GPU.h) ---- */
/* Cosine weight, V63/FDK/kernel_FDK.cu:1-14; multi-source images
 * [view][source] measure rows from the source's height,
 * TB/FDK/kernel_FDK.cu:1-17. */
template <typename T> void fdk_weight_impl(const Geo<T> &d, const T *in, T *out) {
    const int64_t P = (int64_t)d.R * d.ZP;
#pragma omp parallel for collapse(2)
    for (int v = 0; v < d.nviews * d.NS; ++v)
        for (int z = 0; z < d.ZP; ++z)
            for (int x = 0; x < d.R; ++x) {
                T x1 = d.dso;
                T x2 = (d.ymin + (T)x * d.du) * d.dso / (d.dso - d.dod);
                T x3 = d.tb ? (d.zmin + (T)z * d.dv - source_z(d, v % d.NS)) * d.dso / (d.dso - d.dod)
                            : (d.zmin + (T)z * d.dv) * d.dso / (d.dso - d.dod);
                int64_t idx = (int64_t)v * P + (int64_t)z * d.R + x;
                out[idx] = in[idx] * d.dso / std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
            }
}

/* Textbook in-place radix-2 Cooley-Tukey (FP64), sign -1 forward / +1 inverse
 * (unnormalised); non-power-of-two sizes use a direct DFT. */
void fft_inplace(std::vector<std::complex<double>> &a, int sign) {
    const size_t n = a.size();
    if (n & (n - 1)) {
        std::vector<std::complex<double>> out(n);
        for (size_t k = 0; k < n; ++k) {
            std::complex<double> acc = 0;
            for (size_t m = 0; m < n; ++m) {
                double ang = sign * 2.0 * M_PI * (double)((k * m) % n) / (double)n;
                acc += a[m] * std::complex<double>(std::cos(ang), std::sin(ang));
            }
            out[k] = acc;
        }
        a = out;
        return;
    }
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        double ang = sign * 2.0 * M_PI / (double)len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                std::complex<double> w(std::cos(ang * (double)k), std::sin(ang * (double)k));
                std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

/* Filter H[k] of V63/FDK/FFT_1d_GPU.h:243-269 (in float as the reference
 * computes it), optionally Hann-windowed. */
std::vector<double> ramp_filter_taps(const ct_geometry *g, int window) {
    const int pad = 2 * g->ncols;
    float du = (float)g->du, dso = (float)g->dso, dod = (float)g->dod;
    float fmax_ = (float)(1.0 / 2.0 / (double)(du * dso / (dso - dod)));
    float fstep = fmax_ / (float)(pad / 2);
    std::vector<double> H(pad);
    for (int i = 0; i < pad; ++i) {
        int k = i >= pad / 2 ? i - pad : i;
        float h = (float)std::abs(k) * fstep;
        h *= 0.5f;
        double hv = h;
        if (window == CT_WINDOW_HANN) hv *= 0.5 * (1.0 + std::cos(2.0 * M_PI * (double)k / (double)pad));
        H[i] = hv;
    }
    return H;
}

template <typename T>
void ramp_impl(const ct_geometry *g, const T *in, T *out, int nimg, int window) {
    const int R = g->ncols, ZP = g->nrows, pad = 2 * R;
    std::vector<double> H = ramp_filter_taps(g, window);
    const int64_t rows = (int64_t)nimg * ZP;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; ++r) {
        std::vector<std::complex<double>> a(pad, 0.0);
        for (int x = 0; x < R; ++x) a[x] = (double)in[r * R + x];
        fft_inplace(a, -1);
        for (int k = 0; k < pad; ++k) a[k] *= H[k];
        fft_inplace(a, +1);
        for (int x = 0; x < R; ++x) out[r * R + x] = (T)(a[x].real() / pad);
    }
}

/* kernel_FDK_reconstruction, V63/FDK/kernel_FDK.cu:16-62, views in order. */
template <typename T>
void fdk_bp_impl(const Geo<T> &d, const T *filt, T *vol, int v0, int nv) {
    const int64_t P = (int64_t)d.R * d.ZP;
    std::vector<T> sn(nv), cs(nv);
    for (int v = 0; v < nv; ++v) view_angle(d, v0 + v, true, sn[v], cs[v]);
    const T dth = d.us_rate * pi_c<T>() / (T)180;
    const T usn = d.us_rate * (T)d.nviews;
#pragma omp parallel for collapse(2)
    for (int iz = 0; iz < d.Z; ++iz)
        for (int iy = 0; iy < d.N; ++iy)
            for (int ix = 0; ix < d.M; ++ix) {
                int64_t idx = ((int64_t)iz * d.N + iy) * d.M + ix;
                T x = d.bx + d.vx * (T)0.5 + (T)ix * d.vx;
                T y = d.by + d.vy * (T)0.5 + (T)iy * d.vy;
                T z = d.bz + d.vz * (T)0.5 + (T)iz * d.vz;
                T acc = vol[idx];
                for (int v = 0; v < nv; ++v) {
                    T c = cs[v], s = sn[v];
                    T t = x * c + y * s;
                    T sv = -x * s + y * c;
                    T t1 = d.dso * d.dso / ((d.dso - sv) * (d.dso - sv));
                    T ksi = d.dso * z / (d.dso - sv);
                    ksi = (ksi * (d.dso - d.dod) / d.dso - d.zmin) / d.dv;
                    T p = d.dso * t / (d.dso - sv);
                    p = (p * (d.dso - d.dod) / d.dso - d.ymin) / d.du;
                    T t2 = 0;
                    if (((T)0 <= ksi) && (ksi + 1 < (T)d.ZP) && ((T)0 <= p) && (p + 1 < (T)d.R)) {
                        int an = (int)std::floor(ksi);
                        T an0 = ksi - (T)an;
                        int bn = (int)std::floor(p);
                        T bn0 = p - (T)bn;
                        const T *pr = filt + (int64_t)v * P;
                        t2 = ((T)1 - an0) * ((T)1 - bn0) * pr[(int64_t)d.R * an + bn] +
                             ((T)1 - an0) * bn0 * pr[(int64_t)d.R * an + bn + 1] +
                             an0 * ((T)1 - bn0) * pr[(int64_t)(an + 1) * d.R + bn] +
                             an0 * bn0 * pr[(int64_t)(an + 1) * d.R + bn + 1];
                    }
                    // temp1*temp2 * (us_rate*PI/180.0f) * 360.0/(us_rate*Nviews), :60:
                    // the FP32 product, then the double literal 360.0 takes the
                    // rest of the expression and the += to double
                    acc = (T)((double)acc + (double)(t1 * t2 * dth) * 360.0 / (double)usn);
                }
                vol[idx] = acc;
            }
}

/* Multi-source FDK backprojection, one of the reference's three source
 * combinations per view: 0 weights the sources whose stencil is on the
 * detector by 1/(z - z_s)^2 and normalises (TB/FDK/kernel_FDK.cu:37-99);
 * 1 takes the source nearest in z among those whose row lies on the
 * detector (TB/FDK/OneRay/kernel_FDK.cu:37-113); 2 sums the sources
 * (TB/FDK/EqualWeighting/kernel_FDK.cu:37-91, without its unsynchronised
 * update).  Each view's combination is added to the volume
 * (kernel_volume_divide, TB/FDK/kernel_FDK.cu:21-33). */
template <typename T>
void fdk_bp_ms_impl(const Geo<T> &d, const T *filt, T *vol, int v0, int nv, int mode) {
    const int64_t P = (int64_t)d.R * d.ZP;
    std::vector<T> sn(nv), cs(nv);
    for (int v = 0; v < nv; ++v) view_angle(d, v0 + v, true, sn[v], cs[v]);
    const T dth = d.us_rate * pi_c<T>() / (T)180;
    const T usn = d.us_rate * (T)d.nviews;
#pragma omp parallel for collapse(2)
    for (int iz = 0; iz < d.Z; ++iz)
        for (int iy = 0; iy < d.N; ++iy)
            for (int ix = 0; ix < d.M; ++ix) {
                int64_t idx = ((int64_t)iz * d.N + iy) * d.M + ix;
                T x = d.bx + d.vx * (T)0.5 + (T)ix * d.vx;
                T y = d.by + d.vy * (T)0.5 + (T)iy * d.vy;
                T z = d.bz + d.vz * (T)0.5 + (T)iz * d.vz;
                T acc = vol[idx];
                for (int v = 0; v < nv; ++v) {
                    const T c = cs[v], s = sn[v];
                    const T t = x * c + y * s;
                    const T sv = -x * s + y * c;
                    const T t1 = d.dso * d.dso / ((d.dso - sv) * (d.dso - sv));
                    // the view's interpolated value from source i (false: off the detector)
                    auto sample = [&](int i, T &t2) -> bool {
                        const T sz = source_z(d, i);
                        T ksi = d.dso * (z - sz) / (d.dso - sv);
                        ksi = (ksi * (d.dso - d.dod) / d.dso + sz - d.zmin) / d.dv;
                        T p = d.dso * t / (d.dso - sv);
                        p = (p * (d.dso - d.dod) / d.dso - d.ymin) / d.du;
                        if (!(((T)0 <= ksi) && (ksi + 1 < (T)d.ZP) && ((T)0 <= p) && (p + 1 < (T)d.R))) return false;
                        const int an = (int)std::floor(ksi), bn = (int)std::floor(p);
                        const T an0 = ksi - (T)an, bn0 = p - (T)bn;
                        const T *pr = filt + ((int64_t)v * d.NS + i) * P;
                        t2 = ((T)1 - an0) * ((T)1 - bn0) * pr[(int64_t)d.R * an + bn] +
                             ((T)1 - an0) * bn0 * pr[(int64_t)d.R * an + bn + 1] +
                             an0 * ((T)1 - bn0) * pr[(int64_t)(an + 1) * d.R + bn] +
                             an0 * bn0 * pr[(int64_t)(an + 1) * d.R + bn + 1];
                        return true;
                    };
                    T upd = 0;
                    if (mode == 0) {
                        T wsum = 0;
                        for (int i = 0; i < d.NS; ++i) {
                            T t2;
                            if (!sample(i, t2)) continue;
                            const T dz = std::fabs(z - source_z(d, i));
                            const T w = (T)1 / dz / dz;
                            wsum += w;
                            upd += t1 * t2 * dth * (T)360 / usn * w;
                        }
                        if (wsum > (T)0) upd /= wsum;
                    } else if (mode == 1) {
                        int sel = -1;
                        for (int i = 0; i < d.NS; ++i) {
                            const T sz = source_z(d, i);
                            T ksi = d.dso * (z - sz) / (d.dso - sv);
                            ksi = (ksi * (d.dso - d.dod) / d.dso + sz - d.zmin) / d.dv;
                            if (((T)0 <= ksi) && (ksi + 1 < (T)d.ZP)) {
                                if (sel < 0 || std::fabs(sz - z) < std::fabs(source_z(d, sel) - z)) sel = i;
                            }
                        }
                        T t2;
                        if (sel >= 0 && sample(sel, t2)) upd = t1 * t2 * dth * (T)360 / usn;
                    } else {
                        for (int i = 0; i < d.NS; ++i) {
                            T t2;
                            if (sample(i, t2)) upd += t1 * t2 * dth * (T)360 / usn;
                        }
                    }
                    acc = acc + upd;
                }
                vol[idx] = acc;
            }
}

template <typename T> struct Buf { T *p; };

} // namespace

/* ------------------------------------------------------------------------ */
/* extern "C" surface used by oracle/oracle.py.  prec is 32 or 64; arrays are
 * host arrays of that precision. */
extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

#define ORC_DISPATCH(prec, CALL32, CALL64) \
    do { if ((prec) == 64) { CALL64; } else { CALL32; } } while (0)

void orc_view_angle(const ct_geometry *g, int prec, int view, int fdk, double *s, double *c) {
    if (prec == 64) { auto d = derive<double>(g); double a, b; view_angle(d, view, fdk != 0, a, b); *s = a; *c = b; }
    else { auto d = derive<float>(g); float a, b; view_angle(d, view, fdk != 0, a, b); *s = a; *c = b; }
}

void orc_forward(const ct_geometry *g, int prec, const void *vol, void *proj, int v0, int nv) {
    ORC_DISPATCH(prec,
        forward_impl(derive<float>(g), (const float *)vol, (float *)proj, (const float *)nullptr, v0, nv, 0),
        forward_impl(derive<double>(g), (const double *)vol, (double *)proj, (const double *)nullptr, v0, nv, 0));
}
void orc_sart_correction(const ct_geometry *g, int prec, const void *vol, const void *meas, void *corr, int v0, int nv) {
    ORC_DISPATCH(prec,
        forward_impl(derive<float>(g), (const float *)vol, (float *)corr, (const float *)meas, v0, nv, 1),
        forward_impl(derive<double>(g), (const double *)vol, (double *)corr, (const double *)meas, v0, nv, 1));
}
void orc_backproject_voxel(const ct_geometry *g, int prec, void *vol, const void *corr, int v0, int nv, double beta, int mode) {
    ORC_DISPATCH(prec,
        backproject_impl(derive<float>(g), (float *)vol, (const float *)corr, v0, nv, (float)beta, mode),
        backproject_impl(derive<double>(g), (double *)vol, (const double *)corr, v0, nv, beta, mode));
}
void orc_sart_sweep(const ct_geometry *g, int prec, void *vol, const void *meas, double beta, int subset, int order) {
    ORC_DISPATCH(prec,
        sart_sweep_impl(derive<float>(g), (float *)vol, (const float *)meas, (float)beta, subset, order),
        sart_sweep_impl(derive<double>(g), (double *)vol, (const double *)meas, beta, subset, order));
}
void orc_adjoint(const ct_geometry *g, int prec, const void *proj, void *vol, int v0, int nv) {
    ORC_DISPATCH(prec,
        adjoint_impl(derive<float>(g), (const float *)proj, (float *)vol, v0, nv),
        adjoint_impl(derive<double>(g), (const double *)proj, (double *)vol, v0, nv));
}
/* Per-ray Siddon segments of one ray (for the brute-force pin): writes up to
 * cap (index, length) pairs, returns the count. */
int orc_ray_segments(const ct_geometry *g, int prec, int view, int col, int row, int src,
                     int64_t *idx, double *len, int cap) {
    int n = 0;
    if (prec == 64) {
        auto d = derive<double>(g); double s, c; view_angle(d, view, false, s, c);
        auto vis = [&](int64_t i, double l) { if (n < cap) { idx[n] = i; len[n] = l; } n++; };
        siddon_ray(d, s, c, col, row, src, vis);
    } else {
        auto d = derive<float>(g); float s, c; view_angle(d, view, false, s, c);
        auto vis = [&](int64_t i, float l) { if (n < cap) { idx[n] = i; len[n] = l; } n++; };
        siddon_ray(d, s, c, col, row, src, vis);
    }
    return n;
}
void orc_fista_momentum(int prec, void *x, void *xprev, void *y, int64_t n, double gamma) {
    if (prec == 64) {
        double *X = (double *)x, *P = (double *)xprev, *Y = (double *)y;
        for (int64_t i = 0; i < n; ++i) { double xi = X[i]; Y[i] = xi + gamma * (xi - P[i]); P[i] = xi; }
    } else {
        float *X = (float *)x, *P = (float *)xprev, *Y = (float *)y; float g = (float)gamma;
        for (int64_t i = 0; i < n; ++i) { float xi = X[i]; Y[i] = xi + g * (xi - P[i]); P[i] = xi; }
    }
}
void orc_tv_fgp(int prec, int M, int N, int Z, const void *b, void *out, double lambda, int niter) {
    ORC_DISPATCH(prec,
        fgp_impl(M, N, Z, (const float *)b, (float *)out, (float)lambda, niter),
        fgp_impl(M, N, Z, (const double *)b, (double *)out, lambda, niter));
}
void orc_tv_fgp_duals(int prec, int M, int N, int Z, const void *b, void *out, void *duals, double lambda, int niter) {
    ORC_DISPATCH(prec,
        fgp_impl(M, N, Z, (const float *)b, (float *)out, (float)lambda, niter, (float *)duals),
        fgp_impl(M, N, Z, (const double *)b, (double *)out, lambda, niter, (double *)duals));
}
void orc_tv_D(int prec, int M, int N, int Z, const void *x, void *r, void *s, void *t) {
    ORC_DISPATCH(prec,
        fgp_D(M, N, Z, (const float *)x, (float *)r, (float *)s, (float *)t),
        fgp_D(M, N, Z, (const double *)x, (double *)r, (double *)s, (double *)t));
}
void orc_tv_Dt(int prec, int M, int N, int Z, const void *r, const void *s, const void *t, void *out) {
    ORC_DISPATCH(prec,
        fgp_Dt(M, N, Z, (const float *)r, (const float *)s, (const float *)t, (float *)out),
        fgp_Dt(M, N, Z, (const double *)r, (const double *)s, (const double *)t, (double *)out));
}
double orc_tv_norm(int prec, int M, int N, int Z, const void *f) {
    return prec == 64 ? tv_norm_impl(M, N, Z, (const double *)f) : tv_norm_impl(M, N, Z, (const float *)f);
}
double orc_residual_norm2(int prec, const void *a, const void *b, int64_t n) {
    double acc = 0;
    if (prec == 64) {
        const double *A = (const double *)a, *B = (const double *)b;
#pragma omp parallel for reduction(+ : acc)
        for (int64_t i = 0; i < n; ++i) { double e = A[i] - B[i]; acc += e * e; }
    } else {
        const float *A = (const float *)a, *B = (const float *)b;
#pragma omp parallel for reduction(+ : acc)
        for (int64_t i = 0; i < n; ++i) { double e = (double)A[i] - (double)B[i]; acc += e * e; }
    }
    return acc;
}
double orc_tv_eps(int prec, int M, int N, int Z, const void *f, double eps) {
    return prec == 64 ? tv_eps_impl(M, N, Z, (const double *)f, eps) : tv_eps_impl(M, N, Z, (const float *)f, eps);
}
void orc_tv_eps_grad(int prec, int M, int N, int Z, const void *f, double eps, void *g) {
    ORC_DISPATCH(prec,
        tv_eps_grad_impl(M, N, Z, (const float *)f, eps, (float *)g),
        tv_eps_grad_impl(M, N, Z, (const double *)f, eps, (double *)g));
}
void orc_fdk_weight(const ct_geometry *g, int prec, const void *in, void *out) {
    ORC_DISPATCH(prec,
        fdk_weight_impl(derive<float>(g), (const float *)in, (float *)out),
        fdk_weight_impl(derive<double>(g), (const double *)in, (double *)out));
}
void orc_ramp_filter(const ct_geometry *g, int prec, const void *in, void *out, int nimg, int window) {
    ORC_DISPATCH(prec,
        ramp_impl(g, (const float *)in, (float *)out, nimg, window),
        ramp_impl(g, (const double *)in, (double *)out, nimg, window));
}
void orc_ramp_taps(const ct_geometry *g, int window, double *taps) {
    std::vector<double> H = ramp_filter_taps(g, window);
    std::memcpy(taps, H.data(), sizeof(double) * H.size());
}
void orc_fft(double *re, double *im, int n, int sign) {
    std::vector<std::complex<double>> a(n);
    for (int i = 0; i < n; ++i) a[i] = std::complex<double>(re[i], im[i]);
    fft_inplace(a, sign);
    for (int i = 0; i < n; ++i) { re[i] = a[i].real(); im[i] = a[i].imag(); }
}
void orc_fdk_backproject(const ct_geometry *g, int prec, const void *filt, void *vol, int v0, int nv) {
    ORC_DISPATCH(prec,
        fdk_bp_impl(derive<float>(g), (const float *)filt, (float *)vol, v0, nv),
        fdk_bp_impl(derive<double>(g), (const double *)filt, (double *)vol, v0, nv));
}
void orc_fdk_backproject_ms(const ct_geometry *g, int prec, const void *filt, void *vol, int v0, int nv, int mode) {
    ORC_DISPATCH(prec,
        fdk_bp_ms_impl(derive<float>(g), (const float *)filt, (float *)vol, v0, nv, mode),
        fdk_bp_ms_impl(derive<double>(g), (const double *)filt, (double *)vol, v0, nv, mode));
}
/* Reconstruction_FDK, V63/host_FDK.cpp:4-57 (multi-source: TB/FDK/host_FDK.cpp:4-83
 * with the source combination `mode`): vol += FDK(proj). */
void orc_fdk_ms(const ct_geometry *g, int prec, const void *proj, void *vol, int window, int mode) {
    const int ns = g->nsources < 1 ? 1 : g->nsources;
    const int64_t n = (int64_t)g->nviews * ns * g->ncols * g->nrows;
    if (prec == 64) {
        std::vector<double> w(n), f(n);
        fdk_weight_impl(derive<double>(g), (const double *)proj, w.data());
        ramp_impl(g, w.data(), f.data(), g->nviews * ns, window);
        if (ns > 1) fdk_bp_ms_impl(derive<double>(g), f.data(), (double *)vol, 0, g->nviews, mode);
        else fdk_bp_impl(derive<double>(g), f.data(), (double *)vol, 0, g->nviews);
    } else {
        std::vector<float> w(n), f(n);
        fdk_weight_impl(derive<float>(g), (const float *)proj, w.data());
        ramp_impl(g, w.data(), f.data(), g->nviews * ns, window);
        if (ns > 1) fdk_bp_ms_impl(derive<float>(g), f.data(), (float *)vol, 0, g->nviews, mode);
        else fdk_bp_impl(derive<float>(g), f.data(), (float *)vol, 0, g->nviews);
    }
}
void orc_fdk(const ct_geometry *g, int prec, const void *proj, void *vol, int window) {
    orc_fdk_ms(g, prec, proj, vol, window, 0);
}

} // extern "C"
