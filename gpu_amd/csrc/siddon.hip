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
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const bool ok = alive && (r.amax - r.ac > loop_eps<T>()) && (guard > 0) &&
                                (unsigned)r.vi < M && (unsigned)r.vj < N && (unsigned)r.vk < Z;
                const T m = dmin(dmin(r.ax, r.ay), r.az);
                act[u] = ok;
                seg[u] = r.dist * (m - r.ac);
                idx[u] = ok ? (unsigned)r.vk * MN + (unsigned)r.vj * M + (unsigned)r.vi : 0u;
                if (ok) {
                    --guard;
                    r.ac = m;
                    if (r.ax == m) { r.vi += r.sgx; r.nx += r.sgx; r.ax = (d.vx * (T)r.nx + d.bx - r.x1x) * r.ixd; }
                    if (r.ay == m) { r.vj += r.sgy; r.ny += r.sgy; r.ay = (d.vy * (T)r.ny + d.by - r.x1y) * r.iyd; }
                    if (r.az == m) { r.vk += r.sgz; r.nz += r.sgz; r.az = (d.vz * (T)r.nz + d.bz - r.x1z) * r.izd; }
                }
                alive = ok;
            }
            T fv[B];
#pragma unroll
            for (int u = 0; u < B; ++u) fv[u] = (T)ldf(f, idx[u]);
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (act[u]) {
                    if constexpr (COMP) {
                        const T yl = seg[u] - clen;
                        const T tl = len + yl;
                        clen = (tl - len) - yl;
                        len = tl;
                        const T ys = seg[u] * fv[u] - csum;
                        const T ts = sum + ys;
                        csum = (ts - sum) - ys;
                        sum = ts;
                    } else {
                        len += seg[u];
                        sum += seg[u] * fv[u];
                    }
                }
            }
        }
    }
    len_out = len;
    return sum;
}

__device__ __forceinline__ void tile_coords(int &col, int &row) {
    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63;
    col = blockIdx.x * 16 + (w & 1) * 8 + (l & 7);
    row = blockIdx.y * 16 + (w >> 1) * 8 + (l >> 3);
}

// CMD 0: ray sum; CMD 1: SART correction (b - sum)/len
// (V63/kernel_IterativeRecon_universal.cu:554-563).  blockIdx.z = q*NS + src
// for batch slot q; view = first + q*stride; meas is indexed by view - mv0.
template <typename T, typename SV, typename SO, typename SM, int CMD, bool COMP>
__global__ __launch_bounds__(256) void k_forward(DGeo<T> d, const T *__restrict__ ang,
                                                 const SV *__restrict__ f, SO *__restrict__ out,
                                                 const SM *__restrict__ meas, int first, int stride,
                                                 int mv0, int z0) {
    int col, row;
    tile_coords(col, row);
    if (col >= d.R || row >= d.ZP) return;
    const int img = blockIdx.z + z0;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = first + q * stride;
    const T s = ang[2 * v], c = ang[2 * v + 1];
    T len;
    const T sum = siddon_sum<T, SV, 4, COMP>(d, s, c, col, row, src, f, len);
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

template <typename A> __device__ __forceinline__ void atomic_add(A *p, A v);
template <> __device__ __forceinline__ void atomic_add<float>(float *p, float v) { unsafeAtomicAdd(p, v); }
template <> __device__ __forceinline__ void atomic_add<double>(double *p, double v) { unsafeAtomicAdd(p, v); }

// Matched adjoint: scatter the forward's segments (scatter form of
// V63/pixel_driven_backprj/kernel_IterativeRecon_CBCT.cu:563-1116).  A first
// pass measures the ray length so rays the forward zeroes (len < 1e-6*vz)
// scatter nothing: the pair is then an exact transpose.
template <typename T, typename SP, typename A>
__global__ __launch_bounds__(256) void k_adjoint(DGeo<T> d, const T *__restrict__ ang,
                                                 const SP *__restrict__ proj, A *__restrict__ acc,
                                                 int view0, int z0) {
    int col, row;
    tile_coords(col, row);
    if (col >= d.R || row >= d.ZP) return;
    const int img = blockIdx.z + z0;
    const int q = img / d.NS, src = img - q * d.NS;
    const int v = view0 + q;
    const int64_t P = (int64_t)d.R * d.ZP;
    const T y = (T)ldf(proj, (int64_t)img * P + (int64_t)row * d.R + col);
    if (y == (T)0) return;
    const T s = ang[2 * v], c = ang[2 * v + 1];
    const T len = siddon_traverse(d, s, c, col, row, src, [](int, T) {});
    if (len < d.vz * (T)1e-6) return;
    siddon_traverse(d, s, c, col, row, src, [&](int idx, T seg) { atomic_add<A>(acc + idx, (A)seg * (A)y); });
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
        hipLaunchKernelGGL((k_adjoint<T, SP, A>), grid, block, 0, ctx->stream, d, ang, (const SP *)proj, acc, view0, z0);
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
    // rounded once to the output type
    ct_status st;
    double *acc = (double *)slot(ctx, SLOT_TMP_VOL, sizeof(double) * V, &st);
    if (!acc) return st;
    st = launch_fill(ctx, acc, CT_F64, V, 0.0);
    if (st != CT_OK) return st;
    switch (pdt) {
    case CT_F32: st = run_adjoint<float, float, double>(ctx, g, proj, acc, view0, nv); break;
    case CT_F16: st = run_adjoint<float, __half, double>(ctx, g, proj, acc, view0, nv); break;
    case CT_BF16: st = run_adjoint<float, __hip_bfloat16, double>(ctx, g, proj, acc, view0, nv); break;
    default: return fail(ctx, CT_ERR_DTYPE, "unsupported projection dtype");
    }
    if (st != CT_OK) return st;
    return launch_convert(ctx, acc, CT_F64, vol, vdt, V);
}

} // namespace ct
