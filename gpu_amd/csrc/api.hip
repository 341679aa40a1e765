// api.hip - the C ABI of ct_core.h: context, validation, scratch, entry points
// and the device-resident SART / SART-FISTA-TV drivers.
#include "ct_internal.hpp"
#include <cmath>
#include <cstring>
#include <cstdio>
#include <limits>

namespace ct {

ct_status fail(ct_context *ctx, ct_status s, const std::string &msg) {
    if (ctx) ctx->last_error = msg;
    return s;
}

ct_status hip_check(ct_context *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return CT_OK;
    return fail(ctx, CT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

ct_status launch_check(ct_context *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return CT_OK;
    return fail(ctx, CT_ERR_LAUNCH, std::string(what) + " launch: " + hipGetErrorString(e));
}

void *slot(ct_context *ctx, int which, size_t bytes, ct_status *st) {
    DevBuf &b = ctx->slots[which];
    if (bytes == 0) bytes = 1;
    if (b.bytes >= bytes) { *st = CT_OK; return b.ptr; }
    if (b.ptr) {
        // the old buffer may still be read by queued kernels
        hipStreamSynchronize(ctx->stream);
        hipFree(b.ptr);
        ctx->scratch_total -= b.bytes;
        b.ptr = nullptr; b.bytes = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        hipGetLastError();
        *st = fail(ctx, CT_ERR_ALLOC, "device scratch allocation of " + std::to_string(bytes) + " bytes failed");
        return nullptr;
    }
    b.ptr = p; b.bytes = bytes;
    ctx->scratch_total += bytes;
    *st = CT_OK;
    return p;
}

// Per-view angles exactly as the reference host code forms them
// (V63/host_IterativeRecon_CBCT.c:34-41, FDK: V63/host_FDK.cpp:35-40).
template <typename T> static void angle_host(const ct_geometry &g, int j, bool fdk, T &s, T &c) {
    const T pi = std::is_same<T, float>::value ? (T)3.141592653589793f : (T)3.141592653589793;
    T th = (T)((T)j * (T)g.angle_step_deg + (T)g.angle0_deg);
    th = (th + (T)g.angle_shift_deg) * pi / (T)180;
    if (!fdk) th = -th;
    s = std::sin(th); c = std::cos(th);
}
// cos of view j's projector angle (FP32, as the kernels' angle table holds it)
float view_cos(const ct_geometry &g, int j) {
    float s, c;
    angle_host<float>(g, j, false, s, c);
    return c;
}

void fill_angle_table(const ct_geometry &g, bool dbl, bool fdk, void *host) {
    for (int v = 0; v < g.nviews; ++v) {
        if (dbl) { double s, c; angle_host<double>(g, v, fdk, s, c); ((double *)host)[2 * v] = s; ((double *)host)[2 * v + 1] = c; }
        else { float s, c; angle_host<float>(g, v, fdk, s, c); ((float *)host)[2 * v] = s; ((float *)host)[2 * v + 1] = c; }
    }
}

const void *angle_table(ct_context *ctx, const ct_geometry &g, bool dbl, bool fdk, ct_status *st) {
    std::vector<double> key = {(double)g.nviews, g.angle_step_deg, g.angle0_deg, g.angle_shift_deg,
                               dbl ? 1.0 : 0.0, fdk ? 1.0 : 0.0};
    const size_t bytes = (size_t)g.nviews * 2 * (dbl ? 8 : 4);
    if (ctx->angles.ptr && ctx->angle_key == key) { *st = CT_OK; return ctx->angles.ptr; }
    if (ctx->angles.bytes < bytes) {
        if (ctx->angles.ptr) { hipStreamSynchronize(ctx->stream); hipFree(ctx->angles.ptr); ctx->scratch_total -= ctx->angles.bytes; }
        ctx->angles.ptr = nullptr; ctx->angles.bytes = 0;
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            hipGetLastError();
            *st = fail(ctx, CT_ERR_ALLOC, "angle table allocation failed");
            return nullptr;
        }
        ctx->angles.ptr = p; ctx->angles.bytes = bytes; ctx->scratch_total += bytes;
    } else {
        // the table is about to be rewritten: queued kernels may still read it
        hipStreamSynchronize(ctx->stream);
    }
    std::vector<char> host(bytes);
    fill_angle_table(g, dbl, fdk, host.data());
    hipError_t e = hipMemcpy(ctx->angles.ptr, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) { *st = hip_check(ctx, e, "angle table upload"); return nullptr; }
    ctx->angle_key = key;
    *st = CT_OK;
    return ctx->angles.ptr;
}

void prof_drain(ct_context *ctx) {
    if (ctx->prof_recs.empty()) return;
    hipEventSynchronize(ctx->prof_recs.back().b);
    for (auto &r : ctx->prof_recs) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ctx->prof_ms[r.kid] += ms;
            ctx->prof_n[r.kid] += 1;
        } else {
            hipGetLastError();
        }
        ctx->prof_pool.push_back(r.a);
        ctx->prof_pool.push_back(r.b);
    }
    ctx->prof_recs.clear();
}

ct_status check_ptr(ct_context *ctx, const void *p, ct_dtype dt, const char *name) {
    if (!p) return fail(ctx, CT_ERR_NULL, std::string(name) + " is NULL");
    if (((uintptr_t)p) % dtype_size(dt)) return fail(ctx, CT_ERR_ALIGN, std::string(name) + " is not aligned to its element size");
    return CT_OK;
}

static bool finite_pos(double v) { return std::isfinite(v) && v > 0; }

static ct_status validate(ct_context *ctx, const ct_geometry *g) {
    if (!g) return fail(ctx, CT_ERR_NULL, "geometry is NULL");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(ctx, CT_ERR_GEOMETRY, "volume dimensions must be >= 1");
    if (g->ncols < 1 || g->nrows < 1) return fail(ctx, CT_ERR_GEOMETRY, "detector dimensions must be >= 1");
    if (g->nviews < 1) return fail(ctx, CT_ERR_GEOMETRY, "nviews must be >= 1");
    if (g->nsources < 1) return fail(ctx, CT_ERR_GEOMETRY, "nsources must be >= 1");
    if (!finite_pos(g->vx) || !finite_pos(g->vy) || !finite_pos(g->vz)) return fail(ctx, CT_ERR_GEOMETRY, "voxel pitch must be finite and > 0");
    if (!finite_pos(g->du) || !finite_pos(g->dv)) return fail(ctx, CT_ERR_GEOMETRY, "detector pitch must be finite and > 0");
    if (!std::isfinite(g->dso) || !std::isfinite(g->dod) || !std::isfinite(g->det_offset) || !std::isfinite(g->src_z0) ||
        !std::isfinite(g->src_dz) || !std::isfinite(g->angle_step_deg) || !std::isfinite(g->angle0_deg) ||
        !std::isfinite(g->angle_shift_deg))
        return fail(ctx, CT_ERR_GEOMETRY, "geometry holds a non-finite value");
    if (!(g->dso > g->dod)) return fail(ctx, CT_ERR_GEOMETRY, "source must lie beyond the detector plane (dso > dod)");
    if (g->style < CT_CONE || g->style > CT_PARALLEL) return fail(ctx, CT_ERR_GEOMETRY, "unknown beam style");
    if (g->style != CT_CONE && g->nsources != 1) return fail(ctx, CT_ERR_GEOMETRY, "multi-source geometry is cone beam");
    // the source orbit must stay outside the volume's footprint
    const double bx = -g->vx * (g->nx / 2.0 + g->x_offset), ex = bx + g->vx * g->nx;
    const double hy = g->vy * g->ny / 2.0;
    const double rx = std::max(std::fabs(bx), std::fabs(ex));
    const double r = std::sqrt(rx * rx + hy * hy);
    if (!(g->dso > r)) return fail(ctx, CT_ERR_GEOMETRY, "source orbit (dso) lies inside the volume");
    const int64_t V = (int64_t)g->nx * g->ny * g->nz;
    if (V >= (int64_t)1 << 31) return fail(ctx, CT_ERR_OVERFLOW, "volume exceeds 2^31 voxels");
    if ((int64_t)g->nx * g->ny >= (int64_t)1 << 31) return fail(ctx, CT_ERR_OVERFLOW, "volume plane exceeds 2^31 voxels");
    if ((int64_t)g->ncols * g->nrows >= (int64_t)1 << 31) return fail(ctx, CT_ERR_OVERFLOW, "detector exceeds 2^31 pixels");
    if ((int64_t)g->nviews * g->nsources >= (int64_t)1 << 31) return fail(ctx, CT_ERR_OVERFLOW, "too many images");
    return CT_OK;
}

static ct_status views_ok(ct_context *ctx, const ct_geometry *g, int32_t view0, int32_t nviews) {
    if (view0 < 0 || nviews < 0 || (int64_t)view0 + nviews > g->nviews)
        return fail(ctx, CT_ERR_INVALID_ARG, "view range outside [0, nviews)");
    return CT_OK;
}

static ct_status dt_ok(ct_context *ctx, int dt) {
    if (!dtype_valid(dt)) return fail(ctx, CT_ERR_DTYPE, "unknown dtype");
    return CT_OK;
}

static ct_status dims_ok(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz) {
    if (nx < 1 || ny < 1 || nz < 1) return fail(ctx, CT_ERR_INVALID_ARG, "volume dimensions must be >= 1");
    if ((int64_t)nx * ny * nz >= (int64_t)1 << 31) return fail(ctx, CT_ERR_OVERFLOW, "volume exceeds 2^31 voxels");
    return CT_OK;
}

ct_status validate_geometry(ct_context *ctx, const ct_geometry *g) { return validate(ctx, g); }

static int64_t vox(const ct_geometry *g) { return (int64_t)g->nx * g->ny * g->nz; }
static int64_t pix(const ct_geometry *g) { return (int64_t)g->ncols * g->nrows * (g->nsources < 1 ? 1 : g->nsources); }

// SART sweep: subsets in ct_order, corrections of a subset from the current
// volume, one voxel-driven update per subset.
static ct_status sweep(ct_context *ctx, const ct_geometry &g, void *vol, ct_dtype vdt, const void *meas, ct_dtype mdt,
                       double beta, int subset, ct_order order) {
    const bool dbl = vdt == CT_F64;
    const size_t es = dbl ? 8 : 4;
    ct_status st;
    void *corr = slot(ctx, SLOT_CORR, es * (size_t)pix(&g) * subset, &st);
    if (!corr) return st;
    const ct_dtype cdt = dbl ? CT_F64 : CT_F32;
    const int nv = g.nviews;
    if (sweep_zq_eligible(ctx, g, vdt, subset)) {
        // fast mode: the sweep runs on a z-quad-layout copy of the volume
        // (coalesced gathers for the ray walk), converted once each way
        const size_t bytes = (size_t)g.nx * g.ny * (size_t)((g.nz + 3) & ~3) * dtype_size(vdt);
        void *volq = slot(ctx, SLOT_ZQVOL, bytes, &st);
        if (!volq) return st;
        if ((st = launch_zq_pack(ctx, vol, vdt, volq, g.nx, g.ny, g.nz)) != CT_OK) return st;
        // views whose rays run along y take the transposed copy, so that all
        // of them walk by column; the copy follows the sweep's direction
        // changes (four per orbit in view order)
        void *volt = nullptr;
        bool tr = false;
        auto step = [&](int first, int stride, int cnt) {
            const bool want = subset == 1 || cnt == 1 ? sweep_view_transposed(g, first)
                                                      : sweep_view_transposed(g, first + stride * (cnt / 2));
            if (want != tr) {
                if (!volt && !(volt = slot(ctx, SLOT_ZQVOL_T, bytes, &st))) return st;
                ct_status s0 = want ? launch_zq_transpose(ctx, volq, vdt, volt, g.nx, g.ny, g.nz, true)
                                    : launch_zq_transpose(ctx, volt, vdt, volq, g.nx, g.ny, g.nz, false);
                if (s0 != CT_OK) return s0;
                tr = want;
            }
            void *v = tr ? volt : volq;
            ct_status s2 = launch_correction_zq(ctx, g, v, vdt, (float *)corr, meas, mdt, first, stride, cnt, tr);
            if (s2 != CT_OK) return s2;
            return launch_voxel_bp_zq(ctx, g, v, vdt, corr, first, stride, cnt, beta, tr);
        };
        if (order == CT_ORDER_INTERLEAVED && subset > 1) {
            const int nsub = (nv + subset - 1) / subset;
            for (int s = 0; s < nsub; ++s)
                if ((st = step(s, nsub, (nv - s + nsub - 1) / nsub)) != CT_OK) return st;
        } else {
            for (int p0 = 0; p0 < nv; p0 += subset)
                if ((st = step(p0, 1, std::min(subset, nv - p0))) != CT_OK) return st;
        }
        if (tr && (st = launch_zq_transpose(ctx, volt, vdt, volq, g.nx, g.ny, g.nz, false)) != CT_OK) return st;
        return launch_zq_unpack(ctx, volq, vdt, vol, g.nx, g.ny, g.nz);
    }
    if (order == CT_ORDER_INTERLEAVED && subset > 1) {
        const int nsub = (nv + subset - 1) / subset;
        for (int s = 0; s < nsub; ++s) {
            const int cnt = (nv - s + nsub - 1) / nsub;
            st = launch_forward_mv(ctx, g, vol, vdt, corr, cdt, meas, mdt, s, nsub, cnt, 1, 0);
            if (st != CT_OK) return st;
            st = launch_voxel_bp(ctx, g, vol, vdt, corr, s, nsub, cnt, beta, 1);
            if (st != CT_OK) return st;
        }
    } else {
        for (int p0 = 0; p0 < nv; p0 += subset) {
            const int cnt = std::min(subset, nv - p0);
            st = launch_forward_mv(ctx, g, vol, vdt, corr, cdt, meas, mdt, p0, 1, cnt, 1, 0);
            if (st != CT_OK) return st;
            st = launch_voxel_bp(ctx, g, vol, vdt, corr, p0, 1, cnt, beta, 1);
            if (st != CT_OK) return st;
        }
    }
    return CT_OK;
}

ct_status launch_fgp(ct_context *ctx, int M, int N, int Z, void *vol, ct_dtype dt, double lambda, int niter);

} // namespace ct

using namespace ct;

extern "C" {

int ct_abi_version(void) { return CT_ABI_VERSION; }

const char *ct_status_string(ct_status s) {
    switch (s) {
    case CT_OK: return "ok";
    case CT_ERR_INVALID_ARG: return "invalid argument";
    case CT_ERR_GEOMETRY: return "invalid geometry";
    case CT_ERR_DTYPE: return "unsupported dtype";
    case CT_ERR_NULL: return "null pointer";
    case CT_ERR_ALIGN: return "misaligned pointer";
    case CT_ERR_OVERFLOW: return "size overflow";
    case CT_ERR_ALLOC: return "allocation failed";
    case CT_ERR_DEVICE: return "device error";
    case CT_ERR_LAUNCH: return "launch failed";
    case CT_ERR_IO: return "file error";
    }
    return "unknown status";
}

ct_status ct_create(ct_context **out, int device) {
    if (!out) return CT_ERR_NULL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) { hipGetLastError(); return CT_ERR_DEVICE; }
    if (device < 0 || device >= n) return CT_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) { hipGetLastError(); return CT_ERR_DEVICE; }
    ct_context *ctx = new ct_context();
    ctx->device = device;
    ctx->stream = nullptr;
    // the allocation-free entry points' scratch address on this device
    const ct_status st = noalloc_init(ctx);
    if (st != CT_OK) { delete ctx; return st; }
    *out = ctx;
    return CT_OK;
}

ct_status ct_destroy(ct_context *ctx) {
    if (!ctx) return CT_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    prof_drain(ctx);
    for (hipEvent_t e : ctx->prof_pool) hipEventDestroy(e);
    for (auto &b : ctx->slots) if (b.ptr) hipFree(b.ptr);
    if (ctx->angles.ptr) hipFree(ctx->angles.ptr);
    delete ctx;
    return CT_OK;
}

ct_status ct_set_math(ct_context *ctx, ct_math mode) {
    if (!ctx) return CT_ERR_NULL;
    if (mode != CT_MATH_EXACT && mode != CT_MATH_FAST) return fail(ctx, CT_ERR_INVALID_ARG, "unknown math mode");
    ctx->math = mode;
    return CT_OK;
}

ct_status ct_profile_enable(ct_context *ctx, int32_t enable) {
    if (!ctx) return CT_ERR_NULL;
    ctx->prof = enable != 0;
    return CT_OK;
}

ct_status ct_profile_reset(ct_context *ctx) {
    if (!ctx) return CT_ERR_NULL;
    DeviceGuard dg(ctx);
    prof_drain(ctx);
    for (int k = 0; k < CT_K_COUNT; ++k) { ctx->prof_ms[k] = 0; ctx->prof_n[k] = 0; }
    return CT_OK;
}

ct_status ct_profile_read(ct_context *ctx, ct_kernel_id kernel, double *total_ms, int64_t *launches) {
    if (!ctx || !total_ms || !launches) return CT_ERR_NULL;
    if ((int)kernel < 0 || (int)kernel >= CT_K_COUNT) return fail(ctx, CT_ERR_INVALID_ARG, "unknown kernel id");
    DeviceGuard dg(ctx);
    prof_drain(ctx);
    *total_ms = ctx->prof_ms[kernel];
    *launches = ctx->prof_n[kernel];
    return CT_OK;
}

ct_status ct_set_stream(ct_context *ctx, void *stream) {
    if (!ctx) return CT_ERR_NULL;
    ctx->stream = (hipStream_t)stream;
    return CT_OK;
}

ct_status ct_synchronize(ct_context *ctx) {
    if (!ctx) return CT_ERR_NULL;
    DeviceGuard dg(ctx);
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "synchronize");
}

const char *ct_last_error(const ct_context *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

size_t ct_scratch_bytes(const ct_context *ctx) { return ctx ? ctx->scratch_total : 0; }

ct_status ct_validate_geometry(const ct_geometry *g) { return validate(nullptr, g); }

ct_status ct_view_angle(const ct_geometry *g, int32_t view, int32_t fdk, float *s, float *c) {
    ct_status st = validate(nullptr, g);
    if (st != CT_OK) return st;
    if (!s || !c) return CT_ERR_NULL;
    if (view < 0 || view >= g->nviews) return CT_ERR_INVALID_ARG;
    angle_host<float>(*g, view, fdk != 0, *s, *c);
    return CT_OK;
}

#define CT_TRY(expr) do { ct_status _s = (expr); if (_s != CT_OK) return _s; } while (0)
#define CT_CTX(ctx) do { if (!(ctx)) return CT_ERR_NULL; } while (0)

ct_status ct_forward(ct_context *ctx, const ct_geometry *g, const void *vol, ct_dtype vdt, void *proj, ct_dtype pdt,
                     int32_t view0, int32_t nviews) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, pdt));
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, proj, pdt, "proj"));
    CT_TRY(views_ok(ctx, g, view0, nviews));
    DeviceGuard dg(ctx);
    return launch_forward(ctx, *g, vol, vdt, proj, pdt, nullptr, pdt, view0, 1, nviews, 0);
}

ct_status ct_forward_noalloc(ct_context *ctx, const ct_geometry *g, const void *vol, ct_dtype vdt, void *proj,
                             ct_dtype pdt, int32_t view0, int32_t nviews) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, pdt));
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, proj, pdt, "proj"));
    CT_TRY(views_ok(ctx, g, view0, nviews));
    DeviceGuard dg(ctx);
    return launch_forward_noalloc(ctx, *g, vol, vdt, proj, pdt, view0, nviews);
}

ct_status ct_sart_correction(ct_context *ctx, const ct_geometry *g, const void *vol, ct_dtype vdt, const void *meas,
                             ct_dtype mdt, void *corr, int32_t view0, int32_t nviews) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, mdt));
    const ct_dtype cdt = vdt == CT_F64 ? CT_F64 : CT_F32;
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, meas, mdt, "meas")); CT_TRY(check_ptr(ctx, corr, cdt, "corr"));
    CT_TRY(views_ok(ctx, g, view0, nviews));
    DeviceGuard dg(ctx);
    return launch_forward(ctx, *g, vol, vdt, corr, cdt, meas, mdt, view0, 1, nviews, 1);
}

ct_status ct_backproject_voxel(ct_context *ctx, const ct_geometry *g, void *vol, ct_dtype vdt, const void *corr,
                               int32_t view0, int32_t nviews, double beta, int32_t mode) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt));
    const ct_dtype cdt = vdt == CT_F64 ? CT_F64 : CT_F32;
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, corr, cdt, "corr"));
    CT_TRY(views_ok(ctx, g, view0, nviews));
    if (mode != 0 && mode != 1) return fail(ctx, CT_ERR_INVALID_ARG, "mode must be 0 or 1");
    if (!std::isfinite(beta)) return fail(ctx, CT_ERR_INVALID_ARG, "beta must be finite");
    DeviceGuard dg(ctx);
    return launch_voxel_bp(ctx, *g, vol, vdt, corr, view0, 1, nviews, beta, mode);
}

ct_status ct_sart_sweep(ct_context *ctx, const ct_geometry *g, void *vol, ct_dtype vdt, const void *meas, ct_dtype mdt,
                        double beta, int32_t subset_size, ct_order order) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, mdt));
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, meas, mdt, "meas"));
    if (subset_size < 1 || subset_size > g->nviews) return fail(ctx, CT_ERR_INVALID_ARG, "subset_size must be in [1, nviews]");
    if (order != CT_ORDER_ASCENDING && order != CT_ORDER_INTERLEAVED) return fail(ctx, CT_ERR_INVALID_ARG, "unknown order");
    if (!std::isfinite(beta)) return fail(ctx, CT_ERR_INVALID_ARG, "beta must be finite");
    DeviceGuard dg(ctx);
    return sweep(ctx, *g, vol, vdt, meas, mdt, beta, subset_size, order);
}

ct_status ct_adjoint(ct_context *ctx, const ct_geometry *g, const void *proj, ct_dtype pdt, void *vol, ct_dtype vdt,
                     int32_t view0, int32_t nviews) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, pdt));
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, proj, pdt, "proj"));
    CT_TRY(views_ok(ctx, g, view0, nviews));
    DeviceGuard dg(ctx);
    return launch_adjoint(ctx, *g, proj, pdt, vol, vdt, view0, nviews);
}

ct_status ct_adjoint_test(ct_context *ctx, const ct_geometry *g, ct_dtype dt, uint64_t seed, double *lhs, double *rhs) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    if (dt != CT_F32 && dt != CT_F64) return fail(ctx, CT_ERR_DTYPE, "adjoint test runs in F32 or F64");
    if (!lhs || !rhs) return fail(ctx, CT_ERR_NULL, "result pointers are NULL");
    DeviceGuard dg(ctx);
    const size_t es = dtype_size(dt);
    const int64_t V = vox(g), P = pix(g) * g->nviews;
    ct_status st;
    void *x = slot(ctx, SLOT_TMP_VOL2, es * V, &st); if (!x) return st;
    void *aty = slot(ctx, SLOT_MISC, es * V, &st); if (!aty) return st;
    void *y = slot(ctx, SLOT_TMP_PROJ, es * P, &st); if (!y) return st;
    void *ax = slot(ctx, SLOT_TMP_PROJ2, es * P, &st); if (!ax) return st;
    CT_TRY(launch_randn(ctx, x, dt, V, seed));
    CT_TRY(launch_randn(ctx, y, dt, P, seed ^ 0x5DEECE66Dull));
    CT_TRY(launch_forward(ctx, *g, x, dt, ax, dt, nullptr, dt, 0, 1, g->nviews, 0));
    CT_TRY(launch_adjoint(ctx, *g, y, dt, aty, dt, 0, g->nviews));
    CT_TRY(reduce_dot(ctx, ax, y, P, dt, 0, lhs));
    CT_TRY(reduce_dot(ctx, x, aty, V, dt, 0, rhs));
    return CT_OK;
}

ct_status ct_power_iteration(ct_context *ctx, const ct_geometry *g, ct_dtype dt, int32_t niter, uint64_t seed,
                             double *lip) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    if (dt != CT_F32 && dt != CT_F64) return fail(ctx, CT_ERR_DTYPE, "power iteration runs in F32 or F64");
    if (!lip) return fail(ctx, CT_ERR_NULL, "result pointer is NULL");
    if (niter < 1) return fail(ctx, CT_ERR_INVALID_ARG, "niter must be >= 1");
    DeviceGuard dg(ctx);
    const size_t es = dtype_size(dt);
    const int64_t V = vox(g), P = pix(g) * g->nviews;
    ct_status st;
    void *v = slot(ctx, SLOT_TMP_VOL2, es * V, &st); if (!v) return st;
    void *w = slot(ctx, SLOT_MISC, es * V, &st); if (!w) return st;
    void *p = slot(ctx, SLOT_TMP_PROJ, es * P, &st); if (!p) return st;
    CT_TRY(launch_randn(ctx, v, dt, V, seed));
    double nrm2 = 0;
    CT_TRY(reduce_dot(ctx, v, v, V, dt, 0, &nrm2));
    if (!(nrm2 > 0)) return fail(ctx, CT_ERR_INVALID_ARG, "power iteration start vector is zero");
    CT_TRY(launch_axpby(ctx, v, v, dt, V, 1.0 / std::sqrt(nrm2), 0.0));
    double lam = 0;
    for (int it = 0; it < niter; ++it) {
        CT_TRY(launch_forward(ctx, *g, v, dt, p, dt, nullptr, dt, 0, 1, g->nviews, 0));
        CT_TRY(launch_adjoint(ctx, *g, p, dt, w, dt, 0, g->nviews));
        double wn2 = 0;
        CT_TRY(reduce_dot(ctx, w, w, V, dt, 0, &wn2));
        lam = std::sqrt(wn2);
        if (!(lam > 0)) break;
        // v = w / |w|
        CT_TRY(launch_axpby(ctx, w, v, dt, V, 1.0 / lam, 0.0));
    }
    *lip = lam;
    return CT_OK;
}

// Largest eigenvalue of the operator a SART update applies,
// T v = D^-1 B^T M^-1 A v: A the Siddon forward projection of every view, M
// the rays' lengths, B^T the voxel-driven backprojection and D its per-voxel
// chord sums (the reference's correction and update,
// V63/kernel_IterativeRecon_universal.cu:554-563, :785-807, taken over all
// views as one subset).  Each application is a correction pass against zero
// measurements, c = (0 - A v)/len, followed by a backprojection of c into a
// zeroed volume with relaxation -1, which leaves T v.  The operator averages
// along rays and over the rays through a voxel, so its spectral radius is 1
// up to the reconstruction mask and thresholds; the estimate replaces the
// fixed Lip_con of V63/main_recon.cu:58 (the step is 2 / lip with
// lip = 2 * radius).  The iterate starts from 1 + 0.1 N(0, 1) (its Perron
// component is then large) and is normalised with FP64 norms.
ct_status ct_sart_lipschitz(ct_context *ctx, const ct_geometry *g, ct_dtype dt, int32_t subset, int32_t niter,
                            uint64_t seed, double *lip) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    if (dt != CT_F32 && dt != CT_F64) return fail(ctx, CT_ERR_DTYPE, "SART power iteration runs in F32 or F64");
    if (!lip) return fail(ctx, CT_ERR_NULL, "result pointer is NULL");
    if (niter < 1) return fail(ctx, CT_ERR_INVALID_ARG, "niter must be >= 1");
    if (subset < 1 || subset > g->nviews) return fail(ctx, CT_ERR_INVALID_ARG, "subset must be in [1, nviews]");
    DeviceGuard dg(ctx);
    const size_t es = dtype_size(dt);
    const int64_t V = vox(g), P = pix(g) * subset;
    ct_status st;
    void *v = slot(ctx, SLOT_TMP_VOL2, es * V, &st); if (!v) return st;
    void *w = slot(ctx, SLOT_MISC, es * V, &st); if (!w) return st;
    void *c = slot(ctx, SLOT_TMP_PROJ, es * P, &st); if (!c) return st;
    void *zero = slot(ctx, SLOT_TMP_PROJ2, es * P, &st); if (!zero) return st;
    CT_TRY(launch_fill(ctx, zero, dt, P, 0.0));
    // start: 1 + 0.1 N(0, 1), near the operator's non-negative Perron vector
    CT_TRY(launch_randn(ctx, w, dt, V, seed));
    CT_TRY(launch_fill(ctx, v, dt, V, 1.0));
    CT_TRY(launch_axpby(ctx, w, v, dt, V, 0.1, 1.0));
    double nrm2 = 0;
    CT_TRY(reduce_dot(ctx, v, v, V, dt, 0, &nrm2));
    if (!(nrm2 > 0)) return fail(ctx, CT_ERR_INVALID_ARG, "power iteration start vector is zero");
    CT_TRY(launch_axpby(ctx, v, v, dt, V, 1.0 / std::sqrt(nrm2), 0.0));
    double lam = 0;
    for (int it = 0; it < niter; ++it) {
        // w = K v: per subset the correction of zero measurements,
        // c = -M_S^-1 A_S v, then w -= D_S^-1 B_S^T c (mode 1, step -1) into
        // an accumulator that every subset reads from the same v
        CT_TRY(launch_fill(ctx, w, dt, V, 0.0));
        for (int p0 = 0; p0 < g->nviews; p0 += subset) {
            const int cnt = std::min(subset, g->nviews - p0);
            CT_TRY(launch_forward_mv(ctx, *g, v, dt, c, dt, zero, dt, p0, 1, cnt, 1, p0));
            CT_TRY(launch_voxel_bp(ctx, *g, w, dt, c, p0, 1, cnt, -1.0, 1));
        }
        double wn2 = 0;
        CT_TRY(reduce_dot(ctx, w, w, V, dt, 0, &wn2));
        lam = std::sqrt(wn2);
        if (!(lam > 0)) break;
        CT_TRY(launch_axpby(ctx, w, v, dt, V, 1.0 / lam, 0.0));
    }
    *lip = 2.0 * lam;
    return CT_OK;
}

ct_status ct_fista_gamma(double t_prev, double *t_next, double *gamma) {
    if (!t_next || !gamma) return CT_ERR_NULL;
    if (!std::isfinite(t_prev) || t_prev < 1.0) return CT_ERR_INVALID_ARG;
    const double tk = (1.0 + std::sqrt(1.0 + 4.0 * t_prev * t_prev)) / 2.0;
    *t_next = tk;
    *gamma = (t_prev - 1.0) / tk;
    return CT_OK;
}

ct_status ct_fista_momentum(ct_context *ctx, void *x, void *x_prev, void *y, int64_t n, ct_dtype dt, double gamma) {
    CT_CTX(ctx);
    CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, x, dt, "x")); CT_TRY(check_ptr(ctx, x_prev, dt, "x_prev")); CT_TRY(check_ptr(ctx, y, dt, "y"));
    if (n < 0) return fail(ctx, CT_ERR_INVALID_ARG, "n must be >= 0");
    if (!std::isfinite(gamma)) return fail(ctx, CT_ERR_INVALID_ARG, "gamma must be finite");
    if (n == 0) return CT_OK;
    DeviceGuard dg(ctx);
    return launch_fista(ctx, x, x_prev, y, n, dt, gamma);
}

ct_status ct_tv_fgp(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz, void *vol, ct_dtype dt, double lambda,
                    int32_t niter) {
    CT_CTX(ctx);
    CT_TRY(dims_ok(ctx, nx, ny, nz)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, vol, dt, "vol"));
    if (!std::isfinite(lambda) || lambda < 0) return fail(ctx, CT_ERR_INVALID_ARG, "lambda must be finite and >= 0");
    if (niter < 0) return fail(ctx, CT_ERR_INVALID_ARG, "niter must be >= 0");
    if (lambda == 0) return CT_OK;
    DeviceGuard dg(ctx);
    return launch_fgp(ctx, nx, ny, nz, vol, dt, lambda, niter);
}

ct_status ct_tv_fgp_adjoint(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz, const void *r, const void *s,
                            const void *t, void *out, ct_dtype dt) {
    CT_CTX(ctx);
    CT_TRY(dims_ok(ctx, nx, ny, nz)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, r, dt, "p_r")); CT_TRY(check_ptr(ctx, s, dt, "p_s")); CT_TRY(check_ptr(ctx, t, dt, "p_t"));
    CT_TRY(check_ptr(ctx, out, dt, "out"));
    DeviceGuard dg(ctx);
    return launch_fgp_Dt(ctx, nx, ny, nz, r, s, t, out, dt);
}

ct_status ct_tv_fgp_forward_op(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz, const void *x, void *r, void *s,
                               void *t, ct_dtype dt) {
    CT_CTX(ctx);
    CT_TRY(dims_ok(ctx, nx, ny, nz)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, x, dt, "x"));
    CT_TRY(check_ptr(ctx, r, dt, "p_r")); CT_TRY(check_ptr(ctx, s, dt, "p_s")); CT_TRY(check_ptr(ctx, t, dt, "p_t"));
    DeviceGuard dg(ctx);
    return launch_fgp_D(ctx, nx, ny, nz, x, r, s, t, dt);
}

ct_status ct_tv_eps(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz, const void *vol, ct_dtype dt, double eps,
                    double *value) {
    CT_CTX(ctx);
    CT_TRY(dims_ok(ctx, nx, ny, nz)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, vol, dt, "vol"));
    if (!value) return fail(ctx, CT_ERR_NULL, "value is NULL");
    if (!std::isfinite(eps) || eps < 0) return fail(ctx, CT_ERR_INVALID_ARG, "eps must be finite and >= 0");
    DeviceGuard dg(ctx);
    return reduce_tv(ctx, vol, dt, nx, ny, nz, 1, eps, value);
}

ct_status ct_tv_eps_grad(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz, const void *vol, ct_dtype dt, double eps,
                         void *grad) {
    CT_CTX(ctx);
    CT_TRY(dims_ok(ctx, nx, ny, nz)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, vol, dt, "vol")); CT_TRY(check_ptr(ctx, grad, dt, "grad"));
    if (!std::isfinite(eps) || eps <= 0) return fail(ctx, CT_ERR_INVALID_ARG, "eps must be finite and > 0");
    DeviceGuard dg(ctx);
    return launch_tv_eps_grad(ctx, vol, grad, dt, nx, ny, nz, eps);
}

ct_status ct_residual_norm2(ct_context *ctx, const void *a, const void *b, int64_t n, ct_dtype dt, double *value) {
    CT_CTX(ctx);
    CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, a, dt, "a")); CT_TRY(check_ptr(ctx, b, dt, "b"));
    if (!value) return fail(ctx, CT_ERR_NULL, "value is NULL");
    if (n < 0) return fail(ctx, CT_ERR_INVALID_ARG, "n must be >= 0");
    DeviceGuard dg(ctx);
    return reduce_dot(ctx, a, b, n, dt, 1, value);
}

ct_status ct_tv_norm(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz, const void *vol, ct_dtype dt, double *value) {
    CT_CTX(ctx);
    CT_TRY(dims_ok(ctx, nx, ny, nz)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, vol, dt, "vol"));
    if (!value) return fail(ctx, CT_ERR_NULL, "value is NULL");
    DeviceGuard dg(ctx);
    return reduce_tv(ctx, vol, dt, nx, ny, nz, 0, 0.0, value);
}

ct_status ct_dot(ct_context *ctx, const void *a, const void *b, int64_t n, ct_dtype dt, double *value) {
    CT_CTX(ctx);
    CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, a, dt, "a")); CT_TRY(check_ptr(ctx, b, dt, "b"));
    if (!value) return fail(ctx, CT_ERR_NULL, "value is NULL");
    if (n < 0) return fail(ctx, CT_ERR_INVALID_ARG, "n must be >= 0");
    DeviceGuard dg(ctx);
    return reduce_dot(ctx, a, b, n, dt, 0, value);
}

ct_status ct_objective(ct_context *ctx, const ct_geometry *g, const void *vol, ct_dtype vdt, const void *meas,
                       ct_dtype mdt, double lambda_tv, double *data_fidelity, double *tv, double *objective) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g)); CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, mdt));
    CT_TRY(check_ptr(ctx, vol, vdt, "vol")); CT_TRY(check_ptr(ctx, meas, mdt, "meas"));
    if (!data_fidelity || !tv || !objective) return fail(ctx, CT_ERR_NULL, "result pointers are NULL");
    if (!std::isfinite(lambda_tv)) return fail(ctx, CT_ERR_INVALID_ARG, "lambda_tv must be finite");
    DeviceGuard dg(ctx);
    double dfid = 0, tvv = 0;
    CT_TRY(launch_residual(ctx, *g, vol, vdt, meas, mdt, &dfid));
    CT_TRY(reduce_tv(ctx, vol, vdt, g->nx, g->ny, g->nz, 0, 0.0, &tvv));
    *data_fidelity = dfid;
    *tv = tvv;
    *objective = dfid + 2.0 * lambda_tv * tvv;
    return CT_OK;
}

ct_status ct_fdk_weight(ct_context *ctx, const ct_geometry *g, const void *in, void *out, ct_dtype dt) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, in, dt, "proj_in")); CT_TRY(check_ptr(ctx, out, dt, "proj_out"));
    DeviceGuard dg(ctx);
    return launch_fdk_weight(ctx, *g, in, dt, out, dt);
}

ct_status ct_ramp_filter(ct_context *ctx, const ct_geometry *g, const void *in, void *out, ct_dtype dt, int32_t nimg,
                         ct_window window) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g)); CT_TRY(dt_ok(ctx, dt));
    CT_TRY(check_ptr(ctx, in, dt, "in")); CT_TRY(check_ptr(ctx, out, dt, "out"));
    if (nimg < 0) return fail(ctx, CT_ERR_INVALID_ARG, "nimg must be >= 0");
    if (window != CT_WINDOW_RAMP && window != CT_WINDOW_HANN) return fail(ctx, CT_ERR_INVALID_ARG, "unknown window");
    DeviceGuard dg(ctx);
    return launch_ramp(ctx, *g, in, dt, out, dt, nimg, window);
}

ct_status ct_fdk_backproject_ms(ct_context *ctx, const ct_geometry *g, const void *filt, ct_dtype pdt, void *vol,
                                ct_dtype vdt, int32_t view0, int32_t nviews, ct_fdk_sources sources) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g)); CT_TRY(dt_ok(ctx, pdt)); CT_TRY(dt_ok(ctx, vdt));
    CT_TRY(check_ptr(ctx, filt, pdt, "filtered")); CT_TRY(check_ptr(ctx, vol, vdt, "vol"));
    CT_TRY(views_ok(ctx, g, view0, nviews));
    if (g->style != CT_CONE) return fail(ctx, CT_ERR_GEOMETRY, "FDK is cone beam");
    if (sources < CT_FDK_INVERSE_SQUARE || sources > CT_FDK_EQUAL) return fail(ctx, CT_ERR_INVALID_ARG, "unknown source combination");
    DeviceGuard dg(ctx);
    // filtered holds the images of views [view0, view0+nviews), all sources of each
    return launch_fdk_bp(ctx, *g, filt, pdt, vol, vdt, view0, nviews, (int)sources);
}

ct_status ct_fdk_backproject(ct_context *ctx, const ct_geometry *g, const void *filt, ct_dtype pdt, void *vol,
                             ct_dtype vdt, int32_t view0, int32_t nviews) {
    return ct_fdk_backproject_ms(ctx, g, filt, pdt, vol, vdt, view0, nviews, CT_FDK_INVERSE_SQUARE);
}

ct_status ct_fdk_ms(ct_context *ctx, const ct_geometry *g, const void *proj, ct_dtype pdt, void *vol, ct_dtype vdt,
                    ct_window window, ct_fdk_sources sources) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g)); CT_TRY(dt_ok(ctx, pdt)); CT_TRY(dt_ok(ctx, vdt));
    CT_TRY(check_ptr(ctx, proj, pdt, "proj")); CT_TRY(check_ptr(ctx, vol, vdt, "vol"));
    if (g->style != CT_CONE) return fail(ctx, CT_ERR_GEOMETRY, "FDK is cone beam");
    if (window != CT_WINDOW_RAMP && window != CT_WINDOW_HANN) return fail(ctx, CT_ERR_INVALID_ARG, "unknown window");
    if (sources < CT_FDK_INVERSE_SQUARE || sources > CT_FDK_EQUAL) return fail(ctx, CT_ERR_INVALID_ARG, "unknown source combination");
    DeviceGuard dg(ctx);
    const ct_dtype fdt = vdt == CT_F64 ? CT_F64 : CT_F32;
    const int64_t P = pix(g) * g->nviews;
    ct_status st;
    void *w = slot(ctx, SLOT_FDK, dtype_size(fdt) * P, &st);
    if (!w) return st;
    const int nimg = g->nviews * (g->nsources < 1 ? 1 : g->nsources);
    CT_TRY(launch_ramp(ctx, *g, proj, pdt, w, fdt, nimg, window, true)); // cosine weight + filter
    return launch_fdk_bp(ctx, *g, w, fdt, vol, vdt, 0, g->nviews, (int)sources);
}

ct_status ct_fdk(ct_context *ctx, const ct_geometry *g, const void *proj, ct_dtype pdt, void *vol, ct_dtype vdt,
                 ct_window window) {
    return ct_fdk_ms(ctx, g, proj, pdt, vol, vdt, window, CT_FDK_INVERSE_SQUARE);
}

ct_status ct_sart_fista_iteration(ct_context *ctx, const ct_geometry *g, void *x, void *x_prev, ct_dtype vdt,
                                  const void *meas, ct_dtype mdt, double beta, int32_t subset_size, ct_order order,
                                  double fgp_lambda, int32_t fgp_niter, double *t_state) {
    CT_CTX(ctx);
    CT_TRY(validate(ctx, g));
    CT_TRY(dt_ok(ctx, vdt)); CT_TRY(dt_ok(ctx, mdt));
    CT_TRY(check_ptr(ctx, x, vdt, "x")); CT_TRY(check_ptr(ctx, x_prev, vdt, "x_prev")); CT_TRY(check_ptr(ctx, meas, mdt, "meas"));
    if (!t_state) return fail(ctx, CT_ERR_NULL, "t_state is NULL");
    if (subset_size < 1 || subset_size > g->nviews) return fail(ctx, CT_ERR_INVALID_ARG, "subset_size must be in [1, nviews]");
    if (!std::isfinite(beta)) return fail(ctx, CT_ERR_INVALID_ARG, "beta must be finite");
    if (!std::isfinite(fgp_lambda)) return fail(ctx, CT_ERR_INVALID_ARG, "fgp_lambda must be finite");
    if (fgp_lambda > 0 && fgp_niter < 0) return fail(ctx, CT_ERR_INVALID_ARG, "fgp_niter must be >= 0");
    double tk, gamma;
    if (ct_fista_gamma(*t_state, &tk, &gamma) != CT_OK) return fail(ctx, CT_ERR_INVALID_ARG, "t_state must be >= 1");
    DeviceGuard dg(ctx);
    // x holds y_k: SART sweep (V63/main_recon.cu:352), TV (:354-370), momentum (:372-384)
    CT_TRY(sweep(ctx, *g, x, vdt, meas, mdt, beta, subset_size, order));
    if (fgp_lambda > 0) CT_TRY(launch_fgp(ctx, g->nx, g->ny, g->nz, x, vdt, fgp_lambda, fgp_niter));
    CT_TRY(launch_fista(ctx, x, x_prev, x, vox(g), vdt, gamma));
    *t_state = tk;
    return CT_OK;
}

ct_status ct_convert(ct_context *ctx, const void *src, ct_dtype sdt, void *dst, ct_dtype ddt, int64_t n) {
    CT_CTX(ctx);
    CT_TRY(dt_ok(ctx, sdt)); CT_TRY(dt_ok(ctx, ddt));
    CT_TRY(check_ptr(ctx, src, sdt, "src")); CT_TRY(check_ptr(ctx, dst, ddt, "dst"));
    if (n < 0) return fail(ctx, CT_ERR_INVALID_ARG, "n must be >= 0");
    DeviceGuard dg(ctx);
    return launch_convert(ctx, src, sdt, dst, ddt, n);
}

} // extern "C"
