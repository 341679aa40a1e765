// io.hip - the callers on either side of the path: projection files in; volume
// dumps, objective logs and FISTA checkpoints out; the log transform of raw
// detector counts.
//
// The reference reads one raw float32 file per view,
// directory + "Projections/View_%.02f.dat" named by the view's angle, skipping
// leading rows with fseek (V63/main_recon.cu:196-224), and writes the volume
// as raw float32 ".recon" files and the per-iteration (TV, data fidelity,
// objective) triples as doubles (:437-487, SaveDeviceDataToFile
// V63/host_tool_functions.cu:2-15).  Here the files move through two pinned
// host buffers so that reading (several views in parallel threads) overlaps
// the copies to HBM, and storage types other than FP32 convert on the GPU.
#include "ct_internal.hpp"
#include <algorithm>
#include <cerrno>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#define IO_TRY(expr) do { ct_status _s = (expr); if (_s != CT_OK) return _s; } while (0)

namespace ct {

// ------------------------------------------------------------- file names
// File angle of view j (V63/main_recon.cu:202-209): float(j*us_rate +
// initialAngle); in tomosynthesis the second half of the views starts at
// 180 degrees; negative angles wrap by 360.
static float file_angle(const ct_geometry &g, int j, bool tomo) {
    const float us = (float)g.angle_step_deg, a0 = (float)g.angle0_deg;
    float a = (float)j * us + a0;
    if (tomo && j >= g.nviews / 2) a = 180.0f + (float)(j - g.nviews / 2) * us + a0;
    if (a < 0.0f) a = a + 360.0f;
    return a;
}

static std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
static std::string fmt(const char *f, ...) {
    va_list ap;
    va_start(ap, f);
    char tmp[1024];
    const int n = vsnprintf(tmp, sizeof tmp, f, ap);
    va_end(ap);
    if (n < 0) return std::string();
    if ((size_t)n < sizeof tmp) return std::string(tmp, (size_t)n);
    std::string s((size_t)n + 1, '\0');
    va_start(ap, f);
    vsnprintf(&s[0], s.size(), f, ap);
    va_end(ap);
    s.resize((size_t)n);
    return s;
}

static ct_status put(const std::string &s, char *buf, int32_t size) {
    if (!buf) return CT_ERR_NULL;
    if (size < 1 || (int64_t)s.size() + 1 > size) return CT_ERR_INVALID_ARG;
    memcpy(buf, s.c_str(), s.size() + 1);
    return CT_OK;
}

static std::string projection_path(const ct_geometry &g, int view, bool tomo, const char *dir) {
    return std::string(dir) + fmt("Projections/View_%.02f.dat", (double)file_angle(g, view, tomo));
}

// ---------------------------------------------------------- pinned staging
// Two pinned host buffers: file data moves through one while the other's
// copy runs on the context stream.
struct Pinned {
    void *p[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    ~Pinned() {
        for (int i = 0; i < 2; ++i) {
            if (ev[i]) { (void)hipEventSynchronize(ev[i]); (void)hipEventDestroy(ev[i]); }
            if (p[i]) (void)hipHostFree(p[i]);
        }
    }
    ct_status init(ct_context *ctx, size_t bytes) {
        for (int i = 0; i < 2; ++i) {
            if (hipHostMalloc(&p[i], bytes, hipHostMallocDefault) != hipSuccess) {
                p[i] = nullptr;
                return fail(ctx, CT_ERR_ALLOC, "pinned staging allocation failed");
            }
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                ev[i] = nullptr;
                return fail(ctx, CT_ERR_DEVICE, "event creation failed");
            }
        }
        return CT_OK;
    }
    // wait until buffer b's previous copy has finished
    ct_status ready(ct_context *ctx, int b) {
        if (!used[b]) return CT_OK;
        return hip_check(ctx, hipEventSynchronize(ev[b]), "staging copy");
    }
};

constexpr size_t IO_CHUNK = (size_t)64 << 20; // bytes per staging buffer

// Reads nbytes at byte offset `off` of `path` into dst; "" on success.
static std::string read_file(const std::string &path, int64_t off, void *dst, size_t nbytes) {
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return path + ": " + strerror(errno);
    std::string err;
    if (off > 0 && fseeko(fp, (off_t)off, SEEK_SET) != 0) err = path + ": seek failed";
    else if (fread(dst, 1, nbytes, fp) != nbytes) err = path + ": file shorter than the rows requested";
    fclose(fp);
    return err;
}

// Host -> device of `count` FP32 values staged in pinned buffer b, written to
// dst (element offset `at`) in storage type dt.
static ct_status upload(ct_context *ctx, Pinned &pb, int b, void *dst, ct_dtype dt, int64_t at, int64_t count) {
    const size_t bytes = (size_t)count * 4;
    ct_status st;
    if (dt == CT_F32) {
        st = hip_check(ctx, hipMemcpyAsync((float *)dst + at, pb.p[b], bytes, hipMemcpyHostToDevice, ctx->stream),
                       "upload");
        if (st != CT_OK) return st;
    } else {
        float *stage = (float *)slot(ctx, SLOT_IO, IO_CHUNK, &st);
        if (!stage) return st;
        st = hip_check(ctx, hipMemcpyAsync(stage, pb.p[b], bytes, hipMemcpyHostToDevice, ctx->stream), "upload");
        if (st != CT_OK) return st;
        st = launch_convert(ctx, stage, CT_F32, (char *)dst + (size_t)at * dtype_size(dt), dt, count);
        if (st != CT_OK) return st;
    }
    pb.used[b] = true;
    return hip_check(ctx, hipEventRecord(pb.ev[b], ctx->stream), "event");
}

// Device -> host of `count` values of src (element offset `at`, storage type
// dt) into pinned buffer b as FP32.
static ct_status download(ct_context *ctx, Pinned &pb, int b, const void *src, ct_dtype dt, int64_t at,
                          int64_t count) {
    const size_t bytes = (size_t)count * 4;
    ct_status st;
    if (dt == CT_F32) {
        st = hip_check(ctx, hipMemcpyAsync(pb.p[b], (const float *)src + at, bytes, hipMemcpyDeviceToHost,
                                           ctx->stream), "download");
    } else {
        float *stage = (float *)slot(ctx, SLOT_IO, IO_CHUNK, &st);
        if (!stage) return st;
        st = launch_convert(ctx, (const char *)src + (size_t)at * dtype_size(dt), dt, stage, CT_F32, count);
        if (st != CT_OK) return st;
        st = hip_check(ctx, hipMemcpyAsync(pb.p[b], stage, bytes, hipMemcpyDeviceToHost, ctx->stream), "download");
    }
    if (st != CT_OK) return st;
    pb.used[b] = true;
    return hip_check(ctx, hipEventRecord(pb.ev[b], ctx->stream), "event");
}

// ------------------------------------------------------------ log transform
// Line integrals from raw counts: p = -ln(max((I - dark)/(flat - dark), floor)),
// flat and dark one detector image each (flat absent: the scalar i0; dark
// absent: 0).  An HBM-bound stream: the hardware reciprocal and log2
// (relative error ~1e-7) replace the IEEE division and logf, and FP32 images
// whose pixel count is a multiple of 4 move as 16-byte vectors.
__device__ __forceinline__ float neg_log_ratio(float I, float F, float D, float floor_) {
    const float tr = fmaxf((I - D) * frcp(F - D), floor_);
    return -0.69314718055994531f * __builtin_amdgcn_logf(tr); // -ln(tr) = -log2(tr) ln 2
}

template <typename SI, typename SO>
__global__ __launch_bounds__(256) void k_log(const SI *__restrict__ raw, const float *__restrict__ flat,
                                             const float *__restrict__ dark, float i0, SO *__restrict__ out,
                                             int64_t pix, float floor_) {
    const int64_t base = (int64_t)blockIdx.y * pix;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pix; p += (int64_t)gridDim.x * blockDim.x) {
        const float D = dark ? dark[p] : 0.0f;
        const float F = flat ? flat[p] : i0;
        St<SO>::st(out, base + p, neg_log_ratio(ldf(raw, base + p), F, D, floor_));
    }
}

__global__ __launch_bounds__(256) void k_log4(const float4 *__restrict__ raw, const float4 *__restrict__ flat,
                                              const float4 *__restrict__ dark, float i0, float4 *__restrict__ out,
                                              int64_t pix4, float floor_) {
    const int64_t base = (int64_t)blockIdx.y * pix4;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < pix4; p += (int64_t)gridDim.x * blockDim.x) {
        const float4 I = raw[base + p];
        const float4 D = dark ? dark[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 F = flat ? flat[p] : make_float4(i0, i0, i0, i0);
        out[base + p] = make_float4(neg_log_ratio(I.x, F.x, D.x, floor_), neg_log_ratio(I.y, F.y, D.y, floor_),
                                    neg_log_ratio(I.z, F.z, D.z, floor_), neg_log_ratio(I.w, F.w, D.w, floor_));
    }
}

template <typename SI>
static void log_out(ct_context *ctx, const void *raw, const float *flat, const float *dark, float i0, void *out,
                    ct_dtype odt, int64_t pix, int nimg, float floor_) {
    const dim3 block(256);
    if (std::is_same<SI, float>::value && odt == CT_F32 && pix % 4 == 0 && ((uintptr_t)raw | (uintptr_t)out |
        (uintptr_t)flat | (uintptr_t)dark) % 16 == 0) {
        const int64_t pix4 = pix / 4;
        const dim3 grid((unsigned)std::min<int64_t>((pix4 + 255) / 256, 4096), (unsigned)nimg);
        hipLaunchKernelGGL(k_log4, grid, block, 0, ctx->stream, (const float4 *)raw, (const float4 *)flat,
                           (const float4 *)dark, i0, (float4 *)out, pix4, floor_);
        return;
    }
    const dim3 grid((unsigned)std::min<int64_t>((pix + 255) / 256, 4096), (unsigned)nimg);
    switch (odt) {
    case CT_F32: hipLaunchKernelGGL((k_log<SI, float>), grid, block, 0, ctx->stream, (const SI *)raw, flat, dark, i0, (float *)out, pix, floor_); break;
    case CT_F16: hipLaunchKernelGGL((k_log<SI, __half>), grid, block, 0, ctx->stream, (const SI *)raw, flat, dark, i0, (__half *)out, pix, floor_); break;
    case CT_BF16: hipLaunchKernelGGL((k_log<SI, __hip_bfloat16>), grid, block, 0, ctx->stream, (const SI *)raw, flat, dark, i0, (__hip_bfloat16 *)out, pix, floor_); break;
    default: break;
    }
}

// ----------------------------------------------------------- checkpoints
struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t dtype;
    int32_t nx, ny, nz, iteration;
    double vx, vy, vz;
    int32_t ncols, nrows, nviews, nsources;
    double du, dv, det_offset, dso, dod, angle_step_deg, angle0_deg, angle_shift_deg, src_z0, src_dz;
    int32_t x_offset, pad;
    double t_state;
    uint64_t nvox;
    uint64_t checksum;
};
static const char CKPT_MAGIC[8] = {'C', 'T', 'F', 'I', 'S', 'T', 'A', '1'};

static void header_geometry(CkptHeader &h, const ct_geometry &g) {
    h.nx = g.nx; h.ny = g.ny; h.nz = g.nz; h.x_offset = g.x_offset;
    h.vx = g.vx; h.vy = g.vy; h.vz = g.vz;
    h.ncols = g.ncols; h.nrows = g.nrows; h.nviews = g.nviews; h.nsources = g.nsources;
    h.du = g.du; h.dv = g.dv; h.det_offset = g.det_offset; h.dso = g.dso; h.dod = g.dod;
    h.angle_step_deg = g.angle_step_deg; h.angle0_deg = g.angle0_deg; h.angle_shift_deg = g.angle_shift_deg;
    h.src_z0 = g.src_z0; h.src_dz = g.src_dz;
}

// 64-bit word hash of the payload (multiply-xorshift per 8 bytes), chained
// across calls.
static uint64_t hash_words(uint64_t h, const void *p, size_t bytes) {
    const unsigned char *c = (const unsigned char *)p;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        memcpy(&w, c + i, 8);
        h = (h ^ w) * 0x100000001b3ull;
        h ^= h >> 29;
    }
    for (; i < bytes; ++i) h = (h ^ c[i]) * 0x100000001b3ull;
    return h;
}

// Device volume (raw storage bytes) <-> file, through the pinned buffers.
static ct_status save_raw(ct_context *ctx, Pinned &pb, FILE *fp, const void *src, size_t bytes, uint64_t &h) {
    size_t done = 0;
    int b = 0;
    size_t pending = 0; // bytes waiting in buffer 1-b
    while (done < bytes || pending) {
        size_t n = 0;
        if (done < bytes) {
            n = std::min(IO_CHUNK, bytes - done);
            IO_TRY(pb.ready(ctx, b));
            IO_TRY(hip_check(ctx, hipMemcpyAsync(pb.p[b], (const char *)src + done, n, hipMemcpyDeviceToHost,
                                                     ctx->stream), "download"));
            pb.used[b] = true;
            IO_TRY(hip_check(ctx, hipEventRecord(pb.ev[b], ctx->stream), "event"));
        }
        if (pending) {
            IO_TRY(pb.ready(ctx, 1 - b));
            h = hash_words(h, pb.p[1 - b], pending);
            if (fwrite(pb.p[1 - b], 1, pending, fp) != pending) return fail(ctx, CT_ERR_IO, "checkpoint write failed");
        }
        pending = n;
        done += n;
        b = 1 - b;
    }
    return CT_OK;
}

static ct_status load_raw(ct_context *ctx, Pinned &pb, FILE *fp, void *dst, size_t bytes, uint64_t &h) {
    int b = 0;
    for (size_t done = 0; done < bytes; b = 1 - b) {
        const size_t n = std::min(IO_CHUNK, bytes - done);
        IO_TRY(pb.ready(ctx, b));
        if (fread(pb.p[b], 1, n, fp) != n) return fail(ctx, CT_ERR_IO, "checkpoint file is truncated");
        h = hash_words(h, pb.p[b], n);
        IO_TRY(hip_check(ctx, hipMemcpyAsync((char *)dst + done, pb.p[b], n, hipMemcpyHostToDevice, ctx->stream),
                             "upload"));
        pb.used[b] = true;
        IO_TRY(hip_check(ctx, hipEventRecord(pb.ev[b], ctx->stream), "event"));
        done += n;
    }
    return CT_OK;
}

} // namespace ct

using namespace ct;

extern "C" {

ct_status ct_projection_path(const ct_geometry *g, int32_t view, int32_t tomo, const char *directory, char *buf,
                             int32_t size) {
    if (!g || !directory) return CT_ERR_NULL;
    IO_TRY(validate_geometry(nullptr, g));
    if (view < 0 || view >= g->nviews) return CT_ERR_INVALID_ARG;
    return put(projection_path(*g, view, tomo != 0, directory), buf, size);
}

ct_status ct_recon_path(const ct_geometry *g, double unit_um, const char *directory, const char *output_folder,
                        const char *object_name, int32_t iteration, double lambda_tv, double lambda_l1,
                        double lipschitz, char *buf, int32_t size) {
    if (!g || !directory || !output_folder || !object_name) return CT_ERR_NULL;
    IO_TRY(validate_geometry(nullptr, g));
    const float us = (float)g->angle_step_deg, a0 = (float)g->angle0_deg;
    const float a1 = a0 + (float)(g->nviews - 1) * us; // endAngle, V63/main_recon.cu:174
    const float um = (float)g->vx * (float)unit_um;
    std::string s;
    if (iteration < 0)
        s = fmt("%s/%s_%d_%d_%d_%.0fum_new_view_%d_(%.0f,%.0f)_TV_%.2f_L1_%.2f_Lip_%.2f.recon", output_folder,
                object_name, g->nx, g->ny, g->nz, (double)um, g->nviews, (double)a0, (double)a1,
                (double)(float)lambda_tv, (double)(float)lambda_l1, (double)(float)lipschitz);
    else
        s = fmt("%s/%s_%d_%d_%d_%.0fum_iterative_%d_view_%d_(%.0f,%.0f)_TV_%.2f_L1_%.2f_Lip_%.2f.recon",
                output_folder, object_name, g->nx, g->ny, g->nz, (double)um, iteration, g->nviews, (double)a0,
                (double)a1, (double)(float)lambda_tv, (double)(float)lambda_l1, (double)(float)lipschitz);
    return put(std::string(directory) + s, buf, size);
}

ct_status ct_objective_path(const ct_geometry *g, const char *directory, const char *output_folder,
                            const char *object_name, double lambda_tv, double lipschitz, char *buf, int32_t size) {
    if (!g || !directory || !output_folder || !object_name) return CT_ERR_NULL;
    IO_TRY(validate_geometry(nullptr, g));
    const float us = (float)g->angle_step_deg, a0 = (float)g->angle0_deg;
    const float a1 = a0 + (float)(g->nviews - 1) * us;
    return put(std::string(directory) +
                   fmt("%s/object_func_%s_view_%d_(%.0f,%.0f)_TV_%.2f_Lip_%.2f.bin", output_folder, object_name,
                       g->nviews, (double)a0, (double)a1, (double)(float)lambda_tv, (double)(float)lipschitz),
               buf, size);
}

ct_status ct_write_objective(const char *path, const double *values, int64_t count) {
    if (!path || (!values && count > 0)) return CT_ERR_NULL;
    if (count < 0) return CT_ERR_INVALID_ARG;
    FILE *fp = fopen(path, "wb");
    if (!fp) return CT_ERR_IO;
    const bool ok = fwrite(values, sizeof(double), (size_t)count, fp) == (size_t)count;
    return (fclose(fp) == 0 && ok) ? CT_OK : CT_ERR_IO;
}

ct_status ct_read_projections(ct_context *ctx, const ct_geometry *g, const char *directory, int32_t tomo,
                              int32_t row0, void *proj, ct_dtype dt, int32_t view0, int32_t nviews) {
    if (!ctx) return CT_ERR_NULL;
    IO_TRY(validate_geometry(ctx, g));
    if (!directory) return fail(ctx, CT_ERR_NULL, "directory is NULL");
    if (!dtype_valid(dt) || dt == CT_F64) return fail(ctx, CT_ERR_DTYPE, "projections are read into F32/F16/BF16");
    IO_TRY(check_ptr(ctx, proj, dt, "proj"));
    if (g->nsources != 1) return fail(ctx, CT_ERR_GEOMETRY, "projection files hold one source per view");
    if (view0 < 0 || nviews < 0 || (int64_t)view0 + nviews > g->nviews || row0 < 0)
        return fail(ctx, CT_ERR_INVALID_ARG, "view range outside [0, nviews) or negative row offset");
    if (nviews == 0) return CT_OK;
    DeviceGuard dg(ctx);
    const int64_t img = (int64_t)g->ncols * g->nrows; // floats per view
    if ((size_t)img * 4 > IO_CHUNK) return fail(ctx, CT_ERR_OVERFLOW, "one view exceeds the 64 MiB staging buffer");
    const int per = (int)std::min<int64_t>(nviews, (int64_t)(IO_CHUNK / ((size_t)img * 4)));
    Pinned pb;
    IO_TRY(pb.init(ctx, (size_t)per * img * 4));
    const int64_t off = (int64_t)row0 * g->ncols * 4;
    int b = 0;
    for (int v = 0; v < nviews; v += per, b = 1 - b) {
        const int n = std::min(per, nviews - v);
        IO_TRY(pb.ready(ctx, b));
        // read this chunk's views in parallel
        const int nt = std::min(n, 8);
        std::vector<std::string> err((size_t)nt);
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                for (int i = t; i < n; i += nt) {
                    const std::string p = projection_path(*g, view0 + v + i, tomo != 0, directory);
                    std::string e = read_file(p, off, (float *)pb.p[b] + (size_t)i * img, (size_t)img * 4);
                    if (!e.empty() && err[t].empty()) err[t] = e;
                }
            });
        for (auto &x : th) x.join();
        for (auto &e : err)
            if (!e.empty()) {
                (void)hipStreamSynchronize(ctx->stream);
                return fail(ctx, CT_ERR_IO, "projection read: " + e);
            }
        IO_TRY(upload(ctx, pb, b, proj, dt, (int64_t)v * img, (int64_t)n * img));
    }
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "projection read");
}

ct_status ct_write_volume(ct_context *ctx, const ct_geometry *g, const void *vol, ct_dtype dt, const char *path) {
    if (!ctx) return CT_ERR_NULL;
    IO_TRY(validate_geometry(ctx, g));
    if (!path) return fail(ctx, CT_ERR_NULL, "path is NULL");
    if (!dtype_valid(dt)) return fail(ctx, CT_ERR_DTYPE, "unknown dtype");
    IO_TRY(check_ptr(ctx, vol, dt, "vol"));
    DeviceGuard dg(ctx);
    const int64_t V = (int64_t)g->nx * g->ny * g->nz;
    const int64_t per = (int64_t)(IO_CHUNK / 4);
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(ctx, CT_ERR_IO, std::string(path) + ": " + strerror(errno));
    ct_status st = CT_OK;
    {
        Pinned pb;
        st = pb.init(ctx, IO_CHUNK);
        // chunk c downloads into buffer c&1 while chunk c-1 is written out
        int64_t prev_at = -1, prev_n = 0;
        int b = 0;
        for (int64_t at = 0; st == CT_OK && (at < V || prev_n); b = 1 - b) {
            const int64_t n = at < V ? std::min(per, V - at) : 0;
            if (n) {
                st = pb.ready(ctx, b);
                if (st == CT_OK) st = download(ctx, pb, b, vol, dt, at, n);
            }
            if (st == CT_OK && prev_n) {
                st = pb.ready(ctx, 1 - b);
                if (st == CT_OK && fwrite(pb.p[1 - b], 4, (size_t)prev_n, fp) != (size_t)prev_n)
                    st = fail(ctx, CT_ERR_IO, std::string(path) + ": write failed");
            }
            prev_at = at;
            prev_n = n;
            at += n;
        }
        (void)prev_at;
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (fclose(fp) != 0 && st == CT_OK) st = fail(ctx, CT_ERR_IO, std::string(path) + ": close failed");
    return st;
}

ct_status ct_read_volume(ct_context *ctx, const ct_geometry *g, const char *path, void *vol, ct_dtype dt) {
    if (!ctx) return CT_ERR_NULL;
    IO_TRY(validate_geometry(ctx, g));
    if (!path) return fail(ctx, CT_ERR_NULL, "path is NULL");
    if (!dtype_valid(dt)) return fail(ctx, CT_ERR_DTYPE, "unknown dtype");
    IO_TRY(check_ptr(ctx, vol, dt, "vol"));
    DeviceGuard dg(ctx);
    const int64_t V = (int64_t)g->nx * g->ny * g->nz;
    const int64_t per = (int64_t)(IO_CHUNK / 4);
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(ctx, CT_ERR_IO, std::string(path) + ": " + strerror(errno));
    ct_status st = CT_OK;
    {
        Pinned pb;
        st = pb.init(ctx, IO_CHUNK);
        int b = 0;
        for (int64_t at = 0; st == CT_OK && at < V; b = 1 - b) {
            const int64_t n = std::min(per, V - at);
            st = pb.ready(ctx, b);
            if (st == CT_OK && fread(pb.p[b], 4, (size_t)n, fp) != (size_t)n)
                st = fail(ctx, CT_ERR_IO, std::string(path) + ": file holds fewer than nx*ny*nz floats");
            if (st == CT_OK) st = upload(ctx, pb, b, vol, dt, at, n);
            at += n;
        }
        (void)hipStreamSynchronize(ctx->stream);
    }
    fclose(fp);
    return st;
}

ct_status ct_log_transform(ct_context *ctx, const void *raw, ct_dtype raw_dtype, const float *flat, const float *dark,
                           double i0, void *out, ct_dtype out_dtype, int64_t pixels, int32_t nimg, double floor_value) {
    if (!ctx) return CT_ERR_NULL;
    if (!dtype_valid(raw_dtype) || raw_dtype == CT_F64 || !dtype_valid(out_dtype) || out_dtype == CT_F64)
        return fail(ctx, CT_ERR_DTYPE, "log transform: F32/F16/BF16 storage");
    IO_TRY(check_ptr(ctx, raw, raw_dtype, "raw"));
    IO_TRY(check_ptr(ctx, out, out_dtype, "out"));
    if (pixels < 1 || nimg < 0 || nimg > 65535) return fail(ctx, CT_ERR_INVALID_ARG, "pixels >= 1, 0 <= nimg <= 65535");
    if (!(floor_value >= 1.17549435e-38) || !(floor_value <= 1.0))
        return fail(ctx, CT_ERR_INVALID_ARG, "floor must lie in [FLT_MIN, 1]");
    if (!flat && !(i0 > 0 && std::isfinite(i0))) return fail(ctx, CT_ERR_INVALID_ARG, "i0 must be > 0 without a flat field");
    if (nimg == 0) return CT_OK;
    DeviceGuard dg(ctx);
    switch (raw_dtype) {
    case CT_F32: log_out<float>(ctx, raw, flat, dark, (float)i0, out, out_dtype, pixels, nimg, (float)floor_value); break;
    case CT_F16: log_out<__half>(ctx, raw, flat, dark, (float)i0, out, out_dtype, pixels, nimg, (float)floor_value); break;
    default: log_out<__hip_bfloat16>(ctx, raw, flat, dark, (float)i0, out, out_dtype, pixels, nimg, (float)floor_value); break;
    }
    return launch_check(ctx, "log transform");
}

ct_status ct_save_checkpoint(ct_context *ctx, const ct_geometry *g, const char *path, const void *x,
                             const void *x_prev, ct_dtype dt, double t_state, int32_t iteration) {
    if (!ctx) return CT_ERR_NULL;
    IO_TRY(validate_geometry(ctx, g));
    if (!path) return fail(ctx, CT_ERR_NULL, "path is NULL");
    if (!dtype_valid(dt)) return fail(ctx, CT_ERR_DTYPE, "unknown dtype");
    IO_TRY(check_ptr(ctx, x, dt, "x"));
    IO_TRY(check_ptr(ctx, x_prev, dt, "x_prev"));
    DeviceGuard dg(ctx);
    CkptHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, CKPT_MAGIC, 8);
    h.version = 1;
    h.dtype = (uint32_t)dt;
    header_geometry(h, *g);
    h.iteration = iteration;
    h.t_state = t_state;
    h.nvox = (uint64_t)g->nx * g->ny * g->nz;
    const size_t bytes = (size_t)h.nvox * dtype_size(dt);
    const std::string tmp = std::string(path) + ".part";
    FILE *fp = fopen(tmp.c_str(), "wb");
    if (!fp) return fail(ctx, CT_ERR_IO, tmp + ": " + strerror(errno));
    ct_status st = CT_OK;
    uint64_t hh = 0xcbf29ce484222325ull;
    {
        Pinned pb;
        st = pb.init(ctx, IO_CHUNK);
        if (st == CT_OK && fwrite(&h, sizeof h, 1, fp) != 1) st = fail(ctx, CT_ERR_IO, tmp + ": write failed");
        if (st == CT_OK) st = save_raw(ctx, pb, fp, x, bytes, hh);
        if (st == CT_OK) st = save_raw(ctx, pb, fp, x_prev, bytes, hh);
        (void)hipStreamSynchronize(ctx->stream);
    }
    h.checksum = hh;
    if (st == CT_OK && (fseeko(fp, 0, SEEK_SET) != 0 || fwrite(&h, sizeof h, 1, fp) != 1))
        st = fail(ctx, CT_ERR_IO, tmp + ": header write failed");
    if (fclose(fp) != 0 && st == CT_OK) st = fail(ctx, CT_ERR_IO, tmp + ": close failed");
    // the finished file replaces the old one in one step
    if (st == CT_OK && rename(tmp.c_str(), path) != 0) st = fail(ctx, CT_ERR_IO, std::string(path) + ": rename failed");
    if (st != CT_OK) remove(tmp.c_str());
    return st;
}

ct_status ct_load_checkpoint(ct_context *ctx, const ct_geometry *g, const char *path, void *x, void *x_prev,
                             ct_dtype dt, double *t_state, int32_t *iteration) {
    if (!ctx) return CT_ERR_NULL;
    IO_TRY(validate_geometry(ctx, g));
    if (!path || !t_state || !iteration) return fail(ctx, CT_ERR_NULL, "path, t_state or iteration is NULL");
    if (!dtype_valid(dt)) return fail(ctx, CT_ERR_DTYPE, "unknown dtype");
    IO_TRY(check_ptr(ctx, x, dt, "x"));
    IO_TRY(check_ptr(ctx, x_prev, dt, "x_prev"));
    DeviceGuard dg(ctx);
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(ctx, CT_ERR_IO, std::string(path) + ": " + strerror(errno));
    CkptHeader h;
    ct_status st = CT_OK;
    if (fread(&h, sizeof h, 1, fp) != 1 || memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.version != 1)
        st = fail(ctx, CT_ERR_IO, std::string(path) + ": not a checkpoint of this format");
    CkptHeader want;
    memset(&want, 0, sizeof want);
    header_geometry(want, *g);
    if (st == CT_OK && h.dtype != (uint32_t)dt) st = fail(ctx, CT_ERR_DTYPE, "checkpoint holds another storage type");
    if (st == CT_OK &&
        (h.nx != want.nx || h.ny != want.ny || h.nz != want.nz || h.x_offset != want.x_offset || h.vx != want.vx ||
         h.vy != want.vy || h.vz != want.vz || h.ncols != want.ncols || h.nrows != want.nrows ||
         h.nviews != want.nviews || h.nsources != want.nsources || h.du != want.du || h.dv != want.dv ||
         h.det_offset != want.det_offset || h.dso != want.dso || h.dod != want.dod ||
         h.angle_step_deg != want.angle_step_deg || h.angle0_deg != want.angle0_deg ||
         h.angle_shift_deg != want.angle_shift_deg || h.src_z0 != want.src_z0 || h.src_dz != want.src_dz ||
         h.nvox != (uint64_t)g->nx * g->ny * g->nz))
        st = fail(ctx, CT_ERR_GEOMETRY, "checkpoint was written for another geometry");
    uint64_t hh = 0xcbf29ce484222325ull;
    if (st == CT_OK) {
        const size_t bytes = (size_t)h.nvox * dtype_size(dt);
        Pinned pb;
        st = pb.init(ctx, IO_CHUNK);
        if (st == CT_OK) st = load_raw(ctx, pb, fp, x, bytes, hh);
        if (st == CT_OK) st = load_raw(ctx, pb, fp, x_prev, bytes, hh);
        (void)hipStreamSynchronize(ctx->stream);
    }
    fclose(fp);
    if (st == CT_OK && hh != h.checksum) st = fail(ctx, CT_ERR_IO, std::string(path) + ": checksum mismatch");
    if (st == CT_OK) {
        *t_state = h.t_state;
        *iteration = h.iteration;
    }
    return st;
}

} // extern "C"
