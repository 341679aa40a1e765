// compat.cpp - reference-named host wrappers (include/ct_compat.h) over the C ABI.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <mutex>

#include "../../include/ct_compat.h"

namespace {

struct Compat {
    ct_context *ctx = nullptr;
    ct_geometry g{};
    bool have_geom = false;
    int status = 0;
    std::string error;
    double ip1 = 0, ip2 = 0;   // the last check's inner products
    void *dvol = nullptr, *dproj = nullptr, *dvol2 = nullptr;
    size_t vol_bytes = 0, proj_bytes = 0, vol2_bytes = 0;
};

Compat &C() {
    static Compat c;
    return c;
}
std::mutex &mtx() {
    static std::mutex m;
    return m;
}

int64_t nvox() { return (int64_t)C().g.nx * C().g.ny * C().g.nz; }
int64_t npix() { return (int64_t)C().g.ncols * C().g.nrows * C().g.nsources * C().g.nviews; }

bool set_err(int st, const std::string &msg) {
    C().status = st;
    C().error = msg;
    if (st) std::fprintf(stderr, "ct_compat: %s\n", msg.c_str());
    return st == 0;
}

bool ok(ct_status st, const char *what) {
    if (st == CT_OK) return true;
    return set_err(st, std::string(what) + ": " + ct_last_error(C().ctx));
}

bool ready(const char *what) {
    if (!C().have_geom || !C().ctx) return set_err(CT_ERR_INVALID_ARG, std::string(what) + ": call ct_compat_set_geometry first");
    return true;
}

bool is_device(const void *p) {
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) { hipGetLastError(); return false; }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void *grow(void *&p, size_t &have, size_t need) {
    if (have >= need) return p;
    if (p) { hipDeviceSynchronize(); hipFree(p); p = nullptr; have = 0; }
    if (hipMalloc(&p, need) != hipSuccess) { hipGetLastError(); p = nullptr; return nullptr; }
    have = need;
    return p;
}

// Device view of a caller array: the pointer itself if it is device memory,
// else a staging buffer (filled from the host when `upload`).
float *dev_of(float *p, size_t bytes, void *&stage, size_t &have, bool upload, bool *staged) {
    if (is_device(p)) { *staged = false; return p; }
    if (!grow(stage, have, bytes)) { set_err(CT_ERR_ALLOC, "staging allocation failed"); return nullptr; }
    if (upload && hipMemcpy(stage, p, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        hipGetLastError();
        set_err(CT_ERR_DEVICE, "host to device copy failed");
        return nullptr;
    }
    *staged = true;
    return (float *)stage;
}

bool download(float *host, const void *dev, size_t bytes) {
    if (hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        hipGetLastError();
        return set_err(CT_ERR_DEVICE, "device to host copy failed");
    }
    return true;
}

bool sync() {
    return ok(ct_synchronize(C().ctx), "synchronize");
}

// <A0 x, A0 x> and <x, B0 A0 x> for view 0, where B0 is `matched` ? the
// matched adjoint : the voxel-driven backprojection with command 0 into a
// zeroed volume (x = W where L >= 1e-6 vx), as the reference's check runs it
// (V63/host_IterativeRecon_CBCT.c:131-204).  Prints both, like the reference,
// and keeps them for ct_compat_last_inner_products.
void check_joint(float *h_volume, bool matched, const char *what) {
    if (!ready(what)) return;
    Compat &c = C();
    bool sv;
    const int64_t P1 = (int64_t)c.g.ncols * c.g.nrows * c.g.nsources;
    float *v = dev_of(h_volume, 4 * nvox(), c.dvol, c.vol_bytes, true, &sv);
    if (!v) return;
    float *p = (float *)grow(c.dproj, c.proj_bytes, 4 * P1);
    float *w = (float *)grow(c.dvol2, c.vol2_bytes, 4 * nvox());
    if (!p || !w) { set_err(CT_ERR_ALLOC, "allocation failed"); return; }
    if (!ok(ct_forward(c.ctx, &c.g, v, CT_F32, p, CT_F32, 0, 1), "forward")) return;
    if (matched) {
        if (!ok(ct_adjoint(c.ctx, &c.g, p, CT_F32, w, CT_F32, 0, 1), "adjoint")) return;
    } else {
        if (!sync()) return;
        if (hipMemset(w, 0, 4 * nvox()) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            hipGetLastError();
            set_err(CT_ERR_DEVICE, "clear failed");
            return;
        }
        if (!ok(ct_backproject_voxel(c.ctx, &c.g, w, CT_F32, p, 0, 1, 1.0, 0), "voxel backprojection")) return;
    }
    double ip1 = 0, ip2 = 0;
    if (!ok(ct_dot(c.ctx, p, p, P1, CT_F32, &ip1), "dot")) return;
    if (!ok(ct_dot(c.ctx, v, w, nvox(), CT_F32, &ip2), "dot")) return;
    c.ip1 = ip1;
    c.ip2 = ip2;
    std::printf("The 1st inner product = %f\n", ip1);
    std::printf("The 2nd inner product = %f\n", ip2);
    set_err(0, "");
}


} // namespace

extern "C" {

int ct_compat_set_geometry(const ct_geometry *g) {
    std::lock_guard<std::mutex> l(mtx());
    ct_status st = ct_validate_geometry(g);
    if (st != CT_OK) { set_err(st, "invalid geometry"); return st; }
    if (!C().ctx) {
        st = ct_create(&C().ctx, 0);
        if (st != CT_OK) { set_err(st, "no usable HIP device"); return st; }
    }
    C().g = *g;
    C().have_geom = true;
    set_err(0, "");
    return 0;
}

int ct_compat_last_status(void) { return C().status; }
const char *ct_compat_last_error(void) { return C().error.c_str(); }

void ct_compat_release(void) {
    std::lock_guard<std::mutex> l(mtx());
    if (C().ctx) {
        ct_synchronize(C().ctx);
        if (C().dvol) hipFree(C().dvol);
        if (C().dproj) hipFree(C().dproj);
        if (C().dvol2) hipFree(C().dvol2);
        ct_destroy(C().ctx);
    }
    C() = Compat();
}

void Forward_3D_ray_driven_siddon(float *h_volume, float *h_proj_data) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("Forward_3D_ray_driven_siddon")) return;
    Compat &c = C();
    bool sv, sp;
    float *v = dev_of(h_volume, 4 * nvox(), c.dvol, c.vol_bytes, true, &sv);
    float *p = v ? dev_of(h_proj_data, 4 * npix(), c.dproj, c.proj_bytes, false, &sp) : nullptr;
    if (!v || !p) return;
    if (!ok(ct_forward(c.ctx, &c.g, v, CT_F32, p, CT_F32, 0, c.g.nviews), "forward")) return;
    if (!sync()) return;
    if (sp) download(h_proj_data, p, 4 * npix());
    set_err(0, "");
}

void Forward_3D_ray_driven_siddon_TBCT(float *h_volume, float *h_proj_forward) {
    Forward_3D_ray_driven_siddon(h_volume, h_proj_forward);
}

void Reconstruction_3D_ray_driven_CBCT(float *h_volume, float *h_proj_data, float beta_temp) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("Reconstruction_3D_ray_driven_CBCT")) return;
    Compat &c = C();
    bool sv, sp;
    float *v = dev_of(h_volume, 4 * nvox(), c.dvol, c.vol_bytes, true, &sv);
    float *p = v ? dev_of(h_proj_data, 4 * npix(), c.dproj, c.proj_bytes, true, &sp) : nullptr;
    if (!v || !p) return;
    if (!ok(ct_sart_sweep(c.ctx, &c.g, v, CT_F32, p, CT_F32, beta_temp, 1, CT_ORDER_ASCENDING), "SART sweep")) return;
    if (!sync()) return;
    if (sv) download(h_volume, v, 4 * nvox());
    set_err(0, "");
}

void Reconstruction_3D_ray_driven_TBCT(float *h_volume, float *h_proj_data, float beta_temp) {
    Reconstruction_3D_ray_driven_CBCT(h_volume, h_proj_data, beta_temp);
}

void CheckMatchedJointOperator(float *h_volume) {
    std::lock_guard<std::mutex> l(mtx());
    check_joint(h_volume, false, "CheckMatchedJointOperator");
}

void ct_compat_check_matched_adjoint(float *h_volume) {
    std::lock_guard<std::mutex> l(mtx());
    check_joint(h_volume, true, "ct_compat_check_matched_adjoint");
}

int ct_compat_last_inner_products(double *ip1, double *ip2) {
    if (!ip1 || !ip2) return CT_ERR_NULL;
    *ip1 = C().ip1;
    *ip2 = C().ip2;
    return 0;
}



void FGP_denoise_GPUx7_exact(float *h_f_noise, float lambda, int Niter_denoise) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("FGP_denoise_GPUx7_exact")) return;
    Compat &c = C();
    bool sv;
    float *v = dev_of(h_f_noise, 4 * nvox(), c.dvol, c.vol_bytes, true, &sv);
    if (!v) return;
    if (!ok(ct_tv_fgp(c.ctx, c.g.nx, c.g.ny, c.g.nz, v, CT_F32, lambda, Niter_denoise), "FGP")) return;
    if (!sync()) return;
    if (sv) download(h_f_noise, v, 4 * nvox());
    set_err(0, "");
}

void F_Dnoise(float *Anoise, float *Aclean, float lambda, int N) {
    {
        std::lock_guard<std::mutex> l(mtx());
        if (!ready("F_Dnoise")) return;
        if (Aclean != Anoise) {
            const size_t bytes = 4 * nvox();
            const bool dn = is_device(Anoise), dc = is_device(Aclean);
            hipError_t e = hipMemcpy(Aclean, Anoise, bytes,
                                     dn ? (dc ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost)
                                        : (dc ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
            if (e != hipSuccess) { hipGetLastError(); set_err(CT_ERR_DEVICE, "copy failed"); return; }
        }
    }
    FGP_denoise_GPUx7_exact(Aclean, lambda, N);
}

void Reconstruction_FDK(float *h_volume, float *h_proj_data) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("Reconstruction_FDK")) return;
    Compat &c = C();
    bool sv, sp;
    float *v = dev_of(h_volume, 4 * nvox(), c.dvol, c.vol_bytes, true, &sv);
    float *p = v ? dev_of(h_proj_data, 4 * npix(), c.dproj, c.proj_bytes, true, &sp) : nullptr;
    if (!v || !p) return;
    if (!ok(ct_fdk(c.ctx, &c.g, p, CT_F32, v, CT_F32, CT_WINDOW_RAMP), "FDK")) return;
    if (!sync()) return;
    if (sv) download(h_volume, v, 4 * nvox());
    set_err(0, "");
}

void RampFiltering_GPU(float *g_in, float *g_out) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("RampFiltering_GPU")) return;
    Compat &c = C();
    bool si, so;
    float *i = dev_of(g_in, 4 * npix(), c.dproj, c.proj_bytes, true, &si);
    if (!i) return;
    float *o = i;
    if (g_out != g_in) {
        o = dev_of(g_out, 4 * npix(), c.dvol2, c.vol2_bytes, false, &so);
        if (!o) return;
    } else {
        so = si;
    }
    if (!ok(ct_ramp_filter(c.ctx, &c.g, i, o, CT_F32, c.g.nviews * c.g.nsources, CT_WINDOW_RAMP), "ramp filter")) return;
    if (!sync()) return;
    if (so) download(g_out, o, 4 * npix());
    set_err(0, "");
}

double L2_norm_gpu(float *d_a, float *d_b) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("L2_norm_gpu")) return 0.0;
    Compat &c = C();
    bool sa, sb;
    float *a = dev_of(d_a, 4 * npix(), c.dproj, c.proj_bytes, true, &sa);
    float *b = a ? dev_of(d_b, 4 * npix(), c.dvol2, c.vol2_bytes, true, &sb) : nullptr;
    if (!a || !b) return 0.0;
    double v = 0;
    if (!ok(ct_residual_norm2(c.ctx, a, b, npix(), CT_F32, &v), "L2 norm")) return 0.0;
    // the reference leaves (a-b)^2 in d_a (V63/kernel_tool_functions.cu:30-34)
    std::vector<float> ha(npix()), hb(npix());
    if (!download(ha.data(), a, 4 * npix()) || !download(hb.data(), b, 4 * npix())) return 0.0;
    for (int64_t i = 0; i < npix(); ++i) ha[i] = (ha[i] - hb[i]) * (ha[i] - hb[i]);
    if (sa) std::memcpy(d_a, ha.data(), 4 * npix());
    else if (hipMemcpy(d_a, ha.data(), 4 * npix(), hipMemcpyHostToDevice) != hipSuccess) {
        hipGetLastError();
        set_err(CT_ERR_DEVICE, "copy failed");
        return 0.0;
    }
    set_err(0, "");
    return v;
}

double TV_norm_gpu(float *d_volume) {
    std::lock_guard<std::mutex> l(mtx());
    if (!ready("TV_norm_gpu")) return 0.0;
    Compat &c = C();
    bool sv;
    float *v = dev_of(d_volume, 4 * nvox(), c.dvol, c.vol_bytes, true, &sv);
    if (!v) return 0.0;
    double r = 0;
    if (!ok(ct_tv_norm(c.ctx, c.g.nx, c.g.ny, c.g.nz, v, CT_F32, &r), "TV norm")) return 0.0;
    set_err(0, "");
    return r;
}

void SaveDeviceDataToFile(float *d_data, int data_size, char filename[]) {
    std::lock_guard<std::mutex> l(mtx());
    if (!d_data || !filename || data_size < 0) { set_err(CT_ERR_NULL, "SaveDeviceDataToFile: bad arguments"); return; }
    const size_t bytes = sizeof(float) * (size_t)data_size;
    std::vector<float> host;
    const float *src = d_data;
    if (is_device(d_data)) {
        host.resize((size_t)data_size);
        if (!download(host.data(), d_data, bytes)) return;
        src = host.data();
    }
    std::FILE *fp = std::fopen(filename, "wb");
    if (!fp) { set_err(CT_ERR_IO, std::string("can not open file to write data ") + filename); return; }
    const bool w = std::fwrite(src, 1, bytes, fp) == bytes;
    if (std::fclose(fp) != 0 || !w) { set_err(CT_ERR_IO, std::string("write failed: ") + filename); return; }
    set_err(0, "");
}

double L2_norm_cpu(const float *a, const float *b, int SIZE) {
    double s = 0;
    for (int i = 0; i < SIZE; ++i) { double e = (double)a[i] - (double)b[i]; s += e * e; }
    return s;
}

double TV_norm_cpu(float *f) {
    std::lock_guard<std::mutex> l(mtx());
    if (!C().have_geom) { set_err(CT_ERR_INVALID_ARG, "TV_norm_cpu: call ct_compat_set_geometry first"); return 0.0; }
    const int M = C().g.nx, N = C().g.ny, Z = C().g.nz;
    const int64_t MN = (int64_t)M * N;
    double acc = 0;
    for (int k = 0; k < Z; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < M; ++i) {
                const int64_t v = k * MN + (int64_t)j * M + i;
                const double c = f[v];
                const double dx = i < M - 1 ? c - f[v + 1] : 0.0;
                const double dy = j < N - 1 ? c - f[v + M] : 0.0;
                const double dz = k < Z - 1 ? c - f[v + MN] : 0.0;
                acc += std::sqrt(dx * dx + dy * dy + dz * dz);
            }
    return acc;
}

} // extern "C"
