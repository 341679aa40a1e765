// ct_internal.hpp - shared device/host helpers of the MI355X CT core.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <type_traits>
#include <cstddef>
#include <string>
#include <vector>
#include <mutex>

#include "../../include/ct_core.h"

namespace ct {

// ---------------------------------------------------------------- storage
template <typename S> struct St;
template <> struct St<float> {
    __device__ __forceinline__ static float ld(const float *p, int64_t i) { return p[i]; }
    __device__ __forceinline__ static void st(float *p, int64_t i, float v) { p[i] = v; }
};
template <> struct St<double> {
    __device__ __forceinline__ static double ld(const double *p, int64_t i) { return p[i]; }
    __device__ __forceinline__ static void st(double *p, int64_t i, double v) { p[i] = v; }
};
template <> struct St<__half> {
    __device__ __forceinline__ static float ld(const __half *p, int64_t i) { return __half2float(p[i]); }
    __device__ __forceinline__ static void st(__half *p, int64_t i, float v) { p[i] = __float2half_rn(v); }
};
template <> struct St<__hip_bfloat16> {
    __device__ __forceinline__ static float ld(const __hip_bfloat16 *p, int64_t i) { return __bfloat162float(p[i]); }
    __device__ __forceinline__ static void st(__hip_bfloat16 *p, int64_t i, float v) { p[i] = __float2bfloat16(v); }
};

// Load in the arithmetic type of the storage: double for F64, float otherwise.
template <typename S> struct LdT { using type = float; };
template <> struct LdT<double> { using type = double; };
template <typename S> __device__ __forceinline__ typename LdT<S>::type ldf(const S *p, int64_t i) {
    return St<S>::ld(p, i);
}

// Raw register form of a loaded value: 16-bit types stay as their bits until
// use, so a load issued ahead of its use is not waited for by an early
// conversion (which would serialise its latency).
template <typename D> struct RawOf { using type = D; };
template <> struct RawOf<__half> { using type = unsigned short; };
template <> struct RawOf<__hip_bfloat16> { using type = unsigned short; };
template <typename D> __device__ __forceinline__ typename RawOf<D>::type ld_raw(const D *p, int64_t i) {
    return reinterpret_cast<const typename RawOf<D>::type *>(p)[i];
}
template <typename T, typename D> __device__ __forceinline__ T from_raw(typename RawOf<D>::type r) {
    if constexpr (std::is_same<D, __half>::value) return (T)(float)__builtin_bit_cast(_Float16, r);
    else if constexpr (std::is_same<D, __hip_bfloat16>::value) return (T)__uint_as_float((unsigned)r << 16);
    else return (T)r;
}

// hardware reciprocal and square root of the fast kernels
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Detector pixels whose centres lie in a voxel's shadow [a, b] (in pixels
// from the centre of pixel 0): ceil(a - 0.01) .. floor(b + 0.01).  The chord
// of a ray vanishes at the shadow's edge, so this keeps every pixel with a
// nonzero chord and drops the empty margin of the reference's
// floor(x + 0.5) bounding box (fast kernels only).
__device__ __forceinline__ int shadow_lo(float a) { return (int)ceilf(a - 0.01f); }
__device__ __forceinline__ int shadow_hi(float b) { return (int)floorf(b + 0.01f); }

// Hardware min/max for the fast kernels.  v_min_f32/v_max_f32/v_min3_f32
// return the other operand for a quiet NaN input, which is the fminf/fmaxf
// result for every NaN that arithmetic can produce (0*inf); issuing them
// directly keeps the compiler from re-canonicalising loop-carried operands
// on every use.
__device__ __forceinline__ float hmin(float a, float b) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float hmax(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// max(a - b, 0) for a - b < 1 (the VOP3 output clamp saturates to [0, 1])
__device__ __forceinline__ float hsub_sat(float a, float b) {
    float r;
    asm("v_sub_f32_e64 %0, %1, %2 clamp" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ float hmin3(float a, float b, float c) {
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Workgroups are dealt to MI355X's 8 XCDs round-robin by linear id.  This
// remap hands each XCD one contiguous eighth of the tile range (x fastest,
// then y, then z), so the tiles that share an XCD's L2 are neighbours.
__device__ __forceinline__ int3 xcd_tile() {
    const unsigned gx = gridDim.x, gy = gridDim.y, T = gx * gy * gridDim.z;
    const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned t = (T & 7u) == 0 ? (L & 7u) * (T >> 3) + (L >> 3) : L;
    const unsigned z = t / (gx * gy), r = t - z * (gx * gy);
    const unsigned y = r / gx;
    return make_int3((int)(r - y * gx), (int)y, (int)z);
}

// Raw buffer loads over one array of `bytes` bytes (< 4 GiB) with 32-bit byte
// offsets.  The buffer unit range-checks each access: an offset at or past
// the end returns 0 instead of touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename S> __device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ float bload<__half>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __half2float(__builtin_bit_cast(__half, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)));
}
template <> __device__ __forceinline__ float bload<__hip_bfloat16>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __bfloat162float(__builtin_bit_cast(__hip_bfloat16, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)));
}

// The same with a wave-uniform part of the offset in an SGPR (soffset); the
// range check applies to voffset + soffset.
template <typename S> __device__ __forceinline__ float bload_so(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t so);
template <> __device__ __forceinline__ float bload_so<float>(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, so, 0));
}
template <> __device__ __forceinline__ float bload_so<__half>(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t so) {
    return __half2float(__builtin_bit_cast(__half, __builtin_amdgcn_raw_buffer_load_b16(r, off, so, 0)));
}
template <> __device__ __forceinline__ float bload_so<__hip_bfloat16>(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t so) {
    return __bfloat162float(__builtin_bit_cast(__hip_bfloat16, __builtin_amdgcn_raw_buffer_load_b16(r, off, so, 0)));
}

inline size_t dtype_size(ct_dtype d) {
    switch (d) {
    case CT_F32: return 4;
    case CT_F16: return 2;
    case CT_BF16: return 2;
    case CT_F64: return 8;
    }
    return 0;
}
template <typename S> constexpr ct_dtype dtype_of();
template <> constexpr ct_dtype dtype_of<float>() { return CT_F32; }
template <> constexpr ct_dtype dtype_of<__half>() { return CT_F16; }
template <> constexpr ct_dtype dtype_of<__hip_bfloat16>() { return CT_BF16; }
template <> constexpr ct_dtype dtype_of<double>() { return CT_F64; }
inline bool dtype_valid(int d) { return d == CT_F32 || d == CT_F16 || d == CT_BF16 || d == CT_F64; }

// ------------------------------------------------------------ geometry
// Derived constants, computed on the host exactly as V63/main_recon.cu:61-111
// computes its compile-time constants, in the arithmetic type T.
template <typename T> struct DGeo {
    int M, N, Z, R, ZP, NS, moff, tb, style;
    T vx, vy, vz, ivx, ivy, ivz, bx, by, bz;
    T du, dv, idu, idv, ymin, zmin, zmax, dso, dod, sx, sy, sz0, sdz;
    T us_rate, ang0, shift;
};

// Source end of the ray through the detector pixel at lateral coordinate u
// and height v, per CT_style (V63/kernel_IterativeRecon_universal.cu:22-39):
// cone beam one source (per source index for multi-source), fan beam at the
// row's height, parallel beam moving with the pixel.
template <typename T>
__host__ __device__ __forceinline__ void ray_source(const DGeo<T> &d, T s, T c, T u, T v, int src, T &x, T &y,
                                                    T &z) {
    if (d.style == CT_PARALLEL) {
        x = d.sx * c - u * s;
        y = d.sx * s + u * c;
    } else {
        x = d.sx * c - d.sy * s;
        y = d.sx * s + d.sy * c;
    }
    z = d.style == CT_CONE ? (d.tb ? (T)(d.sz0 + (T)src * d.sdz) : d.sz0) : v;
}

template <typename T> DGeo<T> derive(const ct_geometry &g) {
    DGeo<T> d;
    d.M = g.nx; d.N = g.ny; d.Z = g.nz; d.R = g.ncols; d.ZP = g.nrows;
    d.NS = g.nsources < 1 ? 1 : g.nsources; d.moff = g.x_offset; d.tb = d.NS > 1; d.style = g.style;
    d.vx = (T)g.vx; d.vy = (T)g.vy; d.vz = (T)g.vz;
    d.ivx = (T)(1.0 / (double)d.vx); d.ivy = (T)(1.0 / (double)d.vy); d.ivz = (T)(1.0 / (double)d.vz);
    d.bx = -d.vx * ((T)d.M / (T)2 + (T)d.moff);
    d.by = -d.vy * ((T)d.N / (T)2);
    d.bz = -d.vz * ((T)d.Z / (T)2);
    d.du = (T)g.du; d.dv = (T)g.dv;
    d.idu = (T)1 / d.du; d.idv = (T)1 / d.dv;
    const T off = (T)g.det_offset;
    d.ymin = -d.du * ((T)d.R / (T)2 - (T)0.5) + off;
    d.zmin = -d.dv * ((T)d.ZP / (T)2 - (T)0.5);
    d.zmax = +d.dv * ((T)d.ZP / (T)2 - (T)0.5);
    d.dso = (T)g.dso; d.dod = (T)g.dod;
    d.sx = d.dso; d.sy = off;
    d.sz0 = (T)g.src_z0; d.sdz = (T)g.src_dz;
    d.us_rate = (T)g.angle_step_deg; d.ang0 = (T)g.angle0_deg; d.shift = (T)g.angle_shift_deg;
    return d;
}

// --------------------------------------------------------------- context
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

} // namespace ct

struct ct_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    // grow-only scratch slots, named by use
    ct::DevBuf slots[16];
    // cached per-view angle table (sin, cos interleaved) keyed by geometry + precision + fdk
    ct::DevBuf angles;
    std::vector<double> angle_key;
    size_t scratch_total = 0;
    ct_math math = CT_MATH_FAST;
    // kernel timing (ct_profile_*)
    bool prof = false;
    struct ProfRec { int kid; hipEvent_t a, b; };
    std::vector<ProfRec> prof_recs;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[CT_K_COUNT] = {};
    int64_t prof_n[CT_K_COUNT] = {};
    // host copy of the angle table of the allocation-free forward projection
    std::vector<char> noalloc_host;
};

namespace ct {

// Drains recorded event pairs into the per-kernel totals (synchronizes).
void prof_drain(ct_context *ctx);

// Brackets the launches of one profiled kernel call with an event pair.
struct ProfScope {
    ct_context *ctx;
    int kid;
    hipEvent_t a = nullptr;
    ProfScope(ct_context *c, int k) : ctx(c), kid(k) {
        if (!ctx->prof) return;
        a = take();
        hipEventRecord(a, ctx->stream);
    }
    ~ProfScope() {
        if (!a) return;
        hipEvent_t b = take();
        hipEventRecord(b, ctx->stream);
        ctx->prof_recs.push_back({kid, a, b});
        if (ctx->prof_recs.size() >= 8192) prof_drain(ctx);
    }
    hipEvent_t take() {
        if (!ctx->prof_pool.empty()) {
            hipEvent_t e = ctx->prof_pool.back();
            ctx->prof_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate(&e);
        return e;
    }
};

enum Slot {
    SLOT_CORR = 0,     // SART correction images
    SLOT_RED = 1,      // reduction partials
    SLOT_FGP = 2,      // 9 FGP volumes (two dual sets, momentum copies)
    SLOT_FDK = 3,      // weighted + filtered projection stack
    SLOT_TMP_VOL = 4,  // adjoint test / power iteration volumes
    SLOT_TMP_VOL2 = 5,
    SLOT_TMP_PROJ = 6,
    SLOT_TMP_PROJ2 = 7,
    SLOT_FFT = 8,      // twiddles
    SLOT_MISC = 9,
    SLOT_BPTAB = 10,   // per-view tables of the fast voxel-driven backprojector
    SLOT_ZQVOL = 11,   // the sweep's z-quad-layout volume
    SLOT_FDKV = 12,    // per-view constants of the fast FDK backprojector
    SLOT_IO = 13,      // FP32 staging of file reads and writes
    SLOT_ZQVOL_T = 14  // the sweep's volume in the transposed z-quad layout
};

ct_status fail(ct_context *ctx, ct_status s, const std::string &msg);
ct_status hip_check(ct_context *ctx, hipError_t e, const char *what);
ct_status launch_check(ct_context *ctx, const char *what);
void *slot(ct_context *ctx, int which, size_t bytes, ct_status *st);
// Device angle table (T = float or double), sin at [2v], cos at [2v+1].
const void *angle_table(ct_context *ctx, const ct_geometry &g, bool dbl, bool fdk, ct_status *st);
// the same table in host memory (g.nviews (sin, cos) pairs, float or double)
void fill_angle_table(const ct_geometry &g, bool dbl, bool fdk, void *host);
// looks up this device's address of the allocation-free entry points' scratch
ct_status noalloc_init(ct_context *ctx);
ct_status launch_forward_noalloc(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, void *out,
                                 ct_dtype odt, int32_t view0, int32_t nv);
ct_status check_ptr(ct_context *ctx, const void *p, ct_dtype dt, const char *name);
ct_status validate_geometry(ct_context *ctx, const ct_geometry *g);
float view_cos(const ct_geometry &g, int j);   // projector angle of view j (the kernels' table)
// makes the context's device current for the calls of one entry point
struct DeviceGuard {
    explicit DeviceGuard(ct_context *ctx) { hipSetDevice(ctx->device); }
};

// launchers implemented per translation unit
ct_status launch_forward(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt,
                         void *out, ct_dtype odt, const void *meas, ct_dtype mdt,
                         int32_t view_first, int32_t view_stride, int32_t nviews, int cmd);
ct_status launch_voxel_bp(ct_context *ctx, const ct_geometry &g, void *vol, ct_dtype vdt,
                          const void *corr, int32_t view_first, int32_t view_stride,
                          int32_t nviews, double beta, int mode);
ct_status launch_adjoint(ct_context *ctx, const ct_geometry &g, const void *proj, ct_dtype pdt,
                         void *vol, ct_dtype vdt, int32_t view0, int32_t nviews);
ct_status launch_fill(ct_context *ctx, void *p, ct_dtype dt, int64_t n, double value);
ct_status launch_axpby(ct_context *ctx, const void *x, void *y, ct_dtype dt, int64_t n, double a, double b);
ct_status launch_randn(ct_context *ctx, void *p, ct_dtype dt, int64_t n, uint64_t seed);
ct_status reduce_dot(ct_context *ctx, const void *a, const void *b, int64_t n, ct_dtype dt, int mode, double *out);
ct_status reduce_tv(ct_context *ctx, const void *f, ct_dtype dt, int M, int N, int Z, int mode, double eps, double *out);
ct_status launch_convert(ct_context *ctx, const void *in, ct_dtype idt, void *out, ct_dtype odt, int64_t n);
// [z][y][x] <-> z-quad layout [z/4][y][x][z%4] (z padded to a multiple of 4
// with zeros); dtype-preserving copies
ct_status launch_zq_pack(ct_context *ctx, const void *vol, ct_dtype dt, void *volq, int M, int N, int Z);
ct_status launch_zq_unpack(ct_context *ctx, const void *volq, ct_dtype dt, void *vol, int M, int N, int Z);
// z-quad FP64 accumulator -> [z][y][x] volume of type dt
ct_status launch_zq_unpack_f64(ct_context *ctx, const double *accq, void *vol, ct_dtype dt, int M, int N, int Z);
// [z/4][y][x][z%4] <-> [z/4][x][y][z%4] (to_tr: towards the transposed one)
ct_status launch_zq_transpose(ct_context *ctx, const void *in, ct_dtype dt, void *out, int M, int N, int Z,
                              bool to_tr);
bool sweep_zq_eligible(const ct_context *ctx, const ct_geometry &g, ct_dtype vdt, int subset);
// whether the sweep runs view v's correction and update on the transposed
// z-quad layout (its rays run along y and the column walk is admitted)
bool sweep_view_transposed(const ct_geometry &g, int v);
// tr: the volume is in the transposed z-quad layout
ct_status launch_correction_zq(ct_context *ctx, const ct_geometry &g, const void *volq, ct_dtype vdt, float *corr,
                               const void *meas, ct_dtype mdt, int32_t first, int32_t stride, int32_t nv,
                               bool tr = false);
ct_status launch_voxel_bp_zq(ct_context *ctx, const ct_geometry &g, void *volq, ct_dtype vdt, const void *corr,
                             int32_t first, int32_t stride, int32_t nv, double beta, bool tr = false);
ct_status launch_forward_mv(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, void *out,
                            ct_dtype odt, const void *meas, ct_dtype mdt, int32_t first, int32_t stride,
                            int32_t nv, int cmd, int32_t mv0);
ct_status launch_residual(ct_context *ctx, const ct_geometry &g, const void *vol, ct_dtype vdt, const void *meas,
                          ct_dtype mdt, double *out);
ct_status launch_tv_eps_grad(ct_context *ctx, const void *f, void *g, ct_dtype dt, int M, int N, int Z, double eps);
ct_status launch_fista(ct_context *ctx, void *x, void *xp, void *y, int64_t n, ct_dtype dt, double gamma);
ct_status launch_fgp(ct_context *ctx, int M, int N, int Z, void *vol, ct_dtype dt, double lambda, int niter);
ct_status launch_fgp_D(ct_context *ctx, int M, int N, int Z, const void *x, void *r, void *s, void *t, ct_dtype dt);
ct_status launch_fgp_Dt(ct_context *ctx, int M, int N, int Z, const void *r, const void *s, const void *t, void *out, ct_dtype dt);
ct_status launch_fdk_weight(ct_context *ctx, const ct_geometry &g, const void *in, ct_dtype idt, void *out, ct_dtype odt);
// weighted: apply the FDK cosine weight to the input rows first
ct_status launch_ramp(ct_context *ctx, const ct_geometry &g, const void *in, ct_dtype idt, void *out, ct_dtype odt,
                      int32_t nimg, ct_window window, bool weighted = false);
// sources: how a multi-source geometry combines its sources (ct_fdk_sources)
ct_status launch_fdk_bp(ct_context *ctx, const ct_geometry &g, const void *filt, ct_dtype pdt, void *vol, ct_dtype vdt,
                        int32_t view0, int32_t nv, int sources = 0);

} // namespace ct
