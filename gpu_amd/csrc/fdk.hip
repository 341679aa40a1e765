// fdk.hip - FDK for gfx950: cosine weight, row ramp/Hann filter by an LDS
// FFT, and view-batched weighted bilinear backprojection.
//
// Ramp filter (RampFiltering_GPU, V63/FDK/FFT_1d_GPU.h:241-287): each row is
// zero-padded to pad = 2*ncols, transformed, multiplied by the real even
// filter H, transformed back and divided by pad.  H is real and even, so two
// real rows a, b packed as a + i*b come back as a' + i*b': one complex
// transform filters two rows.  One workgroup owns one row pair and keeps the
// whole padded signal in LDS (radix-2 decimation in time with precomputed
// FP64-accurate twiddles).  Non-power-of-two pads use a direct DFT.
#include "ct_internal.hpp"
#include <cmath>
#include <type_traits>

namespace ct {

template <typename S> __device__ __forceinline__ float ldx(const S *p, int64_t i) {
    if constexpr (std::is_same<S, double>::value) return (float)p[i];
    else return ldf(p, i);
}

// cosine weight, kernel_FDK_projection_adjust (V63/FDK/kernel_FDK.cu:1-14)
template <typename T, typename SI, typename SO>
__global__ void k_fdk_weight(DGeo<T> d, const SI *in, SO *out, int nviews) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y, v = blockIdx.z;
    if (x >= d.R) return;
    const T x1 = d.dso;
    const T x2 = (d.ymin + (T)x * d.du) * d.dso / (d.dso - d.dod);
    const T x3 = (d.zmin + (T)z * d.dv) * d.dso / (d.dso - d.dod);
    const int64_t idx = ((int64_t)v * d.ZP + z) * d.R + x;
    T a;
    if constexpr (std::is_same<SI, double>::value) a = (T)in[idx]; else a = (T)ldf(in, idx);
    const T r = a * d.dso / sqrt(x1 * x1 + x2 * x2 + x3 * x3);
    if constexpr (std::is_same<SO, double>::value) out[idx] = (double)r; else St<SO>::st(out, idx, (float)r);
}

// ---------------------------------------------------------------- FFT filter
struct cplx { float x, y; };

// Forward transform, decimation in frequency: natural-order input,
// bit-reversed output; tw[k] = exp(-2 pi i k / n), k < n/2.
__device__ __forceinline__ void fft_dif_lds(cplx *a, int n, int logn, const float2 *tw) {
    for (int s = logn; s >= 1; --s) {
        const int half = 1 << (s - 1);
        const int tstride = n >> s;
        for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
            const int grp = b / half, k = b - grp * half;
            const int i0 = grp * (half << 1) + k, i1 = i0 + half;
            const float2 w = tw[k * tstride];
            const cplx u = a[i0], v = a[i1];
            const float dr = u.x - v.x, di = u.y - v.y;
            a[i0].x = u.x + v.x; a[i0].y = u.y + v.y;
            a[i1].x = dr * w.x - di * w.y;
            a[i1].y = dr * w.y + di * w.x;
        }
        __syncthreads();
    }
}

// Decimation in time: bit-reversed input, natural-order output; the inverse
// transform (unnormalised) uses the conjugate twiddles.
__device__ __forceinline__ void fft_lds(cplx *a, int n, int logn, const float2 *tw, bool inverse) {
    for (int s = 1; s <= logn; ++s) {
        const int half = 1 << (s - 1);
        const int tstride = n >> s;
        for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
            const int grp = b / half, k = b - grp * half;
            const int i0 = grp * (half << 1) + k, i1 = i0 + half;
            float2 w = tw[k * tstride];
            if (inverse) w.y = -w.y;
            const cplx u = a[i0], v = a[i1];
            const float vr = v.x * w.x - v.y * w.y;
            const float vi = v.x * w.y + v.y * w.x;
            a[i0].x = u.x + vr; a[i0].y = u.y + vi;
            a[i1].x = u.x - vr; a[i1].y = u.y - vi;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int bitrev(int x, int logn) { return (int)(__brev((unsigned)x) >> (32 - logn)); }

template <typename SI, typename SO>
__global__ __launch_bounds__(256) void k_ramp_fft(const SI *in, SO *out, int R, int64_t nrows_total, int pad, int logn,
                                                  const float2 *tw, const float *H) {
    extern __shared__ cplx sig[];
    const int64_t ra = (int64_t)blockIdx.x * 2, rb = ra + 1;
    const bool hb = rb < nrows_total;
    const float inv = 1.0f / (float)pad;
    for (int x = threadIdx.x; x < pad; x += blockDim.x) {
        cplx c;
        c.x = x < R ? ldx(in, ra * R + x) : 0.0f;
        c.y = (x < R && hb) ? ldx(in, rb * R + x) : 0.0f;
        sig[x] = c;
    }
    __syncthreads();
    fft_dif_lds(sig, pad, logn, tw);
    // spectrum is in bit-reversed order: position k holds frequency bitrev(k)
    for (int k = threadIdx.x; k < pad; k += blockDim.x) {
        const float h = H[bitrev(k, logn)];
        sig[k].x *= h; sig[k].y *= h;
    }
    __syncthreads();
    fft_lds(sig, pad, logn, tw, true);
    for (int x = threadIdx.x; x < R; x += blockDim.x) {
        const cplx c = sig[x];
        if constexpr (std::is_same<SO, double>::value) {
            out[ra * R + x] = (double)(c.x * inv);
            if (hb) out[rb * R + x] = (double)(c.y * inv);
        } else {
            St<SO>::st(out, ra * R + x, c.x * inv);
            if (hb) St<SO>::st(out, rb * R + x, c.y * inv);
        }
    }
}

// Direct DFT for pads that are not powers of two; tw holds n entries.
template <typename SI, typename SO>
__global__ __launch_bounds__(256) void k_ramp_dft(const SI *in, SO *out, int R, int64_t nrows_total, int pad,
                                                  const float2 *tw, const float *H) {
    extern __shared__ cplx buf[];
    cplx *sig = buf, *spec = buf + pad;
    const int64_t ra = (int64_t)blockIdx.x * 2, rb = ra + 1;
    const bool hb = rb < nrows_total;
    for (int x = threadIdx.x; x < pad; x += blockDim.x) {
        cplx c;
        c.x = x < R ? ldx(in, ra * R + x) : 0.0f;
        c.y = (x < R && hb) ? ldx(in, rb * R + x) : 0.0f;
        sig[x] = c;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < pad; k += blockDim.x) {
        float re = 0, im = 0;
        for (int m = 0; m < pad; ++m) {
            const float2 w = tw[(int)(((int64_t)k * m) % pad)];
            re += sig[m].x * w.x - sig[m].y * w.y;
            im += sig[m].x * w.y + sig[m].y * w.x;
        }
        spec[k].x = re * H[k]; spec[k].y = im * H[k];
    }
    __syncthreads();
    const float inv = 1.0f / (float)pad;
    for (int x = threadIdx.x; x < R; x += blockDim.x) {
        float re = 0, im = 0;
        for (int k = 0; k < pad; ++k) {
            const float2 w = tw[(int)(((int64_t)k * x) % pad)];
            re += spec[k].x * w.x + spec[k].y * w.y;
            im += spec[k].y * w.x - spec[k].x * w.y;
        }
        if constexpr (std::is_same<SO, double>::value) {
            out[ra * R + x] = (double)(re * inv);
            if (hb) out[rb * R + x] = (double)(im * inv);
        } else {
            St<SO>::st(out, ra * R + x, re * inv);
            if (hb) St<SO>::st(out, rb * R + x, im * inv);
        }
    }
}

// H of V63/FDK/FFT_1d_GPU.h:243-269 (computed in float as the reference does),
// optionally Hann-windowed; twiddles in FP64 rounded to float.
static void filter_tables(const ct_geometry &g, ct_window window, bool pow2, std::vector<float2> &tw,
                          std::vector<float> &H) {
    const int pad = 2 * g.ncols;
    const float du = (float)g.du, dso = (float)g.dso, dod = (float)g.dod;
    const float fmax_ = (float)(1.0 / 2.0 / (double)(du * dso / (dso - dod)));
    const float fstep = fmax_ / (float)(pad / 2);
    H.resize(pad);
    for (int i = 0; i < pad; ++i) {
        const int k = i >= pad / 2 ? i - pad : i;
        float h = (float)std::abs(k) * fstep;
        h *= 0.5f;
        double hv = h;
        if (window == CT_WINDOW_HANN) hv *= 0.5 * (1.0 + std::cos(2.0 * M_PI * (double)k / (double)pad));
        H[i] = (float)hv;
    }
    const int nt = pow2 ? pad / 2 : pad;
    tw.resize(nt);
    for (int k = 0; k < nt; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)pad;
        tw[k].x = (float)std::cos(a);
        tw[k].y = (float)std::sin(a);
    }
}

template <typename SI, typename SO>
static ct_status run_ramp(ct_context *ctx, const ct_geometry &g, const void *in, void *out, int nimg, ct_window window) {
    const int R = g.ncols, pad = 2 * R;
    const bool pow2 = (pad & (pad - 1)) == 0;
    int logn = 0;
    while ((1 << logn) < pad) ++logn;
    const size_t lds = sizeof(cplx) * (pow2 ? pad : 2 * pad);
    if (lds > 64 * 1024) return fail(ctx, CT_ERR_INVALID_ARG, "ramp filter: padded row does not fit in LDS");
    std::vector<float2> tw;
    std::vector<float> H;
    filter_tables(g, window, pow2, tw, H);
    ct_status st;
    char *tab = (char *)slot(ctx, SLOT_FFT, sizeof(float2) * tw.size() + sizeof(float) * H.size(), &st);
    if (!tab) return st;
    float2 *dtw = (float2 *)tab;
    float *dH = (float *)(tab + sizeof(float2) * tw.size());
    st = hip_check(ctx, hipMemcpyAsync(dtw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice, ctx->stream), "twiddles");
    if (st != CT_OK) return st;
    st = hip_check(ctx, hipMemcpyAsync(dH, H.data(), sizeof(float) * H.size(), hipMemcpyHostToDevice, ctx->stream), "filter");
    if (st != CT_OK) return st;
    // pageable sources: make sure the copies have consumed the host vectors
    st = hip_check(ctx, hipStreamSynchronize(ctx->stream), "filter sync");
    if (st != CT_OK) return st;
    const int64_t rows = (int64_t)nimg * g.nrows;
    const int64_t pairs = (rows + 1) / 2;
    if (pairs > 2147483647LL) return fail(ctx, CT_ERR_OVERFLOW, "ramp filter: too many rows");
    ProfScope ps(ctx, CT_K_RAMP);
    if (pow2)
        hipLaunchKernelGGL((k_ramp_fft<SI, SO>), dim3((unsigned)pairs), dim3(256), lds, ctx->stream, (const SI *)in,
                           (SO *)out, R, rows, pad, logn, (const float2 *)dtw, (const float *)dH);
    else
        hipLaunchKernelGGL((k_ramp_dft<SI, SO>), dim3((unsigned)pairs), dim3(256), lds, ctx->stream, (const SI *)in,
                           (SO *)out, R, rows, pad, (const float2 *)dtw, (const float *)dH);
    return launch_check(ctx, "ramp filter");
}

template <typename SI>
static ct_status ramp_out(ct_context *ctx, const ct_geometry &g, const void *in, void *out, ct_dtype odt, int nimg,
                          ct_window w) {
    switch (odt) {
    case CT_F32: return run_ramp<SI, float>(ctx, g, in, out, nimg, w);
    case CT_F16: return run_ramp<SI, __half>(ctx, g, in, out, nimg, w);
    case CT_BF16: return run_ramp<SI, __hip_bfloat16>(ctx, g, in, out, nimg, w);
    case CT_F64: return run_ramp<SI, double>(ctx, g, in, out, nimg, w);
    }
    return fail(ctx, CT_ERR_DTYPE, "ramp: bad output dtype");
}

ct_status launch_ramp(ct_context *ctx, const ct_geometry &g, const void *in, ct_dtype idt, void *out, ct_dtype odt,
                      int32_t nimg, ct_window window) {
    if (nimg <= 0) return CT_OK;
    switch (idt) {
    case CT_F32: return ramp_out<float>(ctx, g, in, out, odt, nimg, window);
    case CT_F16: return ramp_out<__half>(ctx, g, in, out, odt, nimg, window);
    case CT_BF16: return ramp_out<__hip_bfloat16>(ctx, g, in, out, odt, nimg, window);
    case CT_F64: return ramp_out<double>(ctx, g, in, out, odt, nimg, window);
    }
    return fail(ctx, CT_ERR_DTYPE, "ramp: bad input dtype");
}

template <typename SI, typename SO>
static ct_status run_weight(ct_context *ctx, const ct_geometry &g, const void *in, void *out) {
    dim3 b(256), gr((g.ncols + 255) / 256, g.nrows, g.nviews);
    if (std::is_same<SI, double>::value && std::is_same<SO, double>::value)
        hipLaunchKernelGGL((k_fdk_weight<double, SI, SO>), gr, b, 0, ctx->stream, derive<double>(g), (const SI *)in, (SO *)out, g.nviews);
    else
        hipLaunchKernelGGL((k_fdk_weight<float, SI, SO>), gr, b, 0, ctx->stream, derive<float>(g), (const SI *)in, (SO *)out, g.nviews);
    return launch_check(ctx, "fdk weight");
}
template <typename SI>
static ct_status weight_out(ct_context *ctx, const ct_geometry &g, const void *in, void *out, ct_dtype odt) {
    switch (odt) {
    case CT_F32: return run_weight<SI, float>(ctx, g, in, out);
    case CT_F16: return run_weight<SI, __half>(ctx, g, in, out);
    case CT_BF16: return run_weight<SI, __hip_bfloat16>(ctx, g, in, out);
    case CT_F64: return run_weight<SI, double>(ctx, g, in, out);
    }
    return fail(ctx, CT_ERR_DTYPE, "fdk weight: bad output dtype");
}
ct_status launch_fdk_weight(ct_context *ctx, const ct_geometry &g, const void *in, ct_dtype idt, void *out, ct_dtype odt) {
    switch (idt) {
    case CT_F32: return weight_out<float>(ctx, g, in, out, odt);
    case CT_F16: return weight_out<__half>(ctx, g, in, out, odt);
    case CT_BF16: return weight_out<__hip_bfloat16>(ctx, g, in, out, odt);
    case CT_F64: return weight_out<double>(ctx, g, in, out, odt);
    }
    return fail(ctx, CT_ERR_DTYPE, "fdk weight: bad input dtype");
}

// ------------------------------------------------------- backprojection
// kernel_FDK_reconstruction (V63/FDK/kernel_FDK.cu:16-62) over a batch of views:
// one voxel per lane accumulates all views in a register and writes once.
// The reference's per-view weight t1*t2*dtheta*360/(us_rate*Nviews) is applied
// with the FP32 factor dtheta*360/(us_rate*Nviews).
template <typename T, typename SP, typename SV>
__global__ __launch_bounds__(256) void k_fdk_bp(DGeo<T> d, const T *__restrict__ ang, const SP *__restrict__ filt,
                                                SV *__restrict__ vol, int view0, int nv, T wscale) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y * 4 + threadIdx.y;
    const int iz = blockIdx.z;
    if (ix >= d.M || iy >= d.N) return;
    const T x = d.bx + d.vx * (T)0.5 + (T)ix * d.vx;
    const T y = d.by + d.vy * (T)0.5 + (T)iy * d.vy;
    const T z = d.bz + d.vz * (T)0.5 + (T)iz * d.vz;
    const int64_t P = (int64_t)d.R * d.ZP;
    const int64_t idx = ((int64_t)iz * d.N + iy) * d.M + ix;
    T acc;
    if constexpr (std::is_same<SV, double>::value) acc = vol[idx]; else acc = (T)ldf(vol, idx);
    const T sdd = d.dso - d.dod;
    for (int q = 0; q < nv; ++q) {
        const int v = view0 + q;
        const T s = ang[2 * v], c = ang[2 * v + 1];
        const T t = x * c + y * s;
        const T sv = -x * s + y * c;
        const T t1 = d.dso * d.dso / ((d.dso - sv) * (d.dso - sv));
        T ksi = d.dso * z / (d.dso - sv);
        ksi = (ksi * sdd / d.dso - d.zmin) / d.dv;
        T p = d.dso * t / (d.dso - sv);
        p = (p * sdd / d.dso - d.ymin) / d.du;
        T t2 = 0;
        if (((T)0 <= ksi) && (ksi + 1 < (T)d.ZP) && ((T)0 <= p) && (p + 1 < (T)d.R)) {
            const int an = (int)floor(ksi);
            const T an0 = ksi - (T)an;
            const int bn = (int)floor(p);
            const T bn0 = p - (T)bn;
            const SP *pr = filt + (int64_t)q * P + (int64_t)an * d.R + bn;
            T p00, p01, p10, p11;
            if constexpr (std::is_same<SP, double>::value) { p00 = pr[0]; p01 = pr[1]; p10 = pr[d.R]; p11 = pr[d.R + 1]; }
            else { p00 = ldf(pr, 0); p01 = ldf(pr, 1); p10 = ldf(pr, d.R); p11 = ldf(pr, d.R + 1); }
            t2 = ((T)1 - an0) * ((T)1 - bn0) * p00 + ((T)1 - an0) * bn0 * p01 + an0 * ((T)1 - bn0) * p10 +
                 an0 * bn0 * p11;
        }
        acc += t1 * t2 * wscale;
    }
    if constexpr (std::is_same<SV, double>::value) vol[idx] = acc; else St<SV>::st(vol, idx, (float)acc);
}

template <typename T, typename SP, typename SV>
static ct_status run_fdk_bp(ct_context *ctx, const ct_geometry &g, const void *filt, void *vol, int view0, int nv) {
    ct_status st;
    const T *ang = (const T *)angle_table(ctx, g, sizeof(T) == 8, true, &st);
    if (!ang) return st;
    DGeo<T> d = derive<T>(g);
    const T pi = sizeof(T) == 8 ? (T)3.141592653589793 : (T)3.141592653589793f;
    const T dth = d.us_rate * pi / (T)180;
    const double w = (double)dth * 360.0 / (double)(d.us_rate * (T)g.nviews);
    dim3 block(64, 4, 1), grid((d.M + 63) / 64, (d.N + 3) / 4, d.Z);
    ProfScope ps(ctx, CT_K_FDK_BP);
    hipLaunchKernelGGL((k_fdk_bp<T, SP, SV>), grid, block, 0, ctx->stream, d, ang, (const SP *)filt, (SV *)vol, view0,
                       nv, (T)w);
    return launch_check(ctx, "fdk backprojection");
}

template <typename SP>
static ct_status fdk_bp_vol(ct_context *ctx, const ct_geometry &g, const void *filt, void *vol, ct_dtype vdt, int v0, int nv) {
    switch (vdt) {
    case CT_F32: return run_fdk_bp<float, SP, float>(ctx, g, filt, vol, v0, nv);
    case CT_F16: return run_fdk_bp<float, SP, __half>(ctx, g, filt, vol, v0, nv);
    case CT_BF16: return run_fdk_bp<float, SP, __hip_bfloat16>(ctx, g, filt, vol, v0, nv);
    case CT_F64: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "fdk backprojection: F64 volumes need F64 projections");
}

ct_status launch_fdk_bp(ct_context *ctx, const ct_geometry &g, const void *filt, ct_dtype pdt, void *vol, ct_dtype vdt,
                        int32_t view0, int32_t nv) {
    if (nv <= 0) return CT_OK;
    if (pdt == CT_F64 || vdt == CT_F64) {
        if (pdt != CT_F64 || vdt != CT_F64) return fail(ctx, CT_ERR_DTYPE, "fdk backprojection: mixed F64 dtypes");
        return run_fdk_bp<double, double, double>(ctx, g, filt, vol, view0, nv);
    }
    switch (pdt) {
    case CT_F32: return fdk_bp_vol<float>(ctx, g, filt, vol, vdt, view0, nv);
    case CT_F16: return fdk_bp_vol<__half>(ctx, g, filt, vol, vdt, view0, nv);
    case CT_BF16: return fdk_bp_vol<__hip_bfloat16>(ctx, g, filt, vol, vdt, view0, nv);
    default: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "fdk backprojection: bad projection dtype");
}

} // namespace ct
