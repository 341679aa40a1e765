// fdk.hip - FDK for gfx950: cosine weight, row ramp/Hann filter by an LDS
// FFT, and view-batched weighted bilinear backprojection.
//
// Ramp filter (RampFiltering_GPU, V63/FDK/FFT_1d_GPU.h:241-287): each row is
// zero-padded to pad = 2*ncols, transformed, multiplied by the real even
// filter H, transformed back and divided by pad.  H is real and even, so two
// real rows a, b packed as a + i*b come back as a' + i*b': one complex
// transform filters two rows.  One workgroup owns one row pair.  Pads of
// 2^9..2^12 run Stockham radix-8 passes in registers with LDS exchanges;
// other powers of two keep the whole signal in LDS for radix-2 stages; both
// use precomputed FP64-accurate twiddles.  Non-power-of-two pads use a
// direct DFT.
#include "ct_internal.hpp"
#include <cmath>
#include <type_traits>

namespace ct {

template <typename S> __device__ __forceinline__ float ldx(const S *p, int64_t i) {
    if constexpr (std::is_same<S, double>::value) return (float)p[i];
    else return ldf(p, i);
}

// cosine weight, kernel_FDK_projection_adjust (V63/FDK/kernel_FDK.cu:1-14);
// multi-source images [view][source] measure the row height from their
// source (TB/FDK/kernel_FDK.cu:1-17)
template <typename T> __device__ __forceinline__ T fdk_row_height(const DGeo<T> &d, int z, int img) {
    return d.tb ? d.zmin + (T)z * d.dv - (T)(d.sz0 + (T)(img % d.NS) * d.sdz) : d.zmin + (T)z * d.dv;
}

template <typename T, typename SI, typename SO>
__global__ void k_fdk_weight(DGeo<T> d, const SI *in, SO *out, int nviews) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y, v = blockIdx.z;
    if (x >= d.R) return;
    const T x1 = d.dso;
    const T x2 = (d.ymin + (T)x * d.du) * d.dso / (d.dso - d.dod);
    const T x3 = fdk_row_height(d, z, v) * d.dso / (d.dso - d.dod);
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

// ------------------------------------------------ Stockham radix-8 filter
// For pads of 2^9..2^12 the row-pair transform runs as Stockham passes
// (Govindaraju et al., SC'08): a thread holds 8 complex values in registers
// and does one radix-8 butterfly per pass (the last pass, radix 2^(LOGN-9),
// does 8/R smaller ones), so a transform is three or four passes with an LDS
// exchange between them instead of LOGN radix-2 stages.  Pass p with radix R
// and sub-transform size Ns reads x[j + r*N/R], multiplies by
// w^(r*(j mod Ns)*N/(Ns*R)), and writes x'[(j - j mod Ns)*R + j mod Ns + r*Ns];
// input and output are in natural order.  The last forward pass leaves every
// thread holding the spectrum at tid + r*N/8 up to a register permutation,
// which is exactly the first inverse pass's input, so the filter multiply
// needs no exchange; the inverse output is written straight to global memory.
// With `weighted` the FDK cosine weight (V63/FDK/kernel_FDK.cu:8-13, the same
// FP32 operations as k_fdk_weight) is applied as the rows are read.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); } // LDS bank spreading

template <bool INV> __device__ __forceinline__ float2 twmul(float2 a, float2 w) {
    // a * w forward, a * conj(w) inverse
    return INV ? make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y)
               : make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (-i) forward, a * (+i) inverse
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
template <bool INV> __device__ __forceinline__ void dft2(float2 *v) {
    const float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}
template <bool INV> __device__ __forceinline__ void dft4(float2 *v) {
    const float2 c0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    const float2 c1 = cadd(v[1], v[3]), d1 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(c0, c1); v[2] = csub(c0, c1);
    v[1] = cadd(d0, d1); v[3] = csub(d0, d1);
}
// X_2m = DFT4(x_n + x_n+4)_m, X_2m+1 = DFT4((x_n - x_n+4) w8^n)_m
template <bool INV> __device__ __forceinline__ void dft8(float2 *v) {
    constexpr float h = 0.70710678118654752f;
    float2 a[4], b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) { a[n] = cadd(v[n], v[n + 4]); b[n] = csub(v[n], v[n + 4]); }
    // b1 *= w8, b2 *= w8^2 = -i, b3 *= w8^3 (conjugates for the inverse)
    b[1] = INV ? make_float2((b[1].x - b[1].y) * h, (b[1].x + b[1].y) * h)
               : make_float2((b[1].x + b[1].y) * h, (b[1].y - b[1].x) * h);
    b[2] = mul_mi<INV>(b[2]);
    b[3] = INV ? make_float2(-(b[3].x + b[3].y) * h, (b[3].x - b[3].y) * h)
               : make_float2((b[3].y - b[3].x) * h, -(b[3].x + b[3].y) * h);
    dft4<INV>(a);
    dft4<INV>(b);
#pragma unroll
    for (int m = 0; m < 4; ++m) { v[2 * m] = a[m]; v[2 * m + 1] = b[m]; }
}
template <int R, bool INV> __device__ __forceinline__ void dftr(float2 *v) {
    if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else dft2<INV>(v);
}

// one pass: twiddle and butterflies for the thread's 8/R butterflies
// (butterfly b is j = tid + b*N/8); lns = log2(Ns)
template <int LOGN, int R, bool INV>
__device__ __forceinline__ void sh_pass(float2 (&v)[8], int tid, int lns, const float2 *tw) {
    constexpr int N = 1 << LOGN, T = N / 8, NB = 8 / R, LR = R == 8 ? 3 : (R == 4 ? 2 : 1);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (lns > 0) {
            const int j = tid + b * T;
            const int k = j & ((1 << lns) - 1);
            const int step = k << (LOGN - LR - lns); // k*N/(Ns*R)
#pragma unroll
            for (int r = 1; r < R; ++r) v[b * R + r] = twmul<INV>(v[b * R + r], tw[step * r]);
        }
        dftr<R, INV>(&v[b * R]);
    }
}
// write the pass outputs (radix R, sub-transform size 2^lns) to LDS, then
// read the next pass's inputs (radix RN)
template <int LOGN, int R, int RN>
__device__ __forceinline__ void sh_exchange(float2 (&v)[8], float2 *buf, int tid, int lns) {
    constexpr int N = 1 << LOGN, T = N / 8;
    __syncthreads(); // the previous pass's reads of buf are done
#pragma unroll
    for (int b = 0; b < 8 / R; ++b) {
        const int j = tid + b * T;
        const int k = j & ((1 << lns) - 1);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[lpad(base + (r << lns))] = v[b * R + r];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 8 / RN; ++b) {
        const int j = tid + b * T;
#pragma unroll
        for (int r = 0; r < RN; ++r) v[b * RN + r] = buf[lpad(j + r * (N / RN))];
    }
}
// full transform; v holds x[tid + r*N/8] on entry and X[tid + perm(r)*N/8]
// on exit, with perm(b*RL + r) = b + r*(8/RL)
template <int LOGN, bool INV>
__device__ __forceinline__ void sh_fft(float2 (&v)[8], float2 *buf, int tid, const float2 *tw) {
    sh_pass<LOGN, 8, INV>(v, tid, 0, tw);
    sh_exchange<LOGN, 8, 8>(v, buf, tid, 0);
    sh_pass<LOGN, 8, INV>(v, tid, 3, tw);
    sh_exchange<LOGN, 8, 8>(v, buf, tid, 3);
    sh_pass<LOGN, 8, INV>(v, tid, 6, tw);
    if constexpr (LOGN > 9) {
        constexpr int RL = 1 << (LOGN - 9);
        sh_exchange<LOGN, 8, RL>(v, buf, tid, 6);
        sh_pass<LOGN, RL, INV>(v, tid, 9, tw);
    }
}
template <int LOGN> __device__ __forceinline__ constexpr int sh_perm(int i) {
    constexpr int RL = LOGN > 9 ? 1 << (LOGN - 9) : 8;
    return (i / RL) + (i % RL) * (8 / RL);
}

template <typename SI>
__device__ __forceinline__ float fdk_weighted(const DGeo<float> &d, const SI *in, int64_t idx, int x, int z, int img) {
    const float x1 = d.dso;
    const float x2 = (d.ymin + (float)x * d.du) * d.dso / (d.dso - d.dod);
    const float x3 = fdk_row_height(d, z, img) * d.dso / (d.dso - d.dod);
    const float a = ldx(in, idx);
    return a * d.dso / sqrtf(x1 * x1 + x2 * x2 + x3 * x3);
}

template <int LOGN, typename SI, typename SO, bool WEIGHT>
__global__ __launch_bounds__(1 << (LOGN - 3)) void k_ramp_sh(const SI *__restrict__ in, SO *__restrict__ out, int R,
                                                            int64_t nrows_total, const float2 *__restrict__ tw,
                                                            const float *__restrict__ H, DGeo<float> d) {
    constexpr int N = 1 << LOGN, T = N / 8;
    __shared__ float2 buf[N + N / 16];
    __shared__ float2 stw[N];
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < 8; ++r) stw[tid + r * T] = tw[tid + r * T];
    const int64_t ra = (int64_t)blockIdx.x * 2, rb = ra + 1;
    const bool hb = rb < nrows_total;
    const int za = (int)(ra % d.ZP), zb = (int)(rb % d.ZP);
    const int ia = (int)(ra / d.ZP), ib = (int)(rb / d.ZP); // image (view, source) of each row
    float2 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int n = tid + r * T;
        float a = 0.0f, b = 0.0f;
        if (n < R) {
            if constexpr (WEIGHT) {
                a = fdk_weighted(d, in, ra * R + n, n, za, ia);
                if (hb) b = fdk_weighted(d, in, rb * R + n, n, zb, ib);
            } else {
                a = ldx(in, ra * R + n);
                if (hb) b = ldx(in, rb * R + n);
            }
        }
        v[r] = make_float2(a, b);
    }
    __syncthreads();
    sh_fft<LOGN, false>(v, buf, tid, stw);
    // H is real and even: filtering a + ib filters both rows
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float h = H[tid + sh_perm<LOGN>(i) * T];
        v[i].x *= h; v[i].y *= h;
    }
    float2 u[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) u[sh_perm<LOGN>(i)] = v[i];
    sh_fft<LOGN, true>(u, buf, tid, stw);
    const float inv = 1.0f / (float)N;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int n = tid + sh_perm<LOGN>(i) * T;
        if (n < R) {
            if constexpr (std::is_same<SO, double>::value) {
                out[ra * R + n] = (double)(u[i].x * inv);
                if (hb) out[rb * R + n] = (double)(u[i].y * inv);
            } else {
                St<SO>::st(out, ra * R + n, u[i].x * inv);
                if (hb) St<SO>::st(out, rb * R + n, u[i].y * inv);
            }
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
static void filter_tables(const ct_geometry &g, ct_window window, std::vector<float2> &tw,
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
    tw.resize(pad);
    for (int k = 0; k < pad; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)pad;
        tw[k].x = (float)std::cos(a);
        tw[k].y = (float)std::sin(a);
    }
}

template <typename SI, typename SO>
static ct_status run_weight(ct_context *ctx, const ct_geometry &g, const void *in, void *out);

template <int LOGN, typename SI, typename SO>
static void launch_sh(ct_context *ctx, const ct_geometry &g, const void *in, void *out, int64_t rows, int64_t pairs,
                      const float2 *tw, const float *H, bool weighted) {
    const DGeo<float> d = derive<float>(g);
    if (weighted)
        hipLaunchKernelGGL((k_ramp_sh<LOGN, SI, SO, true>), dim3((unsigned)pairs), dim3(1 << (LOGN - 3)), 0,
                           ctx->stream, (const SI *)in, (SO *)out, g.ncols, rows, tw, H, d);
    else
        hipLaunchKernelGGL((k_ramp_sh<LOGN, SI, SO, false>), dim3((unsigned)pairs), dim3(1 << (LOGN - 3)), 0,
                           ctx->stream, (const SI *)in, (SO *)out, g.ncols, rows, tw, H, d);
}

template <typename SI, typename SO>
static ct_status run_ramp(ct_context *ctx, const ct_geometry &g, const void *in, void *out, int nimg, ct_window window,
                          bool weighted) {
    const int R = g.ncols, pad = 2 * R;
    const bool pow2 = (pad & (pad - 1)) == 0;
    int logn = 0;
    while ((1 << logn) < pad) ++logn;
    const bool sh = pow2 && logn >= 9 && logn <= 12;
    const size_t lds = sizeof(cplx) * (pow2 ? pad : 2 * pad);
    if (!sh && lds > 64 * 1024) return fail(ctx, CT_ERR_INVALID_ARG, "ramp filter: padded row does not fit in LDS");
    if (weighted && (!sh || std::is_same<SI, double>::value || std::is_same<SO, double>::value)) {
        // cosine weight as its own pass, then filter in place
        const ct_status st = run_weight<SI, SO>(ctx, g, in, out);
        if (st != CT_OK) return st;
        return run_ramp<SO, SO>(ctx, g, out, out, nimg, window, false);
    }
    std::vector<float2> tw;
    std::vector<float> H;
    filter_tables(g, window, tw, H);
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
    if (sh) {
        switch (logn) {
        case 9: launch_sh<9, SI, SO>(ctx, g, in, out, rows, pairs, dtw, dH, weighted); break;
        case 10: launch_sh<10, SI, SO>(ctx, g, in, out, rows, pairs, dtw, dH, weighted); break;
        case 11: launch_sh<11, SI, SO>(ctx, g, in, out, rows, pairs, dtw, dH, weighted); break;
        default: launch_sh<12, SI, SO>(ctx, g, in, out, rows, pairs, dtw, dH, weighted); break;
        }
    } else if (pow2)
        hipLaunchKernelGGL((k_ramp_fft<SI, SO>), dim3((unsigned)pairs), dim3(256), lds, ctx->stream, (const SI *)in,
                           (SO *)out, R, rows, pad, logn, (const float2 *)dtw, (const float *)dH);
    else
        hipLaunchKernelGGL((k_ramp_dft<SI, SO>), dim3((unsigned)pairs), dim3(256), lds, ctx->stream, (const SI *)in,
                           (SO *)out, R, rows, pad, (const float2 *)dtw, (const float *)dH);
    return launch_check(ctx, "ramp filter");
}

template <typename SI>
static ct_status ramp_out(ct_context *ctx, const ct_geometry &g, const void *in, void *out, ct_dtype odt, int nimg,
                          ct_window w, bool weighted) {
    switch (odt) {
    case CT_F32: return run_ramp<SI, float>(ctx, g, in, out, nimg, w, weighted);
    case CT_F16: return run_ramp<SI, __half>(ctx, g, in, out, nimg, w, weighted);
    case CT_BF16: return run_ramp<SI, __hip_bfloat16>(ctx, g, in, out, nimg, w, weighted);
    case CT_F64: return run_ramp<SI, double>(ctx, g, in, out, nimg, w, weighted);
    }
    return fail(ctx, CT_ERR_DTYPE, "ramp: bad output dtype");
}

ct_status launch_ramp(ct_context *ctx, const ct_geometry &g, const void *in, ct_dtype idt, void *out, ct_dtype odt,
                      int32_t nimg, ct_window window, bool weighted) {
    if (nimg <= 0) return CT_OK;
    switch (idt) {
    case CT_F32: return ramp_out<float>(ctx, g, in, out, odt, nimg, window, weighted);
    case CT_F16: return ramp_out<__half>(ctx, g, in, out, odt, nimg, window, weighted);
    case CT_BF16: return ramp_out<__hip_bfloat16>(ctx, g, in, out, odt, nimg, window, weighted);
    case CT_F64: return ramp_out<double>(ctx, g, in, out, odt, nimg, window, weighted);
    }
    return fail(ctx, CT_ERR_DTYPE, "ramp: bad input dtype");
}

template <typename SI, typename SO>
static ct_status run_weight(ct_context *ctx, const ct_geometry &g, const void *in, void *out) {
    dim3 b(256), gr((g.ncols + 255) / 256, g.nrows, g.nviews * (g.nsources < 1 ? 1 : g.nsources));
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
// The reference's per-view term temp1*temp2*dtheta*360.0/(us_rate*Nviews) is
// formed as it is there: the FP32 product temp1*temp2*dtheta, then the
// double literal 360.0 makes the rest of the expression and the += double,
// and the sum is stored back as float (V63/FDK/kernel_FDK.cu:60).
template <typename T, typename SP, typename SV>
__global__ __launch_bounds__(256) void k_fdk_bp(DGeo<T> d, const T *__restrict__ ang, const SP *__restrict__ filt,
                                                SV *__restrict__ vol, int view0, int nv, T dth, T usn) {
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
        // temp1*temp2 * (us_rate*PI/180) * 360.0/(us_rate*Nviews), V63/FDK/kernel_FDK.cu:60
        if constexpr (std::is_same<T, float>::value)
            acc = (float)((double)acc + (double)(t1 * t2 * dth) * 360.0 / (double)usn);
        else
            acc = acc + t1 * t2 * dth * (T)360 / usn;
    }
    if constexpr (std::is_same<SV, double>::value) vol[idx] = acc; else St<SV>::st(vol, idx, (float)acc);
}

// Multi-source FDK backprojection (TB/FDK/*/kernel_FDK.cu), one voxel per
// lane, the reference's operation order in both arithmetic modes: per view
// the sources combine as `mode` says - 0: sources whose stencil is on the
// detector weighted by 1/(z - z_s)^2 and normalised (TB/FDK/kernel_FDK.cu:37-99);
// 1: the source nearest in z among those whose row is on the detector
// (TB/FDK/OneRay/kernel_FDK.cu:37-113); 2: their sum
// (TB/FDK/EqualWeighting/kernel_FDK.cu:37-91, without that kernel's
// unsynchronised per-source updates) - and the view's value is added to the
// volume (kernel_volume_divide, TB/FDK/kernel_FDK.cu:21-33).
template <typename T, typename SP, typename SV>
__global__ __launch_bounds__(256) void k_fdk_bp_ms(DGeo<T> d, const T *__restrict__ ang, const SP *__restrict__ filt,
                                                   SV *__restrict__ vol, int view0, int nv, T dth, T usn, int mode) {
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
    for (int q = 0; q < nv; ++q) {
        const int v = view0 + q;
        const T s = ang[2 * v], c = ang[2 * v + 1];
        const T t = x * c + y * s;
        const T sv = -x * s + y * c;
        const T t1 = d.dso * d.dso / ((d.dso - sv) * (d.dso - sv));
        // ksi of source i (the stencil row), p (column)
        auto ksi_of = [&](T sz) {
            const T k = d.dso * (z - sz) / (d.dso - sv);
            return (k * (d.dso - d.dod) / d.dso + sz - d.zmin) / d.dv;
        };
        T p = d.dso * t / (d.dso - sv);
        p = (p * (d.dso - d.dod) / d.dso - d.ymin) / d.du;
        const bool pin = ((T)0 <= p) && (p + (T)1 < (T)d.R);
        auto sample = [&](int i, T ksi) {
            const int an = (int)floor(ksi), bn = (int)floor(p);
            const T an0 = ksi - (T)an, bn0 = p - (T)bn;
            const SP *pr = filt + ((int64_t)q * d.NS + i) * P + (int64_t)an * d.R + bn;
            T p00, p01, p10, p11;
            if constexpr (std::is_same<SP, double>::value) { p00 = pr[0]; p01 = pr[1]; p10 = pr[d.R]; p11 = pr[d.R + 1]; }
            else { p00 = ldf(pr, 0); p01 = ldf(pr, 1); p10 = ldf(pr, d.R); p11 = ldf(pr, d.R + 1); }
            return ((T)1 - an0) * ((T)1 - bn0) * p00 + ((T)1 - an0) * bn0 * p01 + an0 * ((T)1 - bn0) * p10 +
                   an0 * bn0 * p11;
        };
        T upd = 0;
        if (mode == 1) {
            int sel = -1;
            T zsel = 0, ksel = 0;
            for (int i = 0; i < d.NS; ++i) {
                const T sz = (T)(d.sz0 + (T)i * d.sdz);
                const T ksi = ksi_of(sz);
                if (((T)0 <= ksi) && (ksi + (T)1 < (T)d.ZP) && (sel < 0 || fabs(sz - z) < fabs(zsel - z))) {
                    sel = i; zsel = sz; ksel = ksi;
                }
            }
            if (sel >= 0 && pin) upd = t1 * sample(sel, ksel) * dth * (T)360 / usn;
        } else if (pin) {
            T wsum = 0;
            for (int i = 0; i < d.NS; ++i) {
                const T sz = (T)(d.sz0 + (T)i * d.sdz);
                const T ksi = ksi_of(sz);
                if (!(((T)0 <= ksi) && (ksi + (T)1 < (T)d.ZP))) continue;
                const T term = t1 * sample(i, ksi) * dth * (T)360 / usn;
                if (mode == 0) {
                    const T dz = fabs(z - sz);
                    const T w = (T)1 / dz / dz;
                    wsum += w;
                    upd += term * w;
                } else {
                    upd += term;
                }
            }
            if (mode == 0 && wsum > (T)0) upd /= wsum;
        }
        acc = acc + upd;
    }
    if constexpr (std::is_same<SV, double>::value) vol[idx] = acc; else St<SV>::st(vol, idx, (float)acc);
}

// Fast FP32 form (CT_MATH_FAST) of the same projection.  A lane walks a
// column of voxels along z: for a fixed (x, y) and view, s, the magnification
// weight (DSO/(DSO-s))^2 and the detector column p are shared by the whole
// column and ksi is linear in z, so a voxel costs one FMA for ksi plus the
// bilinear gather.  The reference's five IEEE divisions per voxel and view
// (V63/FDK/kernel_FDK.cu:36-42) become one hardware reciprocal of (DSO - s)
// per column and view: ksi = (z*(DSO-DOD)/(DSO-s) - zmin)/dv and
// p = (t*(DSO-DOD)/(DSO-s) - ymin)/du with the DSO factors cancelled.  The
// in-detector test is the reference's (the whole 2x2 stencil inside).
// Per-view constants come from a table: cos and sin times (DSO-DOD)/du, then
// sin and cos.
__global__ void k_fdk_views(const float *__restrict__ ang, float4 *__restrict__ tab, int view0, int nv, float ku) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nv) return;
    const float s = ang[2 * (view0 + q)], c = ang[2 * (view0 + q) + 1];
    tab[q] = make_float4(c * ku, s * ku, s, c);
}

struct FdkCol {
    float inv, wv, fb, bn0, k0, dk;
};
// Column constants for one view; wv is 0 for a column off the detector.
__device__ __forceinline__ FdkCol fdk_column(const DGeo<float> &d, float4 vt, float x, float y, float wk, float pm,
                                             float kz0, float kdz, float km) {
    FdkCol r;
    r.inv = frcp(fmaf(x, vt.z, fmaf(-y, vt.w, d.dso))); // 1/(DSO - s)
    const float p = fmaf(fmaf(x, vt.x, y * vt.y), r.inv, pm);
    r.wv = ((0.0f <= p) && (p + 1.0f < (float)d.R)) ? r.inv * r.inv * wk : 0.0f;
    r.fb = floorf(p);
    r.bn0 = p - r.fb;
    r.k0 = fmaf(r.inv, kz0, km); // ksi = k0 + k*dk
    r.dk = r.inv * kdz;
    return r;
}

// Global-memory form for any geometry and view dtype: a lane owns FZC voxels
// of a column; gathers are range-checked buffer loads with 32-bit offsets
// into one view, indices clamped into the view and off-detector terms masked.
constexpr int FZC = 8;
template <typename SP, typename SV>
__global__ __launch_bounds__(256) void k_fdk_bp_col(DGeo<float> d, const float4 *__restrict__ vtab,
                                                    const SP *__restrict__ filt, SV *__restrict__ vol, int nv,
                                                    float wscale) {
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int iy = blockIdx.y * 4 + threadIdx.y;
    const int iz0 = blockIdx.z * FZC;
    if (ix >= d.M || iy >= d.N) return;
    const int nz = min(FZC, d.Z - iz0);
    const float x = d.bx + d.vx * 0.5f + (float)ix * d.vx;
    const float y = d.by + d.vy * 0.5f + (float)iy * d.vy;
    const float z0 = d.bz + d.vz * 0.5f + (float)iz0 * d.vz;
    const int64_t MN = (int64_t)d.M * d.N;
    const int64_t idx0 = ((int64_t)iz0 * d.N + iy) * d.M + ix;
    float acc[FZC];
#pragma unroll
    for (int k = 0; k < FZC; ++k) acc[k] = k < nz ? ldf(vol, idx0 + k * MN) : 0.0f;
    const float sdd = d.dso - d.dod;
    const int64_t P = (int64_t)d.R * d.ZP;
    const uint32_t vbytes = (uint32_t)(P * (int64_t)sizeof(SP));
    const uint32_t rowb = (uint32_t)d.R * (uint32_t)sizeof(SP);
    const float wk = d.dso * d.dso * wscale, pm = -d.ymin * d.idu;
    const float kz0 = z0 * sdd * d.idv, kdz = d.vz * sdd * d.idv, km = -d.zmin * d.idv;
    const float zhi = (float)(d.ZP - 2), chi = (float)(d.R - 2);
    for (int q = 0; q < nv; ++q) {
        const FdkCol cc = fdk_column(d, vtab[q], x, y, wk, pm, kz0, kdz, km);
        const uint32_t cb = (uint32_t)(int)__builtin_amdgcn_fmed3f(cc.fb, 0.0f, chi) * (uint32_t)sizeof(SP);
        const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(filt + (int64_t)q * P, vbytes);
#pragma unroll
        for (int k = 0; k < FZC; ++k) {
            const float ksi = fmaf((float)k, cc.dk, cc.k0);
            const float fa = floorf(ksi), an0 = ksi - fa;
            const float fc = __builtin_amdgcn_fmed3f(fa, 0.0f, zhi);
            const uint32_t off = (uint32_t)(int)fc * rowb + cb;
            const float p00 = bload<SP>(rs, off), p01 = bload<SP>(rs, off + (uint32_t)sizeof(SP));
            const float p10 = bload<SP>(rs, off + rowb), p11 = bload<SP>(rs, off + rowb + (uint32_t)sizeof(SP));
            const float a0 = fmaf(cc.bn0, p01 - p00, p00), a1 = fmaf(cc.bn0, p11 - p10, p10);
            acc[k] = fmaf(cc.wv, fc == fa ? fmaf(an0, a1 - a0, a0) : 0.0f, acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < FZC; ++k)
        if (k < nz) St<SV>::st(vol, idx0 + k * MN, acc[k]);
}

// LDS-staged form for FP32 views.  A workgroup owns a tile of FTX x FTY
// columns, FTZ voxels tall, and takes views FVB at a time: wave j bounds the
// tile's detector footprint in view j by interval arithmetic on t, s and z
// (a 0.5-pixel margin absorbs rounding), then the workgroup copies that
// window into LDS, one 64-pixel row segment per wave instruction, by buffer
// loads that write LDS directly.  Window rows sit FRS floats apart, so a
// lane's 2x2 stencil is two ds_read2 with constant offsets; the stride of 80
// spreads a wave's rows over the LDS banks, and 48 rows of it for two views
// keep five workgroups per CU.  The host admits a geometry only when a bound
// on every tile's window fits (fdk_lds_fits), so there is no slow path here;
// the indices are still clamped into the window, which only moves reads of
// voxels whose term the in-detector test zeroes.  Tiles clear of the
// detector's top and bottom rows in every view skip the test and the clamp.
constexpr int FTX = 16, FTY = 16, FTZ = 16, FVB = 2, FRS = 80, FROWS = 48;

__device__ __forceinline__ void lds_row_dma(__amdgpu_buffer_rsrc_t rs, float *dst, uint32_t off) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)dst, 4, off, 0, 0, 0);
}

template <typename SV>
__global__ __launch_bounds__(256) void k_fdk_bp_lds(DGeo<float> d, const float4 *__restrict__ vtab,
                                                    const float *__restrict__ filt, SV *__restrict__ vol, int nv,
                                                    float wscale) {
    __shared__ float win[FVB][FROWS * FRS];
    // per view of the batch: float bounds {first row, last stencil row, first
    // column, last stencil column}; {first row, first column, rows, rows clear
    // of both detector edges}
    __shared__ float4 wlim[FVB];
    __shared__ int4 wdim[FVB];
    const int3 tile = xcd_tile();
    const int tx = threadIdx.x & (FTX - 1), ty = threadIdx.x / FTX;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ix = tile.x * FTX + tx, iy = tile.y * FTY + ty, iz0 = tile.z * FTZ;
    const bool live = ix < d.M && iy < d.N;
    const int nz = min(FTZ, d.Z - iz0);
    const float x = d.bx + d.vx * 0.5f + (float)ix * d.vx;
    const float y = d.by + d.vy * 0.5f + (float)iy * d.vy;
    const float z0 = d.bz + d.vz * 0.5f + (float)iz0 * d.vz;
    // voxel-centre extent of the tile
    const float xa = d.bx + d.vx * 0.5f + (float)(tile.x * FTX) * d.vx;
    const float xb = d.bx + d.vx * 0.5f + (float)(min(tile.x * FTX + FTX, d.M) - 1) * d.vx;
    const float ya = d.by + d.vy * 0.5f + (float)(tile.y * FTY) * d.vy;
    const float yb = d.by + d.vy * 0.5f + (float)(min(tile.y * FTY + FTY, d.N) - 1) * d.vy;
    const float zb = d.bz + d.vz * 0.5f + (float)(iz0 + nz - 1) * d.vz;
    const float xm = 0.5f * (xa + xb), ym = 0.5f * (ya + yb), hx = 0.5f * (xb - xa), hy = 0.5f * (yb - ya);
    const int64_t MN = (int64_t)d.M * d.N;
    const int64_t idx0 = ((int64_t)iz0 * d.N + iy) * d.M + ix;
    float acc[FTZ];
#pragma unroll
    for (int k = 0; k < FTZ; ++k) acc[k] = (live && k < nz) ? ldf(vol, idx0 + k * MN) : 0.0f;
    const float sdd = d.dso - d.dod;
    const int64_t P = (int64_t)d.R * d.ZP;
    const uint32_t vbytes = (uint32_t)(P * (int64_t)sizeof(float));
    const float wk = d.dso * d.dso * wscale, pm = -d.ymin * d.idu;
    const float kz0 = z0 * sdd * d.idv, kdz = d.vz * sdd * d.idv, km = -d.zmin * d.idv;
    // Tiles whose voxels stay clear of the detector's top and bottom rows in
    // every view (for any s in [-r, r], r the tile's largest distance from
    // the axis) skip the in-detector test; one code path per tile keeps the
    // accumulators in place.
    const float rt = sqrtf(fmaxf(xa * xa, xb * xb) + fmaxf(ya * ya, yb * yb));
    const float ilo = frcp(d.dso + rt), ihi = frcp(d.dso - rt);
    const float kall_lo = (fminf(z0 * ilo, z0 * ihi) * sdd - d.zmin) * d.idv;
    const float kall_hi = (fmaxf(zb * ilo, zb * ihi) * sdd - d.zmin) * d.idv;
    const bool clear = (d.dso - rt > 1e-3f * d.dso) && kall_lo >= 1.0f && kall_hi <= (float)(d.ZP - 3);
    auto views = [&](auto tile_interior) {
        for (int qb = 0; qb < nv; qb += FVB) {
            const int nb = min(FVB, nv - qb);
            __syncthreads(); // the previous batch's gathers are done
            if (wave < nb) {
                const float4 vt = vtab[qb + wave];
                const float s = vt.z, c = vt.w;
                const float tc = xm * c + ym * s, et = hx * fabsf(c) + hy * fabsf(s);
                const float sc = -xm * s + ym * c, es = hx * fabsf(s) + hy * fabsf(c);
                const float i0 = frcp(d.dso - (sc - es)), i1 = frcp(d.dso - (sc + es));
                const float tlo = tc - et, thi = tc + et;
                const float plo = (fminf(tlo * i0, tlo * i1) * sdd - d.ymin) * d.idu;
                const float phi = (fmaxf(thi * i0, thi * i1) * sdd - d.ymin) * d.idu;
                const float klo = (fminf(z0 * i0, z0 * i1) * sdd - d.zmin) * d.idv;
                const float khi = (fmaxf(zb * i0, zb * i1) * sdd - d.zmin) * d.idv;
                const float Rf = (float)d.R, Zf = (float)d.ZP;
                const float fra = floorf(fminf(fmaxf(klo - 0.5f, -1.0f), Zf));
                const float frb = floorf(fminf(fmaxf(khi + 0.5f, -1.0f), Zf)) + 1.0f;
                const int ca = (int)floorf(fminf(fmaxf(plo - 0.5f, 0.0f), Rf));
                const int cb = min(max((int)floorf(fminf(fmaxf(phi + 0.5f, 0.0f), Rf)) + 1, ca + 1), min(d.R - 1, ca + 63));
                const int ra = max((int)fra, 0);
                const int rb = min(max((int)frb, ra + 1), min(d.ZP - 1, ra + FROWS - 1));
                if (lane == 0) {
                    wdim[wave] = make_int4(ra, ca, rb - ra + 1, (fra >= 0.0f && frb <= (float)(d.ZP - 1)) ? 1 : 0);
                    wlim[wave] = make_float4((float)ra, (float)(rb - 1), (float)ca, (float)(cb - 1));
                }
            }
            __syncthreads();
    #pragma unroll
            for (int j = 0; j < FVB; ++j) {
                if (j >= nb) break;
                const int4 dim = wdim[j];
                const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(filt + (int64_t)(qb + j) * P, vbytes);
                const uint32_t col = (uint32_t)min(dim.y + lane, d.R - 1);
                const int rows = __builtin_amdgcn_readfirstlane(dim.z);
                for (int r = wave; r < rows; r += 4)
                    lds_row_dma(rs, &win[j][r * FRS], ((uint32_t)(dim.x + r) * (uint32_t)d.R + col) * 4u);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (!live) continue;
    #pragma unroll
            for (int j = 0; j < FVB; ++j) {
                if (j >= nb) break;
                const FdkCol cc = fdk_column(d, vtab[qb + j], x, y, wk, pm, kz0, kdz, km);
                // Stencil rows clamp into [ra, rb-1], which lies inside [0, ZP-2]
                // and holds the row of every voxel that passes the in-detector
                // test, so "the clamp left the row unchanged" is that test
                // (0 <= ksi < ZP-1); for a window clear of both detector edges
                // every voxel of the tile passes it.
                const int4 dim = wdim[j];
                const float4 lim = wlim[j];
                const float fbc = __builtin_amdgcn_fmed3f(cc.fb, lim.z, lim.w);
                // byte offset of the window origin, view j's window base folded in
                int base4 = ((int)fbc - dim.y - dim.x * FRS + j * (FROWS * FRS)) * 4;
                asm("" : "+v"(base4)); // keep the window base inside the per-voxel address
                const char *wj = (const char *)&win[0][0];
                if constexpr (decltype(tile_interior)::value) {
                    // every voxel of the tile passes the in-detector test in
                    // every view, and the margin keeps its stencil in the window
    #pragma unroll
                    for (int k = 0; k < FTZ; ++k) {
                        const float ksi = fmaf((float)k, cc.dk, cc.k0);
                        const float fa = floorf(ksi), an0 = ksi - fa;
                        const int o = __mul24((int)fa, FRS * 4) + base4;
                        const float p00 = *(const float *)(wj + o), p01 = *(const float *)(wj + o + 4);
                        const float p10 = *(const float *)(wj + o + FRS * 4), p11 = *(const float *)(wj + o + FRS * 4 + 4);
                        const float a0 = fmaf(cc.bn0, p01 - p00, p00), a1 = fmaf(cc.bn0, p11 - p10, p10);
                        acc[k] = fmaf(cc.wv, fmaf(an0, a1 - a0, a0), acc[k]);
                    }
                } else {
    #pragma unroll
                    for (int k = 0; k < FTZ; ++k) {
                        const float ksi = fmaf((float)k, cc.dk, cc.k0);
                        const float fa = floorf(ksi), an0 = ksi - fa;
                        const float fc = __builtin_amdgcn_fmed3f(fa, lim.x, lim.y);
                        const int o = __mul24((int)fc, FRS * 4) + base4;
                        const float p00 = *(const float *)(wj + o), p01 = *(const float *)(wj + o + 4);
                        const float p10 = *(const float *)(wj + o + FRS * 4), p11 = *(const float *)(wj + o + FRS * 4 + 4);
                        const float a0 = fmaf(cc.bn0, p01 - p00, p00), a1 = fmaf(cc.bn0, p11 - p10, p10);
                        acc[k] = fmaf(cc.wv, fc == fa ? fmaf(an0, a1 - a0, a0) : 0.0f, acc[k]);
                    }
                }
            }
        }
    };
    if (clear) views(std::true_type{});
    else views(std::false_type{});
    if (!live) return;
#pragma unroll
    for (int k = 0; k < FTZ; ++k)
        if (k < nz) St<SV>::st(vol, idx0 + k * MN, acc[k]);
}

// Host bound on the LDS window of every tile and view.  A tile's voxel
// centres lie within hd (the half-diagonal) of its centre, and DSO - s over
// the tile lies in [D - hd, D + hd]; the window then spans at most
// (2*hd/(D-hd) + (|t|+hd)*(1/(D-hd) - 1/(D+hd)))*(DSO-DOD)/du columns and
// ((FTZ-1)*vz/(D-hd) + |z|max*(1/(D-hd) - 1/(D+hd)))*(DSO-DOD)/dv rows,
// plus the margins.  Both grow with the distance r of the tile centre from
// the axis, so the bound takes r at the grid corner and scans the angle.
// False sends the call to k_fdk_bp_col.
static bool fdk_lds_fits(const DGeo<float> &d) {
    const double vx = d.vx, vy = d.vy, vz = d.vz, dso = d.dso, sdd = (double)d.dso - (double)d.dod;
    const double x0 = d.bx + 0.5 * vx, x1 = d.bx + (d.M - 0.5) * vx;
    const double y0 = d.by + 0.5 * vy, y1 = d.by + (d.N - 0.5) * vy;
    const double rmax = std::sqrt(std::max(x0 * x0, x1 * x1) + std::max(y0 * y0, y1 * y1));
    const double hd = 0.5 * std::sqrt(std::pow((FTX - 1) * vx, 2) + std::pow((FTY - 1) * vy, 2));
    const double zmax = std::max(std::fabs(d.bz + 0.5 * vz), std::fabs(d.bz + (d.Z - 0.5) * vz));
    if (!(dso - rmax - hd > 2e-3 * dso)) return false;
    double cols = 0.0, rows = 0.0;
    for (int i = 0; i <= 720; ++i) {
        const double ph = 3.141592653589793 * i / 720.0;
        const double t = rmax * std::fabs(std::cos(ph)), D = dso - rmax * std::sin(ph);
        const double lo = D - hd, hi = D + hd, spread = 1.0 / lo - 1.0 / hi;
        cols = std::max(cols, (2.0 * hd / lo + (t + hd) * spread) * sdd / d.du);
        rows = std::max(rows, ((FTZ - 1) * vz / lo + zmax * spread) * sdd / d.dv);
    }
    return cols + 5.0 <= 64.0 && rows + 5.0 <= (double)FROWS;
}

template <typename T, typename SP, typename SV>
static ct_status run_fdk_bp(ct_context *ctx, const ct_geometry &g, const void *filt, void *vol, int view0, int nv,
                            int sources) {
    ct_status st;
    const T *ang = (const T *)angle_table(ctx, g, sizeof(T) == 8, true, &st);
    if (!ang) return st;
    DGeo<T> d = derive<T>(g);
    const T pi = sizeof(T) == 8 ? (T)3.141592653589793 : (T)3.141592653589793f;
    const T dth = d.us_rate * pi / (T)180;
    const T usn = d.us_rate * (T)g.nviews;
    const double w = (double)dth * 360.0 / (double)usn;
    dim3 block(64, 4, 1), grid((d.M + 63) / 64, (d.N + 3) / 4, d.Z);
    ProfScope ps(ctx, CT_K_FDK_BP);
    if (d.NS > 1) {
        hipLaunchKernelGGL((k_fdk_bp_ms<T, SP, SV>), grid, block, 0, ctx->stream, d, ang, (const SP *)filt, (SV *)vol,
                           view0, nv, dth, usn, sources);
        return launch_check(ctx, "fdk backprojection");
    }
    if constexpr (std::is_same<T, float>::value && !std::is_same<SP, double>::value) {
        if (ctx->math == CT_MATH_FAST && (uint64_t)g.ncols * g.nrows * sizeof(SP) < (1ull << 32)) {
            float4 *vt = (float4 *)slot(ctx, SLOT_FDKV, sizeof(float4) * (size_t)nv, &st);
            if (!vt) return st;
            hipLaunchKernelGGL(k_fdk_views, dim3((nv + 255) / 256), dim3(256), 0, ctx->stream, (const float *)ang, vt,
                               view0, nv, (d.dso - d.dod) * d.idu);
            if (std::is_same<SP, float>::value && fdk_lds_fits(d)) {
                dim3 tg((d.M + FTX - 1) / FTX, (d.N + FTY - 1) / FTY, (d.Z + FTZ - 1) / FTZ);
                hipLaunchKernelGGL((k_fdk_bp_lds<SV>), tg, dim3(256), 0, ctx->stream, d, (const float4 *)vt,
                                   (const float *)filt, (SV *)vol, nv, (float)w);
            } else {
                dim3 cg((d.M + 63) / 64, (d.N + 3) / 4, (d.Z + FZC - 1) / FZC);
                hipLaunchKernelGGL((k_fdk_bp_col<SP, SV>), cg, block, 0, ctx->stream, d, (const float4 *)vt,
                                   (const SP *)filt, (SV *)vol, nv, (float)w);
            }
            return launch_check(ctx, "fdk backprojection");
        }
    }
    hipLaunchKernelGGL((k_fdk_bp<T, SP, SV>), grid, block, 0, ctx->stream, d, ang, (const SP *)filt, (SV *)vol, view0,
                       nv, dth, usn);
    return launch_check(ctx, "fdk backprojection");
}

template <typename SP>
static ct_status fdk_bp_vol(ct_context *ctx, const ct_geometry &g, const void *filt, void *vol, ct_dtype vdt, int v0,
                            int nv, int sources) {
    switch (vdt) {
    case CT_F32: return run_fdk_bp<float, SP, float>(ctx, g, filt, vol, v0, nv, sources);
    case CT_F16: return run_fdk_bp<float, SP, __half>(ctx, g, filt, vol, v0, nv, sources);
    case CT_BF16: return run_fdk_bp<float, SP, __hip_bfloat16>(ctx, g, filt, vol, v0, nv, sources);
    case CT_F64: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "fdk backprojection: F64 volumes need F64 projections");
}

ct_status launch_fdk_bp(ct_context *ctx, const ct_geometry &g, const void *filt, ct_dtype pdt, void *vol, ct_dtype vdt,
                        int32_t view0, int32_t nv, int sources) {
    if (nv <= 0) return CT_OK;
    if (pdt == CT_F64 || vdt == CT_F64) {
        if (pdt != CT_F64 || vdt != CT_F64) return fail(ctx, CT_ERR_DTYPE, "fdk backprojection: mixed F64 dtypes");
        return run_fdk_bp<double, double, double>(ctx, g, filt, vol, view0, nv, sources);
    }
    switch (pdt) {
    case CT_F32: return fdk_bp_vol<float>(ctx, g, filt, vol, vdt, view0, nv, sources);
    case CT_F16: return fdk_bp_vol<__half>(ctx, g, filt, vol, vdt, view0, nv, sources);
    case CT_BF16: return fdk_bp_vol<__hip_bfloat16>(ctx, g, filt, vol, vdt, view0, nv, sources);
    default: break;
    }
    return fail(ctx, CT_ERR_DTYPE, "fdk backprojection: bad projection dtype");
}

} // namespace ct
