// ops.hip - fused elementwise, stencil and FP64 reduction kernels:
// FISTA momentum, TV-FGP, the FGP difference pair D / D^T, TV_eps and its
// gradient, TV / residual / dot norms, fills, conversions and seeded normals.
#include "ct_internal.hpp"
#include <cmath>
#include <type_traits>

namespace ct {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;

// --------------------------------------------------------------- basics
template <typename S>
__global__ void k_fill(S *p, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        St<S>::st(p, i, v);
}
__global__ void k_fill_f64(double *p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static dim3 grid_for(int64_t n, int threads = 256) {
    int64_t b = (n + threads - 1) / threads;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

ct_status launch_fill(ct_context *ctx, void *p, ct_dtype dt, int64_t n, double value) {
    if (n <= 0) return CT_OK;
    dim3 g = grid_for(n), b(256);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_fill<float>, g, b, 0, ctx->stream, (float *)p, n, (float)value); break;
    case CT_F16: hipLaunchKernelGGL(k_fill<__half>, g, b, 0, ctx->stream, (__half *)p, n, (float)value); break;
    case CT_BF16: hipLaunchKernelGGL(k_fill<__hip_bfloat16>, g, b, 0, ctx->stream, (__hip_bfloat16 *)p, n, (float)value); break;
    case CT_F64: hipLaunchKernelGGL(k_fill_f64, g, b, 0, ctx->stream, (double *)p, n, value); break;
    }
    return launch_check(ctx, "fill");
}

template <typename SI, typename SO>
__global__ void k_convert(const SI *in, SO *out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (std::is_same<SI, double>::value && std::is_same<SO, double>::value) out[i] = in[i];
        else if constexpr (std::is_same<SI, double>::value) St<SO>::st(out, i, (float)in[i]);
        else if constexpr (std::is_same<SO, double>::value) out[i] = (double)ldf(in, i);
        else St<SO>::st(out, i, ldf(in, i));
    }
}

template <typename SI>
static ct_status conv_out(ct_context *ctx, const void *in, void *out, ct_dtype odt, int64_t n) {
    dim3 g = grid_for(n), b(256);
    switch (odt) {
    case CT_F32: hipLaunchKernelGGL((k_convert<SI, float>), g, b, 0, ctx->stream, (const SI *)in, (float *)out, n); break;
    case CT_F16: hipLaunchKernelGGL((k_convert<SI, __half>), g, b, 0, ctx->stream, (const SI *)in, (__half *)out, n); break;
    case CT_BF16: hipLaunchKernelGGL((k_convert<SI, __hip_bfloat16>), g, b, 0, ctx->stream, (const SI *)in, (__hip_bfloat16 *)out, n); break;
    case CT_F64: hipLaunchKernelGGL((k_convert<SI, double>), g, b, 0, ctx->stream, (const SI *)in, (double *)out, n); break;
    }
    return launch_check(ctx, "convert");
}

ct_status launch_convert(ct_context *ctx, const void *in, ct_dtype idt, void *out, ct_dtype odt, int64_t n) {
    if (n <= 0) return CT_OK;
    switch (idt) {
    case CT_F32: return conv_out<float>(ctx, in, out, odt, n);
    case CT_F16: return conv_out<__half>(ctx, in, out, odt, n);
    case CT_BF16: return conv_out<__hip_bfloat16>(ctx, in, out, odt, n);
    case CT_F64: return conv_out<double>(ctx, in, out, odt, n);
    }
    return fail(ctx, CT_ERR_DTYPE, "convert: bad dtype");
}

// z-quad layout conversion: one thread per (x, y, quad); the four z values of
// a quad are read at stride M*N (coalesced over x) and written contiguously.
template <typename S, bool PACK>
__global__ __launch_bounds__(256) void k_zq(const S *__restrict__ in, S *__restrict__ out, int M, int N, int Z) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, q = blockIdx.z;
    if (x >= M) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t lin = (int64_t)y * M + x;
    const int64_t qb = ((int64_t)q * MN + lin) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int z = 4 * q + k;
        if constexpr (PACK) out[qb + k] = z < Z ? in[(int64_t)z * MN + lin] : S(0.0f);
        else if (z < Z) out[(int64_t)z * MN + lin] = in[qb + k];
    }
}
template <bool PACK>
static ct_status zq_copy(ct_context *ctx, const void *in, ct_dtype dt, void *out, int M, int N, int Z) {
    dim3 g((M + 255) / 256, N, (Z + 3) / 4), b(256);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL((k_zq<float, PACK>), g, b, 0, ctx->stream, (const float *)in, (float *)out, M, N, Z); break;
    case CT_F16: hipLaunchKernelGGL((k_zq<__half, PACK>), g, b, 0, ctx->stream, (const __half *)in, (__half *)out, M, N, Z); break;
    case CT_BF16:
        hipLaunchKernelGGL((k_zq<__hip_bfloat16, PACK>), g, b, 0, ctx->stream, (const __hip_bfloat16 *)in,
                           (__hip_bfloat16 *)out, M, N, Z);
        break;
    default: return fail(ctx, CT_ERR_DTYPE, "z-quad copy: unsupported dtype");
    }
    return launch_check(ctx, "z-quad copy");
}
ct_status launch_zq_pack(ct_context *ctx, const void *vol, ct_dtype dt, void *volq, int M, int N, int Z) {
    return zq_copy<true>(ctx, vol, dt, volq, M, N, Z);
}
ct_status launch_zq_unpack(ct_context *ctx, const void *volq, ct_dtype dt, void *vol, int M, int N, int Z) {
    return zq_copy<false>(ctx, volq, dt, vol, M, N, Z);
}

// Transpose of the lateral axes of a z-quad-layout volume: every z-quad
// plane holds an R x C matrix of quads (Q = 16 bytes for FP32, 8 for 16-bit
// storage), written out as C x R, so [z/4][y][x][z%4] becomes
// [z/4][x][y][z%4] and back.  32 x 32 quad tiles go through LDS so that both
// the reads and the writes are row-contiguous.
template <typename Q>
__global__ __launch_bounds__(256) void k_zq_transpose(const Q *__restrict__ in, Q *__restrict__ out, int R, int C) {
    __shared__ Q tile[32][33];
    const int64_t plane = (int64_t)blockIdx.z * R * C;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = r0 + ty + k, c = c0 + tx;
        if (r < R && c < C) tile[ty + k][tx] = in[plane + (int64_t)r * C + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int c = c0 + ty + k, r = r0 + tx;
        if (r < R && c < C) out[plane + (int64_t)c * R + r] = tile[tx][ty + k];
    }
}
// to_tr: [z/4][y][x][z%4] -> [z/4][x][y][z%4]; otherwise the reverse
ct_status launch_zq_transpose(ct_context *ctx, const void *in, ct_dtype dt, void *out, int M, int N, int Z,
                              bool to_tr) {
    const int R = to_tr ? N : M, C = to_tr ? M : N;
    dim3 g((C + 31) / 32, (R + 31) / 32, (Z + 3) / 4), b(256);
    switch (dt) {
    case CT_F32:
        hipLaunchKernelGGL(k_zq_transpose<uint4>, g, b, 0, ctx->stream, (const uint4 *)in, (uint4 *)out, R, C);
        break;
    case CT_F16:
    case CT_BF16:
        hipLaunchKernelGGL(k_zq_transpose<uint2>, g, b, 0, ctx->stream, (const uint2 *)in, (uint2 *)out, R, C);
        break;
    default: return fail(ctx, CT_ERR_DTYPE, "z-quad transpose: unsupported dtype");
    }
    return launch_check(ctx, "z-quad transpose");
}

template <typename SO>
__global__ __launch_bounds__(256) void k_zq_unpack_f64(const double *__restrict__ in, SO *__restrict__ out, int M, int N,
                                                       int Z) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, q = blockIdx.z;
    if (x >= M) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t lin = (int64_t)y * M + x;
    const int64_t qb = ((int64_t)q * MN + lin) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int z = 4 * q + k;
        if (z < Z) St<SO>::st(out, (int64_t)z * MN + lin, (typename LdT<SO>::type)in[qb + k]);
    }
}
ct_status launch_zq_unpack_f64(ct_context *ctx, const double *accq, void *vol, ct_dtype dt, int M, int N, int Z) {
    dim3 g((M + 255) / 256, N, (Z + 3) / 4), b(256);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL((k_zq_unpack_f64<float>), g, b, 0, ctx->stream, accq, (float *)vol, M, N, Z); break;
    case CT_F16: hipLaunchKernelGGL((k_zq_unpack_f64<__half>), g, b, 0, ctx->stream, accq, (__half *)vol, M, N, Z); break;
    case CT_BF16:
        hipLaunchKernelGGL((k_zq_unpack_f64<__hip_bfloat16>), g, b, 0, ctx->stream, accq, (__hip_bfloat16 *)vol, M, N, Z);
        break;
    case CT_F64: hipLaunchKernelGGL((k_zq_unpack_f64<double>), g, b, 0, ctx->stream, accq, (double *)vol, M, N, Z); break;
    }
    return launch_check(ctx, "z-quad unpack");
}

// y = a*x + b*y
template <typename S>
__global__ void k_axpby(const S *x, S *y, int64_t n, float a, float b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        St<S>::st(y, i, a * ldf(x, i) + b * ldf(y, i));
}
__global__ void k_axpby_f64(const double *x, double *y, int64_t n, double a, double b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}
ct_status launch_axpby(ct_context *ctx, const void *x, void *y, ct_dtype dt, int64_t n, double a, double b) {
    dim3 g = grid_for(n), bl(256);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_axpby<float>, g, bl, 0, ctx->stream, (const float *)x, (float *)y, n, (float)a, (float)b); break;
    case CT_F16: hipLaunchKernelGGL(k_axpby<__half>, g, bl, 0, ctx->stream, (const __half *)x, (__half *)y, n, (float)a, (float)b); break;
    case CT_BF16: hipLaunchKernelGGL(k_axpby<__hip_bfloat16>, g, bl, 0, ctx->stream, (const __hip_bfloat16 *)x, (__hip_bfloat16 *)y, n, (float)a, (float)b); break;
    case CT_F64: hipLaunchKernelGGL(k_axpby_f64, g, bl, 0, ctx->stream, (const double *)x, (double *)y, n, a, b); break;
    }
    return launch_check(ctx, "axpby");
}

// Seeded N(0,1): splitmix64 of (seed, index) -> two uniforms -> Box-Muller.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
template <typename S>
__global__ void k_randn(S *p, int64_t n, uint64_t seed) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t a = splitmix64(seed ^ splitmix64((uint64_t)i * 2 + 1));
        uint64_t b = splitmix64(a + 0x632BE59BD9B4E019ull);
        double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        double u2 = ((double)(b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        if constexpr (std::is_same<S, double>::value) p[i] = r;
        else St<S>::st(p, i, (float)r);
    }
}
ct_status launch_randn(ct_context *ctx, void *p, ct_dtype dt, int64_t n, uint64_t seed) {
    dim3 g = grid_for(n), b(256);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_randn<float>, g, b, 0, ctx->stream, (float *)p, n, seed); break;
    case CT_F16: hipLaunchKernelGGL(k_randn<__half>, g, b, 0, ctx->stream, (__half *)p, n, seed); break;
    case CT_BF16: hipLaunchKernelGGL(k_randn<__hip_bfloat16>, g, b, 0, ctx->stream, (__hip_bfloat16 *)p, n, seed); break;
    case CT_F64: hipLaunchKernelGGL(k_randn<double>, g, b, 0, ctx->stream, (double *)p, n, seed); break;
    }
    return launch_check(ctx, "randn");
}

// ------------------------------------------------------------ reductions
// Per-block FP64 partial sums (wave64 shuffles, then LDS across the 4 waves),
// then one block folds the partials in fixed order: deterministic.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[RED_THREADS / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) part[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < RED_THREADS / 64; ++w) r += part[w];
    return r;
}

__global__ __launch_bounds__(RED_THREADS) void k_fold(const double *part, int n, double *out) {
    double v = 0;
    for (int i = threadIdx.x; i < n; i += RED_THREADS) v += part[i];
    double r = block_sum(v);
    if (threadIdx.x == 0) out[0] = r;
}

// mode 0: sum a*b; mode 1: sum (a-b)^2; mode 2: sum a (b unused)
template <typename S>
__global__ __launch_bounds__(RED_THREADS) void k_dot(const S *a, const S *b, int64_t n, int mode, double *part) {
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
        double x, y = 0.0;
        if constexpr (std::is_same<S, double>::value) { x = a[i]; if (mode != 2) y = b[i]; }
        else { x = (double)ldf(a, i); if (mode != 2) y = (double)ldf(b, i); }
        if (mode == 0) acc += x * y;
        else if (mode == 1) { double e = x - y; acc += e * e; }
        else acc += x;
    }
    double r = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

static ct_status finish_reduce(ct_context *ctx, double *part, int nblk, double *out) {
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(RED_THREADS), 0, ctx->stream, part, nblk, part + RED_BLOCKS);
    ct_status st = launch_check(ctx, "reduce");
    if (st != CT_OK) return st;
    st = hip_check(ctx, hipMemcpyAsync(out, part + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "reduce copy");
    if (st != CT_OK) return st;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "reduce sync");
}

static int red_blocks(int64_t n) {
    int64_t b = (n + RED_THREADS - 1) / RED_THREADS;
    if (b > RED_BLOCKS) b = RED_BLOCKS;
    return b < 1 ? 1 : (int)b;
}

ct_status reduce_dot(ct_context *ctx, const void *a, const void *b, int64_t n, ct_dtype dt, int mode, double *out) {
    ct_status st;
    double *part = (double *)slot(ctx, SLOT_RED, sizeof(double) * (RED_BLOCKS + 8), &st);
    if (!part) return st;
    const int nb = red_blocks(n);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_dot<float>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const float *)a, (const float *)b, n, mode, part); break;
    case CT_F16: hipLaunchKernelGGL(k_dot<__half>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const __half *)a, (const __half *)b, n, mode, part); break;
    case CT_BF16: hipLaunchKernelGGL(k_dot<__hip_bfloat16>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const __hip_bfloat16 *)a, (const __hip_bfloat16 *)b, n, mode, part); break;
    case CT_F64: hipLaunchKernelGGL(k_dot<double>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const double *)a, (const double *)b, n, mode, part); break;
    }
    st = launch_check(ctx, "dot");
    if (st != CT_OK) return st;
    return finish_reduce(ctx, part, nb, out);
}

template <typename S> __device__ __forceinline__ double ld64(const S *p, int64_t i) {
    if constexpr (std::is_same<S, double>::value) return p[i];
    else return (double)ldf(p, i);
}

// mode 0: isotropic forward-difference TV (V63/kernel_tool_functions.cu:39-65);
// mode 1: TV_eps with the backward differences that exist.
template <typename S>
__global__ __launch_bounds__(RED_THREADS) void k_tv(const S *f, int M, int N, int Z, int mode, double eps, double *part) {
    const int64_t MN = (int64_t)M * N, V = MN * Z;
    double acc = 0;
    for (int64_t v = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; v < V; v += (int64_t)gridDim.x * RED_THREADS) {
        const int i = (int)(v % M), j = (int)((v / M) % N), k = (int)(v / MN);
        const double c = ld64(f, v);
        if (mode == 0) {
            double dx = i < M - 1 ? c - ld64(f, v + 1) : 0.0;
            double dy = j < N - 1 ? c - ld64(f, v + M) : 0.0;
            double dz = k < Z - 1 ? c - ld64(f, v + MN) : 0.0;
            acc += sqrt(dx * dx + dy * dy + dz * dz);
        } else {
            double q = eps;
            if (i > 0) { double e = c - ld64(f, v - 1); q += e * e; }
            if (j > 0) { double e = c - ld64(f, v - M); q += e * e; }
            if (k > 0) { double e = c - ld64(f, v - MN); q += e * e; }
            acc += sqrt(q);
        }
    }
    double r = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

ct_status reduce_tv(ct_context *ctx, const void *f, ct_dtype dt, int M, int N, int Z, int mode, double eps, double *out) {
    ct_status st;
    double *part = (double *)slot(ctx, SLOT_RED, sizeof(double) * (RED_BLOCKS + 8), &st);
    if (!part) return st;
    const int nb = red_blocks((int64_t)M * N * Z);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_tv<float>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const float *)f, M, N, Z, mode, eps, part); break;
    case CT_F16: hipLaunchKernelGGL(k_tv<__half>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const __half *)f, M, N, Z, mode, eps, part); break;
    case CT_BF16: hipLaunchKernelGGL(k_tv<__hip_bfloat16>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const __hip_bfloat16 *)f, M, N, Z, mode, eps, part); break;
    case CT_F64: hipLaunchKernelGGL(k_tv<double>, dim3(nb), dim3(RED_THREADS), 0, ctx->stream, (const double *)f, M, N, Z, mode, eps, part); break;
    }
    st = launch_check(ctx, "tv");
    if (st != CT_OK) return st;
    return finish_reduce(ctx, part, nb, out);
}

// d TV_eps / d x_v: own term minus the terms of the three forward neighbours.
template <typename S>
__global__ void k_tv_eps_grad(const S *f, S *g, int M, int N, int Z, double eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= M) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t v = (int64_t)k * MN + (int64_t)j * M + i;
    auto phi = [&](int a, int b, int cc, int64_t w) {
        const double c = ld64(f, w);
        double q = eps;
        if (a > 0) { double e = c - ld64(f, w - 1); q += e * e; }
        if (b > 0) { double e = c - ld64(f, w - M); q += e * e; }
        if (cc > 0) { double e = c - ld64(f, w - MN); q += e * e; }
        return sqrt(q);
    };
    const double c = ld64(f, v);
    double num = 0;
    if (i > 0) num += c - ld64(f, v - 1);
    if (j > 0) num += c - ld64(f, v - M);
    if (k > 0) num += c - ld64(f, v - MN);
    double gv = num / phi(i, j, k, v);
    if (i < M - 1) gv -= (ld64(f, v + 1) - c) / phi(i + 1, j, k, v + 1);
    if (j < N - 1) gv -= (ld64(f, v + M) - c) / phi(i, j + 1, k, v + M);
    if (k < Z - 1) gv -= (ld64(f, v + MN) - c) / phi(i, j, k + 1, v + MN);
    if constexpr (std::is_same<S, double>::value) g[v] = gv;
    else St<S>::st(g, v, (float)gv);
}

ct_status launch_tv_eps_grad(ct_context *ctx, const void *f, void *g, ct_dtype dt, int M, int N, int Z, double eps) {
    dim3 b(256), gr((M + 255) / 256, N, Z);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_tv_eps_grad<float>, gr, b, 0, ctx->stream, (const float *)f, (float *)g, M, N, Z, eps); break;
    case CT_F16: hipLaunchKernelGGL(k_tv_eps_grad<__half>, gr, b, 0, ctx->stream, (const __half *)f, (__half *)g, M, N, Z, eps); break;
    case CT_BF16: hipLaunchKernelGGL(k_tv_eps_grad<__hip_bfloat16>, gr, b, 0, ctx->stream, (const __hip_bfloat16 *)f, (__hip_bfloat16 *)g, M, N, Z, eps); break;
    case CT_F64: hipLaunchKernelGGL(k_tv_eps_grad<double>, gr, b, 0, ctx->stream, (const double *)f, (double *)g, M, N, Z, eps); break;
    }
    return launch_check(ctx, "tv_eps_grad");
}

// ------------------------------------------------------------ FISTA
// y = x + gamma*(x - x_prev); x_prev = x (V63/main_recon.cu:375-382), one pass.
template <typename S, typename T>
__global__ void k_fista(S *x, S *xp, S *y, int64_t n, T gamma) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T xi, pi;
        if constexpr (std::is_same<S, double>::value) { xi = x[i]; pi = xp[i]; }
        else { xi = ldf(x, i); pi = ldf(xp, i); }
        const T yi = xi + gamma * (xi - pi);
        if constexpr (std::is_same<S, double>::value) { xp[i] = xi; y[i] = yi; }
        else { St<S>::st(xp, i, xi); St<S>::st(y, i, yi); }
    }
}
ct_status launch_fista(ct_context *ctx, void *x, void *xp, void *y, int64_t n, ct_dtype dt, double gamma) {
    dim3 g = grid_for(n), b(256);
    ProfScope ps(ctx, CT_K_FISTA);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL((k_fista<float, float>), g, b, 0, ctx->stream, (float *)x, (float *)xp, (float *)y, n, (float)gamma); break;
    case CT_F16: hipLaunchKernelGGL((k_fista<__half, float>), g, b, 0, ctx->stream, (__half *)x, (__half *)xp, (__half *)y, n, (float)gamma); break;
    case CT_BF16: hipLaunchKernelGGL((k_fista<__hip_bfloat16, float>), g, b, 0, ctx->stream, (__hip_bfloat16 *)x, (__hip_bfloat16 *)xp, (__hip_bfloat16 *)y, n, (float)gamma); break;
    case CT_F64: hipLaunchKernelGGL((k_fista<double, double>), g, b, 0, ctx->stream, (double *)x, (double *)xp, (double *)y, n, gamma); break;
    }
    return launch_check(ctx, "fista");
}

// ------------------------------------------------------------ TV-FGP
// Dual fields r,s,t (and the momentum copies o,p,q) are volume-sized arrays in
// the arithmetic type: r[v] is the x-face between v and v+e_x, zero on the last
// x plane (the reference's staggered d_r[i+1], V63/host_FGP_Denoise_GPUx7.c:8).
// L = b - lambda * D^T(r,s,t)   (FGP_operator_L_Pc, V63/kernel_FGP_Denoise_GPUx7.cu:2-24)
// Dual fields stored as D (the arithmetic type T, or the 16-bit storage type
// of a 16-bit volume: the duals lie in the unit ball, so FP16/BF16 storage
// moves the result by ~lambda * 1e-3 at most); arithmetic stays in T.
template <typename T, typename D> __device__ __forceinline__ T ldd(const D *p, int64_t i) {
    if constexpr (std::is_same<T, D>::value) return p[i];
    else return (T)ldf(p, i);
}
template <typename T, typename D> __device__ __forceinline__ void std_(D *p, int64_t i, T v) {
    if constexpr (std::is_same<T, D>::value) p[i] = v;
    else St<D>::st(p, i, (float)v);
}

template <typename S, typename T, typename D = T>
__global__ __launch_bounds__(256) void k_fgp_L(const S *b, const D *r, const D *s, const D *t, S *Lout, T *Lt,
                                               int M, int N, int Z, T lambda) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= M) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t v = (int64_t)k * MN + (int64_t)j * M + i;
    const T rm = i > 0 ? ldd<T>(r, v - 1) : (T)0;
    const T sm = j > 0 ? ldd<T>(s, v - M) : (T)0;
    const T tm = k > 0 ? ldd<T>(t, v - MN) : (T)0;
    const T tl = (ldd<T>(r, v) - rm) + (ldd<T>(s, v) - sm) + (ldd<T>(t, v) - tm);
    T bv;
    if constexpr (std::is_same<S, double>::value) bv = b[v]; else bv = ldf(b, v);
    const T L = bv - lambda * tl;
    if (Lt) Lt[v] = L;
    if (Lout) { if constexpr (std::is_same<S, double>::value) Lout[v] = L; else St<S>::st(Lout, v, L); }
}

// One FGP inner iteration in one pass (L never goes to memory).  The dual
// fields are double-buffered: the pass reads r, s, t of the previous
// iteration from one set and writes the next iteration's to the other, so a
// workgroup can read its neighbours' old values at the tile borders whatever
// their progress.  A workgroup owns a 128 x 8 (x, y) tile and marches a chunk
// of z slices: it stages the old r, s, t of the next slice in LDS with a
// one-voxel border, forms L of that slice on the tile and its right column
// and bottom row (the neighbours the dual step needs), then updates the
// current slice.  Per voxel the arithmetic is that of k_fgp_L followed by the
// dual gradient step, unit-ball projection and momentum of FGP_operator_LT_Pp
// (V63/kernel_FGP_Denoise_GPUx7.cu:27-88).  HBM traffic ~13.8 words per voxel (13 for b, the old duals, o,
// p, q read and the new duals, o, p, q written; the rest is the staged
// border) against 18 for the two-kernel form.
constexpr int FGX = 128, FGY = 8;            // tile (threads)
constexpr int FEX = FGX + 2, FEY = FGY + 2;  // staged extent: one voxel of border
constexpr int FLX = FGX + 1, FLY = FGY + 1;  // L extent: tile + right column + bottom row
constexpr int FGZ = 64;                      // slices per workgroup

template <typename S, typename T, typename D = T>
__global__ __launch_bounds__(FGX * FGY) void k_fgp_step(const S *__restrict__ b, const D *__restrict__ rA,
                                                        const D *__restrict__ sA, const D *__restrict__ tA,
                                                        D *__restrict__ rB, D *__restrict__ sB, D *__restrict__ tB,
                                                        D *__restrict__ o, D *__restrict__ p, D *__restrict__ q, int M,
                                                        int N, int Z, T lambda, T g, T a12) {
    // staged old duals: three slices in flight (written one step ahead of
    // their first reader, so no barrier guards the write); L: two slices
    __shared__ T sr[3][FEY * FEX], ss[3][FEY * FEX], stt[3][FEY * FEX];
    __shared__ T sL[2][FLY * FLX];
    constexpr int NT = FGX * FGY;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * FGX + tx;
    // x/y neighbours on one XCD: their border lines meet in its L2
    const int3 tile = xcd_tile();
    const int i0 = tile.x * FGX, j0 = tile.y * FGY;
    const int k0 = tile.z * FGZ, k1 = min(k0 + FGZ, Z);
    if (k0 >= k1) return;
    const int64_t MN = (int64_t)M * N;
    // this thread's staged and L entries (the same in every slice): offset
    // in the slice, or -1 outside the volume / past the extent
    int soff[2], loff[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int e = tid + u * NT;
        soff[u] = -1; loff[u] = -1;
        if (e < FEX * FEY) {
            const int ly = e / FEX - 1, lx = e - (ly + 1) * FEX - 1;
            const int i = i0 + lx, j = j0 + ly;
            if (i >= 0 && i < M && j >= 0 && j < N) soff[u] = j * M + i;
        }
        if (e < FLX * FLY) {
            const int ly = e / FLX, lx = e - ly * FLX;
            const int i = i0 + lx, j = j0 + ly;
            if (i < M && j < N) loff[u] = j * M + i;
        }
    }
    // every load of a slice is issued one step before its use: the staged
    // duals (extended tile, zero outside the volume), b on the L extent and
    // the owner's o, p, q; held in their raw storage form until used
    using RD = typename RawOf<D>::type;
    using RS = typename RawOf<S>::type;
    RD gr[2], gs[2], gt[2];
    RS gb[2];
    auto fetch_stage = [&](int k) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            gr[u] = 0; gs[u] = 0; gt[u] = 0;
            if (soff[u] >= 0) {
                const int64_t v = (int64_t)k * MN + soff[u];
                gr[u] = ld_raw(rA, v); gs[u] = ld_raw(sA, v); gt[u] = ld_raw(tA, v);
            }
        }
    };
    auto put_stage = [&](int k) {
        const int bf = k % 3;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = tid + u * NT;
            if (e < FEX * FEY) {
                sr[bf][e] = from_raw<T, D>(gr[u]); ss[bf][e] = from_raw<T, D>(gs[u]);
                stt[bf][e] = from_raw<T, D>(gt[u]);
            }
        }
    };
    auto fetch_b = [&](int k) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            gb[u] = 0;
            if (loff[u] >= 0) gb[u] = ld_raw(b, (int64_t)k * MN + loff[u]);
        }
    };
    // L of slice k on the tile + right column + bottom row (k_fgp_L)
    auto formL = [&](int k, const RS (&bv)[2]) {
        const int bf = k % 3, pf = (k + 2) % 3;   // slices k and k - 1
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = tid + u * NT;
            if (e >= FLX * FLY) continue;
            const int ly = e / FLX, lx = e - ly * FLX;
            const int i = i0 + lx, j = j0 + ly;
            T L = 0;
            if (i < M && j < N) {
                const int c = (ly + 1) * FEX + (lx + 1);
                const T rm = i > 0 ? sr[bf][c - 1] : (T)0;
                const T sm = j > 0 ? ss[bf][c - FEX] : (T)0;
                const T tm = k > 0 ? stt[pf][c] : (T)0;
                const T tl = (sr[bf][c] - rm) + (ss[bf][c] - sm) + (stt[bf][c] - tm);
                L = from_raw<T, S>(bv[u]) - lambda * tl;
            }
            sL[k & 1][e] = L;
        }
    };
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < M && j < N;
    RD po = 0, pp = 0, pq = 0;
    auto fetch_own = [&](int k) {
        if (own) {
            const int64_t v = (int64_t)k * MN + (int64_t)j * M + i;
            po = ld_raw(o, v); pp = ld_raw(p, v); pq = ld_raw(q, v);
        }
    };
    // prologue: slices k0 - 1 and k0 staged, L(k0) formed, slice k0 + 1's
    // duals, b of k0 + 1 and the owner's o, p, q of k0 in flight
    if (k0 > 0) { fetch_stage(k0 - 1); put_stage(k0 - 1); }
    fetch_stage(k0); put_stage(k0);
    fetch_b(k0);
    __syncthreads();
    formL(k0, gb);
    if (k0 + 1 < Z) { fetch_stage(k0 + 1); fetch_b(k0 + 1); }
    fetch_own(k0);
    for (int k = k0; k < k1; ++k) {
        const bool nx = k + 1 < Z;
        if (nx) put_stage(k + 1);
        const RD co = po, cp = pp, cq = pq;
        const RS bn[2] = {gb[0], gb[1]};
        if (k + 2 < Z && k + 1 < k1) { fetch_stage(k + 2); fetch_b(k + 2); }
        if (k + 1 < k1) fetch_own(k + 1);
        __syncthreads();   // slice k + 1 staged; L(k) complete
        if (nx) formL(k + 1, bn);
        __syncthreads();   // L(k + 1) complete
        if (!own) continue;
        // dual step, projection and momentum (FGP_operator_LT_Pp)
        const int64_t v = (int64_t)k * MN + (int64_t)j * M + i;
        const int c = (ty + 1) * FEX + (tx + 1), l = ty * FLX + tx;
        const int bf = k % 3;
        const T r0 = sr[bf][c], s0 = ss[bf][c], t0 = stt[bf][c];
        T tr = r0, ts = s0, tt = t0;
        const T Lv = sL[k & 1][l];
        if (i < M - 1) tr = tr + a12 * (Lv - sL[k & 1][l + 1]);
        if (j < N - 1) ts = ts + a12 * (Lv - sL[k & 1][l + FLX]);
        if (k < Z - 1) tt = tt + a12 * (Lv - sL[(k + 1) & 1][l]);
        const T nrm = sqrt(tr * tr + ts * ts + tt * tt);
        const T den = nrm > (T)1 ? nrm : (T)1;
        if (i < M - 1) { const T on = tr / den; std_(rB, v, on + g * (on - from_raw<T, D>(co))); std_(o, v, on); }
        else std_(rB, v, r0);
        if (j < N - 1) { const T pn = ts / den; std_(sB, v, pn + g * (pn - from_raw<T, D>(cp))); std_(p, v, pn); }
        else std_(sB, v, s0);
        if (k < Z - 1) { const T qn = tt / den; std_(tB, v, qn + g * (qn - from_raw<T, D>(cq))); std_(q, v, qn); }
        else std_(tB, v, t0);
    }
}

template <typename S, typename T, typename D = T>
static ct_status run_fgp(ct_context *ctx, int M, int N, int Z, void *vol, double lambda, int niter) {
    const int64_t V = (int64_t)M * N * Z;
    ct_status st;
    // (the step keeps in-slice offsets in 32 bits)
    if ((int64_t)M * N >= (1ll << 31)) return fail(ctx, CT_ERR_OVERFLOW, "fgp: slice of 2^31 voxels or more");
    // two sets of duals (read one, write the other) and the momentum copies
    D *w = (D *)slot(ctx, SLOT_FGP, sizeof(D) * V * 9, &st);
    if (!w) return st;
    D *dual[2][3] = {{w, w + V, w + 2 * V}, {w + 6 * V, w + 7 * V, w + 8 * V}};
    D *o = w + 3 * V, *p = w + 4 * V, *q = w + 5 * V;
    st = launch_fill(ctx, w, dtype_of<D>(), V * 9, 0.0);
    if (st != CT_OK) return st;
    dim3 b(256), gr((M + 255) / 256, N, Z);
    dim3 fb(FGX, FGY), fg((M + FGX - 1) / FGX, (N + FGY - 1) / FGY, (Z + FGZ - 1) / FGZ);
    const T lam = (T)lambda;
    const T a12 = (T)1 / (lam * (T)12);
    T tk1 = 1;
    for (int it = 0; it < niter; ++it) {
        const T tk = ((T)1 + std::sqrt((T)1 + (T)4 * tk1 * tk1)) / (T)2;
        const T g = (tk1 - (T)1) / tk;
        ProfScope ps(ctx, CT_K_FGP);
        D *const *A = dual[it & 1], *const *B = dual[(it + 1) & 1];
        hipLaunchKernelGGL((k_fgp_step<S, T, D>), fg, fb, 0, ctx->stream, (const S *)vol, (const D *)A[0],
                           (const D *)A[1], (const D *)A[2], B[0], B[1], B[2], o, p, q, M, N, Z, lam, g, a12);
        tk1 = tk;
    }
    hipLaunchKernelGGL((k_fgp_L<S, T, D>), gr, b, 0, ctx->stream, (const S *)vol, (const D *)o, (const D *)p,
                       (const D *)q, (S *)vol, (T *)nullptr, M, N, Z, lam);
    return launch_check(ctx, "fgp");
}

ct_status launch_fgp(ct_context *ctx, int M, int N, int Z, void *vol, ct_dtype dt, double lambda, int niter) {
    switch (dt) {
    case CT_F32: return run_fgp<float, float>(ctx, M, N, Z, vol, lambda, niter);
    // 16-bit volumes keep their dual fields in the same 16-bit type
    case CT_F16: return run_fgp<__half, float, __half>(ctx, M, N, Z, vol, lambda, niter);
    case CT_BF16: return run_fgp<__hip_bfloat16, float, __hip_bfloat16>(ctx, M, N, Z, vol, lambda, niter);
    case CT_F64: return run_fgp<double, double>(ctx, M, N, Z, vol, lambda, niter);
    }
    return fail(ctx, CT_ERR_DTYPE, "fgp: bad dtype");
}

// D: x -> forward differences (zero on the last plane); D^T: divergence form.
template <typename S>
__global__ void k_fgp_D(const S *x, S *r, S *s, S *t, int M, int N, int Z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= M) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t v = (int64_t)k * MN + (int64_t)j * M + i;
    const double c = ld64(x, v);
    double a = i < M - 1 ? c - ld64(x, v + 1) : 0.0;
    double bb = j < N - 1 ? c - ld64(x, v + M) : 0.0;
    double e = k < Z - 1 ? c - ld64(x, v + MN) : 0.0;
    if constexpr (std::is_same<S, double>::value) { r[v] = a; s[v] = bb; t[v] = e; }
    else { St<S>::st(r, v, (float)a); St<S>::st(s, v, (float)bb); St<S>::st(t, v, (float)e); }
}
template <typename S>
__global__ void k_fgp_Dt(const S *r, const S *s, const S *t, S *out, int M, int N, int Z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= M) return;
    const int64_t MN = (int64_t)M * N;
    const int64_t v = (int64_t)k * MN + (int64_t)j * M + i;
    double rv = i < M - 1 ? ld64(r, v) : 0.0, sv = j < N - 1 ? ld64(s, v) : 0.0, tv = k < Z - 1 ? ld64(t, v) : 0.0;
    double rm = i > 0 ? ld64(r, v - 1) : 0.0, sm = j > 0 ? ld64(s, v - M) : 0.0, tm = k > 0 ? ld64(t, v - MN) : 0.0;
    double o = (rv - rm) + (sv - sm) + (tv - tm);
    if constexpr (std::is_same<S, double>::value) out[v] = o; else St<S>::st(out, v, (float)o);
}

ct_status launch_fgp_D(ct_context *ctx, int M, int N, int Z, const void *x, void *r, void *s, void *t, ct_dtype dt) {
    dim3 b(256), gr((M + 255) / 256, N, Z);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_fgp_D<float>, gr, b, 0, ctx->stream, (const float *)x, (float *)r, (float *)s, (float *)t, M, N, Z); break;
    case CT_F16: hipLaunchKernelGGL(k_fgp_D<__half>, gr, b, 0, ctx->stream, (const __half *)x, (__half *)r, (__half *)s, (__half *)t, M, N, Z); break;
    case CT_BF16: hipLaunchKernelGGL(k_fgp_D<__hip_bfloat16>, gr, b, 0, ctx->stream, (const __hip_bfloat16 *)x, (__hip_bfloat16 *)r, (__hip_bfloat16 *)s, (__hip_bfloat16 *)t, M, N, Z); break;
    case CT_F64: hipLaunchKernelGGL(k_fgp_D<double>, gr, b, 0, ctx->stream, (const double *)x, (double *)r, (double *)s, (double *)t, M, N, Z); break;
    }
    return launch_check(ctx, "fgp D");
}
ct_status launch_fgp_Dt(ct_context *ctx, int M, int N, int Z, const void *r, const void *s, const void *t, void *out, ct_dtype dt) {
    dim3 b(256), gr((M + 255) / 256, N, Z);
    switch (dt) {
    case CT_F32: hipLaunchKernelGGL(k_fgp_Dt<float>, gr, b, 0, ctx->stream, (const float *)r, (const float *)s, (const float *)t, (float *)out, M, N, Z); break;
    case CT_F16: hipLaunchKernelGGL(k_fgp_Dt<__half>, gr, b, 0, ctx->stream, (const __half *)r, (const __half *)s, (const __half *)t, (__half *)out, M, N, Z); break;
    case CT_BF16: hipLaunchKernelGGL(k_fgp_Dt<__hip_bfloat16>, gr, b, 0, ctx->stream, (const __hip_bfloat16 *)r, (const __hip_bfloat16 *)s, (const __hip_bfloat16 *)t, (__hip_bfloat16 *)out, M, N, Z); break;
    case CT_F64: hipLaunchKernelGGL(k_fgp_Dt<double>, gr, b, 0, ctx->stream, (const double *)r, (const double *)s, (const double *)t, (double *)out, M, N, Z); break;
    }
    return launch_check(ctx, "fgp Dt");
}

} // namespace ct
