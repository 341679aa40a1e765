/*
 * ct_core.h - C ABI of the MI355X (gfx950) cone-beam CT reconstruction core.
 *
 * Every entry point takes DEVICE pointers (HBM-resident buffers owned by the
 * caller), runs asynchronously on the context's HIP stream unless it returns a
 * host scalar, and returns a ct_status.  No entry point exits the process or
 * stages host memory implicitly; the file calls (projections in, volumes and
 * checkpoints out) move data through pinned host buffers of their own and
 * return when it is in place.  Host-array wrappers with the reference's own
 * function names are declared in ct_compat.h.
 *
 * Reference call surface replaced (paths relative to the reference root,
 * V63 = CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823,
 * TB  = TBCT/TBCT_1024V_1024P_v4.2_20171012):
 *   compile-time geometry            V63/main_recon.cu:54-111, TB/main_recon.cu:83-117
 *   Forward_3D_ray_driven_siddon     V63/host_IterativeRecon_CBCT.c:70
 *   Reconstruction_3D_ray_driven_CBCT V63/host_IterativeRecon_CBCT.c:1
 *   CheckMatchedJointOperator        V63/host_IterativeRecon_CBCT.c:131
 *   pixel-driven (matched) SART      V63/pixel_driven_backprj/host_IterativeRecon_CBCT.c:1
 *   FISTA momentum (inline)          V63/main_recon.cu:372-384
 *   FGP_denoise_GPUx7_exact          V63/host_FGP_Denoise_GPUx7.c:1
 *   Reconstruction_FDK               V63/host_FDK.cpp:4
 *   RampFiltering_GPU                V63/FDK/FFT_1d_GPU.h:241
 *   L2_norm_gpu / TV_norm_gpu        V63/host_tool_functions.cu:28,61
 *   tv_gradient_matrix_3d_kernel     V63/kernel_IterativeRecon_universal.cu:912
 *   Reconstruction_3D_ray_driven_TBCT TB/voxel/host_IterativeRecon_TBCT_voxel.c:1
 *   Forward_3D_ray_driven_siddon_TBCT TB/voxel/host_IterativeRecon_TBCT_voxel.c:111
 *   projection files / outputs       V63/main_recon.cu:199-224 / :437-487
 */
#ifndef CT_CORE_H
#define CT_CORE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_ABI_VERSION 1

typedef enum ct_status {
    CT_OK = 0,
    CT_ERR_INVALID_ARG = 1,   /* bad scalar argument (beta, niter, view range ...) */
    CT_ERR_GEOMETRY = 2,      /* geometry fails validation */
    CT_ERR_DTYPE = 3,         /* unsupported dtype or dtype combination */
    CT_ERR_NULL = 4,          /* null pointer */
    CT_ERR_ALIGN = 5,         /* pointer not aligned to its element size */
    CT_ERR_OVERFLOW = 6,      /* a size exceeds the kernels' 32-bit index range */
    CT_ERR_ALLOC = 7,         /* device scratch allocation failed */
    CT_ERR_DEVICE = 8,        /* no usable HIP device / HIP runtime error */
    CT_ERR_LAUNCH = 9,        /* kernel launch failed */
    CT_ERR_IO = 10            /* a file could not be opened, read or written, or is malformed */
} ct_status;

/* Storage types.  Arithmetic is FP32 for F32/F16/BF16 storage and FP64 for
 * F64 storage; ray sums and backprojection sums are always accumulated in the
 * arithmetic type, never in the storage type. */
typedef enum ct_dtype {
    CT_F32 = 0,
    CT_F16 = 1,
    CT_BF16 = 2,
    CT_F64 = 3
} ct_dtype;

/* Runtime replacement of the reference's compile-time constants
 * (V63/main_recon.cu:54-111).  Lengths are in any one unit (the reference uses
 * metres, the synthetic workloads millimetres); every threshold of the
 * algorithms scales with the voxel pitch, so the unit does not matter.
 *
 * Coordinates: volume index f[k*ny*nx + j*nx + i] (x fastest), voxel centre
 *   x_i = bx + (i+0.5)*vx with bx = -vx*(nx/2 + x_offset), by = -vy*ny/2,
 *   bz = -vz*nz/2.
 * Detector: ncols x nrows, stored [row][col]; column centre
 *   u_c = -du*(ncols/2 - 0.5) + det_offset + c*du, row centre
 *   v_r = -dv*(nrows/2 - 0.5) + r*dv.  The reference uses one pitch for both.
 * Source (view angle phi): (dso*cos - det_offset*sin, dso*sin + det_offset*cos, z_s)
 *   with z_s = src_z0 + s*src_dz for source s; detector pixel at
 *   (dod*cos - u*sin, dod*sin + u*cos, v).  dod < 0 puts the detector across
 *   the isocentre (SDD = dso - dod).
 * Angles: phi_j = -((float)(j*angle_step_deg + angle0_deg) + angle_shift_deg)*pi/180
 *   for the iterative operators (V63/host_IterativeRecon_CBCT.c:34-36) and +theta
 *   for FDK (V63/host_FDK.cpp:35-37).
 * nsources > 1 selects the multi-source (TB) variants of the SART kernels:
 *   projection stacks are [view][source][row][col].
 */
typedef struct ct_geometry {
    int32_t nx, ny, nz;           /* M, N, ZETA */
    int32_t x_offset;             /* M_Offset (half-detector use) */
    double vx, vy, vz;            /* volumn_x, volumn_y, volumn_z */
    int32_t ncols, nrows;         /* R, Z_prj */
    double du, dv;                /* Detector_pixel_x (columns, rows) */
    double det_offset;            /* Offset (= Source_y) */
    double dso, dod;              /* DSO, DOD */
    int32_t nviews;               /* Nviews */
    int32_t nsources;             /* 1 for CBCT; N_source for TB */
    double angle_step_deg;        /* us_rate */
    double angle0_deg;            /* initialAngle */
    double angle_shift_deg;       /* shiftAngle */
    double src_z0, src_dz;        /* Source_z (CBCT) / Source_z_min, Source_interval (TB) */
    int32_t style;                /* CT_style (V63/main_recon.cu:28): a ct_style value */
    int32_t reserved;             /* 0 */
} ct_geometry;

/* Beam geometry (CT_style, V63/kernel_IterativeRecon_universal.cu:22-39,
 * 656-670, 709-724).  CT_FAN: multi-slice fan beam, every detector row has
 * its own source at the row's height, so rays run in z-planes.  CT_PARALLEL:
 * the source moves with the detector pixel, at (dso*cos - u*sin,
 * dso*sin + u*cos, v), so all rays of a view are parallel.  In the voxel
 * backprojection both take the voxel's own height for the source and project
 * without magnification along z (CT_PARALLEL also without it along u).
 * Multi-source (nsources > 1) and FDK are cone-beam only. */
typedef enum ct_style {
    CT_CONE = 0,
    CT_FAN = 1,
    CT_PARALLEL = 2
} ct_style;

/* Ramp-filter window for FDK. */
typedef enum ct_window {
    CT_WINDOW_RAMP = 0,   /* pure ramp x 0.5, as V63/FDK/FFT_1d_GPU.h:256-267 */
    CT_WINDOW_HANN = 1    /* ramp x 0.5 x 0.5*(1+cos(2*pi*k/pad)), k in [-pad/2, pad/2) */
} ct_window;

/* View order inside a SART sweep. */
typedef enum ct_order {
    CT_ORDER_ASCENDING = 0,   /* subset s = views {s*S, ..., s*S+S-1}: S=1 is the reference order */
    CT_ORDER_INTERLEAVED = 1  /* subset s = views {s, s+n_sub, s+2*n_sub, ...} */
} ct_order;

typedef struct ct_context ct_context;

/* ---- context --------------------------------------------------------- */
int         ct_abi_version(void);
ct_status   ct_create(ct_context **ctx, int device);
ct_status   ct_destroy(ct_context *ctx);
/* Use a caller-owned hipStream_t (NULL = the legacy default stream). */
ct_status   ct_set_stream(ct_context *ctx, void *hip_stream);
ct_status   ct_synchronize(ct_context *ctx);
const char *ct_last_error(const ct_context *ctx);
const char *ct_status_string(ct_status s);
/* Validates a geometry without touching the GPU. */
ct_status   ct_validate_geometry(const ct_geometry *g);
/* Per-view angle (radians) exactly as the kernels use it; fdk selects +theta. */
ct_status   ct_view_angle(const ct_geometry *g, int32_t view, int32_t fdk,
                          float *sin_out, float *cos_out);
/* Bytes of device scratch the context currently holds. */
size_t      ct_scratch_bytes(const ct_context *ctx);

/* ---- projector pair ---------------------------------------------------- */
/* Siddon ray sums (command 0 of forward_ray_driven_3d_kernel_correction,
 * V63/kernel_IterativeRecon_universal.cu:7-567) for views [view0, view0+nviews):
 * proj[(v-view0)][src][row][col]. */
ct_status ct_forward(ct_context *ctx, const ct_geometry *g,
                     const void *vol, ct_dtype vol_dtype,
                     void *proj, ct_dtype proj_dtype,
                     int32_t view0, int32_t nviews);

/* No counterpart in the reference: ct_forward (CMD 0 ray sums) for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or allocating library routine: its one scratch buffer, the per-view
 * angle table, is a static 256 KiB device array per GPU, so g->nviews is
 * limited to 16384 (larger geometries get CT_ERR_INVALID_ARG).  It always
 * takes the per-ray walk on the volume as given (no z-quad copy); results
 * agree with ct_forward to FP32 rounding (identical for fewer than four
 * images).  The call first synchronizes its stream, then overwrites the
 * table: calls that overlap on one GPU from different streams or contexts
 * race on it. */
ct_status ct_forward_noalloc(ct_context *ctx, const ct_geometry *g,
                             const void *vol, ct_dtype vol_dtype,
                             void *proj, ct_dtype proj_dtype,
                             int32_t view0, int32_t nviews);

/* SART correction (command 1, same kernel): corr = (b - sum l*f)/sum l, 0 where
 * sum l < 1e-6*vz.  meas holds the views' measured projections; corr is F32
 * (F64 when vol_dtype is F64). */
ct_status ct_sart_correction(ct_context *ctx, const ct_geometry *g,
                             const void *vol, ct_dtype vol_dtype,
                             const void *meas, ct_dtype meas_dtype,
                             void *corr,
                             int32_t view0, int32_t nviews);

/* Voxel-driven backprojection (backprj_ray_driven_3d_kernel,
 * V63/kernel_IterativeRecon_universal.cu:571-810), summed over the nviews
 * correction images: mode 1: x += beta*W/L where L >= 1e-6*vx (SART update);
 * mode 0: x = W where L >= 1e-6*vx (unchanged elsewhere).  In fast mode a
 * CT_F16/CT_BF16 volume stores the mode-1 update with stochastic rounding
 * (a dither taken from W and L, so runs are reproducible): small updates are
 * kept in expectation instead of being rounded away. */
ct_status ct_backproject_voxel(ct_context *ctx, const ct_geometry *g,
                               void *vol, ct_dtype vol_dtype,
                               const void *corr,
                               int32_t view0, int32_t nviews,
                               double beta, int32_t mode);

/* One SART sweep over all views, device resident (Reconstruction_3D_ray_driven_CBCT,
 * V63/host_IterativeRecon_CBCT.c:1-67).  subset_size = 1 with CT_ORDER_ASCENDING
 * reproduces the reference's view-by-view order; larger subsets accumulate W and
 * L over the subset's views before one update. */
ct_status ct_sart_sweep(ct_context *ctx, const ct_geometry *g,
                        void *vol, ct_dtype vol_dtype,
                        const void *meas, ct_dtype meas_dtype,
                        double beta, int32_t subset_size, ct_order order);

/* Matched adjoint: vol = A^T proj, the exact transpose of ct_forward's Siddon
 * segment lengths (scatter form of V63/pixel_driven_backprj/kernel_IterativeRecon_CBCT.cu:563-1116,
 * with the forward's 1e-16 loop tolerance and ray-length cut so that the pair
 * is transposed exactly).  Overwrites vol. */
ct_status ct_adjoint(ct_context *ctx, const ct_geometry *g,
                     const void *proj, ct_dtype proj_dtype,
                     void *vol, ct_dtype vol_dtype,
                     int32_t view0, int32_t nviews);

/* <A x, y> and <x, A^T y> in FP64 for seeded x, y ~ N(0,1) over all views
 * (CheckMatchedJointOperator, V63/host_IterativeRecon_CBCT.c:131-204). */
ct_status ct_adjoint_test(ct_context *ctx, const ct_geometry *g, ct_dtype dtype,
                          uint64_t seed, double *lhs, double *rhs);

/* Largest eigenvalue of A^T A by niter power iterations (FP64 norms). */
ct_status ct_power_iteration(ct_context *ctx, const ct_geometry *g, ct_dtype dtype,
                             int32_t niter, uint64_t seed, double *lipschitz);

/* Lipschitz estimate of the SART-FISTA step: 2 x the largest eigenvalue of
 * the operator that one SART sweep with subsets of `subset` consecutive
 * views applies to first order in its step,
 *     K = sum over subsets S of D_S^-1 B_S^T M_S^-1 A_S
 * (A_S the subset's forward projection, M_S its rays' lengths, B_S^T the
 * voxel-driven backprojection, D_S its per-voxel chord sums), by niter power
 * iterations with FP64 norms; dtype is the arithmetic of the projections
 * (F32 or F64).  K has non-negative entries, so its largest eigenvalue is its
 * spectral radius, with a non-negative eigenvector.  The step
 * beta = 2 / lipschitz = 1 / rho(K) keeps a sweep inside FISTA's 1/L bound on
 * the gradient step, and the FGP argument is 2 lambda_TV / lipschitz; they
 * replace the constant Lip_con of V63/main_recon.cu:58,161-162.  subset =
 * nviews gives the simultaneous operator (radius at most 1); view by view
 * (subset 1) it grows with the number of views. */
ct_status ct_sart_lipschitz(ct_context *ctx, const ct_geometry *g, ct_dtype dtype, int32_t subset,
                            int32_t niter, uint64_t seed, double *lipschitz);

/* ---- FISTA / TV ------------------------------------------------------- */
/* y = x + gamma*(x - x_prev); x_prev = x (V63/main_recon.cu:375-382).  y may
 * alias x, which fuses the reference's F_recon <- F_Y_k copy away. */
ct_status ct_fista_momentum(ct_context *ctx, void *x, void *x_prev, void *y,
                            int64_t n, ct_dtype dtype, double gamma);

/* FISTA scalar t_k = (1 + sqrt(1 + 4 t_{k-1}^2))/2 and gamma = (t_{k-1}-1)/t_k in FP64. */
ct_status ct_fista_gamma(double t_prev, double *t_next, double *gamma);

/* TV-FGP denoise in place (FGP_denoise_GPUx7_exact, V63/host_FGP_Denoise_GPUx7.c:1-111
 * and kernels V63/kernel_FGP_Denoise_GPUx7.cu:2-88): x* = argmin ||x-b||^2 + 2*lambda*TV(x).
 * lambda == 0 leaves vol unchanged. */
ct_status ct_tv_fgp(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz,
                    void *vol, ct_dtype dtype, double lambda, int32_t niter);

/* The linear pair inside FGP: D maps a volume to its three forward differences
 * (r,s,t)_v = (x_v - x_{v+e}) where v+e is inside, else 0 (the L^T operator,
 * V63/kernel_FGP_Denoise_GPUx7.cu:58-65); D^T maps (r,s,t) back:
 * (D^T p)_v = p_v - p_{v-e} summed over axes (the L operator, :14).  Dual
 * fields are volume-sized arrays, one per axis. */
ct_status ct_tv_fgp_adjoint(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz,
                            const void *p_r, const void *p_s, const void *p_t,
                            void *out, ct_dtype dtype);
ct_status ct_tv_fgp_forward_op(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz,
                               const void *x, void *p_r, void *p_s, void *p_t,
                               ct_dtype dtype);

/* TV_eps(x) = sum_v sqrt(eps + sum_d (x_v - x_{v-e_d})^2) over the backward
 * differences that exist inside the volume (the definition of
 * tv_gradient_matrix_3d_kernel, V63/kernel_IterativeRecon_universal.cu:912-949,
 * extended to the boundary with missing neighbours dropped).  FP64 result. */
ct_status ct_tv_eps(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz,
                    const void *vol, ct_dtype dtype, double eps, double *value);
ct_status ct_tv_eps_grad(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz,
                         const void *vol, ct_dtype dtype, double eps, void *grad);

/* ---- norms (FP64 on the GPU) ------------------------------------------ */
/* sum (a-b)^2 (L2_norm_gpu, V63/host_tool_functions.cu:28-58, without the
 * overwrite of a). */
ct_status ct_residual_norm2(ct_context *ctx, const void *a, const void *b,
                            int64_t n, ct_dtype dtype, double *value);
/* Isotropic forward-difference TV (TV_norm_kernel, V63/kernel_tool_functions.cu:39-65). */
ct_status ct_tv_norm(ct_context *ctx, int32_t nx, int32_t ny, int32_t nz,
                     const void *vol, ct_dtype dtype, double *value);
/* sum a*b in FP64. */
ct_status ct_dot(ct_context *ctx, const void *a, const void *b, int64_t n,
                 ct_dtype dtype, double *value);

/* Objective of V63/main_recon.cu:401-428: data = sum (Ax - b)^2 over all views
 * (the residual is fused into the ray traversal; Ax is never stored), tv =
 * ct_tv_norm(x), objective = data + 2*lambda_tv*tv.  All FP64. */
ct_status ct_objective(ct_context *ctx, const ct_geometry *g,
                       const void *vol, ct_dtype vol_dtype,
                       const void *meas, ct_dtype meas_dtype,
                       double lambda_tv, double *data_fidelity, double *tv, double *objective);

/* ---- FDK -------------------------------------------------------------- */
/* Cosine weight (kernel_FDK_projection_adjust, V63/FDK/kernel_FDK.cu:1-14) of
 * all views, out-of-place allowed in == out. */
ct_status ct_fdk_weight(ct_context *ctx, const ct_geometry *g,
                        const void *proj_in, void *proj_out, ct_dtype dtype);
/* Row ramp filter by a zero-padded 2*ncols FFT (RampFiltering_GPU,
 * V63/FDK/FFT_1d_GPU.h:241-287) of nimg images of nrows rows. */
ct_status ct_ramp_filter(ct_context *ctx, const ct_geometry *g,
                         const void *in, void *out, ct_dtype dtype,
                         int32_t nimg, ct_window window);
/* Weighted bilinear backprojection of filtered views [view0, view0+nviews)
 * accumulated into vol (kernel_FDK_reconstruction, V63/FDK/kernel_FDK.cu:16-62).
 * A multi-source geometry takes the CT_FDK_INVERSE_SQUARE combination. */
ct_status ct_fdk_backproject(ct_context *ctx, const ct_geometry *g,
                             const void *filtered, ct_dtype proj_dtype,
                             void *vol, ct_dtype vol_dtype,
                             int32_t view0, int32_t nviews);
/* Whole FDK (Reconstruction_FDK, V63/host_FDK.cpp:4-57): vol += FDK(proj).
 * Scratch for the filtered stack is held by the context.  Cone beam only. */
ct_status ct_fdk(ct_context *ctx, const ct_geometry *g,
                 const void *proj, ct_dtype proj_dtype,
                 void *vol, ct_dtype vol_dtype, ct_window window);

/* Multi-source FDK (TB/FDK/host_FDK.cpp:4-83): the stack holds
 * [view][source][row][col]; each image is weighted (rows measured from its
 * source's height) and filtered, and per view and voxel the sources combine
 * as the reference's three variants do. */
typedef enum ct_fdk_sources {
    CT_FDK_INVERSE_SQUARE = 0,  /* sources on the detector weighted by 1/(z - z_s)^2, normalised (TB/FDK/kernel_FDK.cu:37-99) */
    CT_FDK_NEAREST = 1,         /* the source nearest in z whose row is on the detector (TB/FDK/OneRay/kernel_FDK.cu:37-113) */
    CT_FDK_EQUAL = 2            /* the sum over sources (TB/FDK/EqualWeighting/kernel_FDK.cu:37-91) */
} ct_fdk_sources;
ct_status ct_fdk_backproject_ms(ct_context *ctx, const ct_geometry *g,
                                const void *filtered, ct_dtype proj_dtype,
                                void *vol, ct_dtype vol_dtype,
                                int32_t view0, int32_t nviews, ct_fdk_sources sources);
ct_status ct_fdk_ms(ct_context *ctx, const ct_geometry *g,
                    const void *proj, ct_dtype proj_dtype,
                    void *vol, ct_dtype vol_dtype, ct_window window, ct_fdk_sources sources);

/* ---- driver ----------------------------------------------------------- */
/* One outer iteration of SART + TV-FGP + FISTA momentum (V63/main_recon.cu:344-385),
 * device resident.  On entry x holds y_k (the extrapolated point, initially 0)
 * and x_prev holds x_{k-1}; on return x holds y_{k+1} and x_prev holds x_k.
 * *t_state carries t_{k-1} (1.0 before the first iteration).  fgp_lambda <= 0
 * skips the TV step. */
ct_status ct_sart_fista_iteration(ct_context *ctx, const ct_geometry *g,
                                  void *x, void *x_prev, ct_dtype vol_dtype,
                                  const void *meas, ct_dtype meas_dtype,
                                  double beta, int32_t subset_size, ct_order order,
                                  double fgp_lambda, int32_t fgp_niter,
                                  double *t_state);

/* ---- arithmetic mode ---------------------------------------------------- */
/* CT_MATH_EXACT keeps the reference's FP32 operation sequence (IEEE division
 * and square root, no contraction): results are bit-identical to the CPU
 * restatement.  CT_MATH_FAST (the default) lets the FP32 voxel-driven
 * backprojector use the hardware reciprocal / square root and fused
 * multiply-adds; it stays within the parity tolerances (1e-5 relative L2 per
 * update, 1e-4 on reconstructions).  FP64 paths are always exact. */
typedef enum ct_math {
    CT_MATH_EXACT = 0,
    CT_MATH_FAST = 1
} ct_math;
ct_status ct_set_math(ct_context *ctx, ct_math mode);

/* ---- kernel timing ------------------------------------------------------ */
/* HIP-event timing of every launch of the hot kernels, on the context's
 * stream (the reference brackets its host calls with cudaEvent timers,
 * V63/host_IterativeRecon_CBCT.c:5-8,55-61).  While enabled, each launch of a
 * profiled kernel is bracketed by an event pair; ct_profile_read synchronizes
 * and returns the summed milliseconds and launch count of one kernel class. */
typedef enum ct_kernel_id {
    CT_K_FORWARD = 0,       /* Siddon ray sum / SART correction */
    CT_K_VOXEL_BP = 1,      /* voxel-driven backprojection + update */
    CT_K_ADJOINT = 2,       /* matched adjoint scatter */
    CT_K_FGP = 3,           /* one TV-FGP inner iteration (one fused launch) */
    CT_K_FISTA = 4,         /* FISTA momentum */
    CT_K_FDK_BP = 5,        /* FDK backprojection */
    CT_K_RAMP = 6,          /* ramp filter */
    CT_K_COUNT = 7
} ct_kernel_id;
ct_status ct_profile_enable(ct_context *ctx, int32_t enable);
ct_status ct_profile_reset(ct_context *ctx);
ct_status ct_profile_read(ct_context *ctx, ct_kernel_id kernel, double *total_ms, int64_t *launches);

/* ---- files: projections in, volumes / objective / checkpoints out ------- */
/* These calls move data between files and device buffers: they stage through
 * pinned host memory and return once the data is in place (synchronous). */

/* Path of view `view`'s raw float32 projection file: directory +
 * "Projections/View_%.02f.dat" with the reference's file angle
 * float(view*angle_step + angle0), the second half of the views at
 * 180 + (view - nviews/2)*angle_step + angle0 when tomo != 0, and +360 for a
 * negative angle (commented-out reader, V63/main_recon.cu:199-224).  No
 * context; fails with CT_ERR_INVALID_ARG if `size` bytes do not hold it. */
ct_status ct_projection_path(const ct_geometry *g, int32_t view, int32_t tomo,
                             const char *directory, char *buf, int32_t size);

/* Reads views [view0, view0+nviews) from their files into the device stack
 * proj[view-view0][row][col] of storage type dtype (F32/F16/BF16).  Each file
 * holds float32 rows of ncols values; the first row0 rows are skipped (the
 * reference's fseek, V63/main_recon.cu:219) and nrows rows are read.  A
 * missing or short file fails with CT_ERR_IO naming the file. */
ct_status ct_read_projections(ct_context *ctx, const ct_geometry *g, const char *directory,
                              int32_t tomo, int32_t row0, void *proj, ct_dtype dtype,
                              int32_t view0, int32_t nviews);

/* Line integrals from raw detector counts, p = -ln(max((I - dark)/(flat - dark), floor))
 * for nimg images of `pixels` pixels; flat and dark are one FP32 image of
 * `pixels` each, flat NULL means the scalar i0 and dark NULL means 0; floor
 * lies in [FLT_MIN, 1].  A preprocessing step the reference leaves to the
 * caller. */
ct_status ct_log_transform(ct_context *ctx, const void *raw, ct_dtype raw_dtype,
                           const float *flat, const float *dark, double i0,
                           void *out, ct_dtype out_dtype, int64_t pixels, int32_t nimg,
                           double floor_value);

/* The volume as a raw float32 [z][y][x] file, the ".recon" format
 * (V63/main_recon.cu:437-463; SaveDeviceDataToFile, V63/host_tool_functions.cu:2-15),
 * from any storage type; and the reverse, e.g. to load an initial guess
 * (V63/main_recon.cu:252-266). */
ct_status ct_write_volume(ct_context *ctx, const ct_geometry *g, const void *vol,
                          ct_dtype dtype, const char *path);
ct_status ct_read_volume(ct_context *ctx, const ct_geometry *g, const char *path,
                         void *vol, ct_dtype dtype);

/* The reference's output names (V63/main_recon.cu:438,452,468): directory +
 * "<folder>/<object>_<nx>_<ny>_<nz>_<vx>um_new_view_<nviews>_(<a0>,<a1>)_TV_.._L1_.._Lip_...recon"
 * (iteration < 0) or "..._iterative_<k>_view_..." (iteration >= 0), and
 * "<folder>/object_func_<object>_view_<nviews>_(<a0>,<a1>)_TV_.._Lip_...bin".
 * unit_um converts the geometry's length unit to micrometres (1e6 for metres,
 * 1e3 for millimetres); a1 = a0 + (nviews-1)*angle_step. */
ct_status ct_recon_path(const ct_geometry *g, double unit_um, const char *directory,
                        const char *output_folder, const char *object_name, int32_t iteration,
                        double lambda_tv, double lambda_l1, double lipschitz, char *buf, int32_t size);
ct_status ct_objective_path(const ct_geometry *g, const char *directory, const char *output_folder,
                            const char *object_name, double lambda_tv, double lipschitz,
                            char *buf, int32_t size);
/* Writes `count` doubles, the (tv, data fidelity, objective) triples of the
 * iterations so far, as the reference rewrites its object_func file each
 * iteration (V63/main_recon.cu:467-477).  Host memory; no context. */
ct_status ct_write_objective(const char *path, const double *values, int64_t count);

/* FISTA state for exact resume, which the reference does not keep: x (the
 * extrapolated point y_k), x_prev (x_{k-1}), t_state (t_{k-1}) and the
 * iteration count, with the geometry and a checksum.  Continuing from a
 * loaded checkpoint with ct_sart_fista_iteration gives the same bits as the
 * uninterrupted run.  The file is written to path + ".part" and renamed into
 * place.  Loading checks the format, storage type (CT_ERR_DTYPE), geometry
 * (CT_ERR_GEOMETRY) and checksum (CT_ERR_IO). */
ct_status ct_save_checkpoint(ct_context *ctx, const ct_geometry *g, const char *path,
                             const void *x, const void *x_prev, ct_dtype dtype,
                             double t_state, int32_t iteration);
ct_status ct_load_checkpoint(ct_context *ctx, const ct_geometry *g, const char *path,
                             void *x, void *x_prev, ct_dtype dtype,
                             double *t_state, int32_t *iteration);

/* ---- utilities -------------------------------------------------------- */
/* Element conversion between storage types (n elements, device). */
ct_status ct_convert(ct_context *ctx, const void *src, ct_dtype src_dtype,
                     void *dst, ct_dtype dst_dtype, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* CT_CORE_H */
