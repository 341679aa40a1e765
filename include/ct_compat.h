/*
 * ct_compat.h - the reference's host call surface, re-exported with the same
 * names and signatures on FP32 arrays, so code written against the reference
 * links against libctcore.so unchanged.
 *
 * The reference fixes its geometry at compile time; here it is set once with
 * ct_compat_set_geometry() (device 0).  Array arguments may be host or device
 * pointers (detected per call); host arrays are staged through device buffers
 * held by the compatibility context.  The functions return nothing, like the
 * reference; failures are reported by ct_compat_last_status() /
 * ct_compat_last_error() and never exit the process.
 *
 * Paths are relative to the reference root, V63 = CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823,
 * TB = TBCT/TBCT_1024V_1024P_v4.2_20171012.
 */
#ifndef CT_COMPAT_H
#define CT_COMPAT_H

#include "ct_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces the compile-time constants of V63/main_recon.cu:54-111. */
int         ct_compat_set_geometry(const ct_geometry *g);
int         ct_compat_last_status(void);
const char *ct_compat_last_error(void);
/* Frees the compatibility context and its staging buffers. */
void        ct_compat_release(void);

/* V63/host_IterativeRecon_CBCT.c:70 - all views; h_proj_data may be host or device. */
void Forward_3D_ray_driven_siddon(float *h_volume, float *h_proj_data);
/* V63/host_IterativeRecon_CBCT.c:1 - one SART sweep, view by view, volume updated in place. */
void Reconstruction_3D_ray_driven_CBCT(float *h_volume, float *h_proj_data, float beta_temp);
/* V63/host_IterativeRecon_CBCT.c:131 - prints <A0 x, A0 x> and <x, B0 A0 x> for view 0,
 * B0 the voxel-driven backprojection with command 0 into a zeroed volume, as
 * the reference runs it: the gap shows how far that backprojector is from
 * the transpose. */
void CheckMatchedJointOperator(float *h_volume);
/* No counterpart in the reference: the same check with the matched adjoint
 * (ct_adjoint) as B0, whose two inner products agree to rounding. */
void ct_compat_check_matched_adjoint(float *h_volume);
/* The two inner products of the last check (either function above). */
int  ct_compat_last_inner_products(double *ip1, double *ip2);
/* V63/host_FGP_Denoise_GPUx7.c:1 - in place. */
void FGP_denoise_GPUx7_exact(float *h_f_noise, float lambda, int Niter_denoise);
/* V63/host_FDK.cpp:4 - h_volume += FDK(h_proj_data) (caller zeroes h_volume). */
void Reconstruction_FDK(float *h_volume, float *h_proj_data);
/* V63/FDK/FFT_1d_GPU.h:241 - ramp filter of all Nviews images (may alias). */
void RampFiltering_GPU(float *g_in, float *g_out);
/* V63/host_tool_functions.cu:28 - overwrites d_a with (d_a - d_b)^2 like the
 * reference and returns the sum (accumulated in FP64). */
double L2_norm_gpu(float *d_a, float *d_b);
/* V63/host_tool_functions.cu:61 */
double TV_norm_gpu(float *d_volume);
/* V63/host_tool_functions.cu:2 - writes data_size floats (host or device)
 * to a raw file; a failure sets the compat status instead of exiting. */
void SaveDeviceDataToFile(float *d_data, int data_size, char filename[]);
/* V63/host_tool_functions.cu:18 (host arithmetic, FP64 sum) */
double L2_norm_cpu(const float *a, const float *b, int SIZE);
/* V63/host_tool_functions.cu:105 (host arithmetic, FP64 sum) */
double TV_norm_cpu(float *d_volume);
/* V63/host_FGP_Denoise_CPU.h:242 - runs the GPU FGP (GPUx7 semantics) on Anoise into Aclean. */
void F_Dnoise(float *Anoise, float *Aclean, float lambda, int N);
/* TB/voxel/host_IterativeRecon_TBCT_voxel.c:1 - multi-source SART sweep (nsources > 1). */
void Reconstruction_3D_ray_driven_TBCT(float *h_volume, float *h_proj_data, float beta_temp);
/* TB/voxel/host_IterativeRecon_TBCT_voxel.c:111 - multi-source forward. */
void Forward_3D_ray_driven_siddon_TBCT(float *h_volume, float *h_proj_forward);

#ifdef __cplusplus
}
#endif

#endif /* CT_COMPAT_H */
