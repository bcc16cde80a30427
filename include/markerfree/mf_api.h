/*
 * C-ABI of the MI355X-native Markerfree alignment library
 * (libmarkerfree_hip.so).
 *
 * The reference exposes no plugin or FFI API: its boundary is the process
 * (Markerfree -i -o -a -g -p -s, main/main.cpp:6-44) and, inside it, the
 * method-layer C++ classes that Process::DoIt drives (main/Process.cpp:34-92).
 * Each entry point below replaces one of those seams; the comment on each
 * names the reference interface (file:line) it stands in for.
 *
 * Conventions: plain pointers and sizes only; host arrays are caller-owned and
 * copied in or out; "d_" pointers are device (HBM) pointers obtained from
 * mf_dev_malloc or any HIP allocation on the context's device.  No C++
 * exception crosses this boundary; failures return an mf_status and leave a
 * message in mf_last_error(ctx).  One mf_ctx per host thread.
 */
#ifndef MARKERFREE_MF_API_H
#define MARKERFREE_MF_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MF_OK = 0,
  MF_EINVAL = 1,  /* bad argument or shape */
  MF_ENOMEM = 2,  /* device allocation or HBM budget exceeded */
  MF_EDEVICE = 3, /* HIP runtime error */
  MF_EIO = 4,     /* file read/write error */
  MF_EBUSY = 5,   /* admission refused (serving) */
  MF_ESTATE = 6   /* call out of order (e.g. iteration before setup) */
} mf_status;

typedef struct mf_ctx mf_ctx;

/* ---------------------------------------------------------------- context */

/* Replaces cudaSetDevice(opts.GPU) (main/main.cpp:26).  hbm_budget_bytes = 0
 * means "whole device".  Creates the context's HIP stream. */
mf_status mf_ctx_create(int device, size_t hbm_budget_bytes, mf_ctx** out);
void mf_ctx_destroy(mf_ctx* ctx);
const char* mf_last_error(const mf_ctx* ctx);
mf_status mf_ctx_synchronize(mf_ctx* ctx);
/* Number of visible HIP devices (0 when none). */
int mf_device_count(void);
/* Library version string. */
const char* mf_version(void);

/* ---------------------------------------------------------- device memory */
/* Plumbing for callers that own device buffers (tests, benchmarks). */
mf_status mf_dev_malloc(mf_ctx* ctx, size_t bytes, void** d_out);
mf_status mf_dev_free(mf_ctx* ctx, void* d_ptr);
mf_status mf_h2d(mf_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
mf_status mf_d2h(mf_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
mf_status mf_d2d(mf_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---------------------------------------------------------- kernel timing */
/* When enabled, every launch of a named kernel is bracketed by HIP events on
 * the stream it runs on; mf_kernel_time sums the elapsed times (ms) and counts
 * launches.  Names: "bp_slab", "rp_slab", "bprp_fused", "fft_rows",
 * "fft_cols", "correct", "radon", "cc_peak". */
mf_status mf_timing_enable(mf_ctx* ctx, int enable);
mf_status mf_timing_reset(mf_ctx* ctx);
mf_status mf_kernel_time(mf_ctx* ctx, const char* name, double* total_ms,
                         long long* launches);

/* ------------------------------------------------- reprojection (PA core) */

/* CalcReprojFBP::FindProjRange (method/CalcReproj.cu:281-334): the contiguous
 * range [range2[0], range2[1]] of reference projections for 'target'
 * (range2[0] = -1 when none qualifies).  Host-only. */
mf_status mf_find_proj_range(const float* tilts, int nz,
                             const unsigned char* skip, int target, int nproj,
                             int* range2);

/* CalcReprojFBP::DoIt (method/CalcReproj.cu:255-279) for one target:
 * backproject the in-range rows of the corrected stack d_stack [nz][by][bx]
 * into per-row xz slabs of 2bx x tb voxels (mGBackProj, :26-64) and ray-sum
 * them at tilts[target] (mGForProj, :68-124) into d_reproj [by][bx].
 * mode 0: materialised slab (d_slab [by][tb][2bx] if non-NULL, else a
 * library workspace); mode 1: fused backprojection+reprojection that samples
 * the slab on the fly (identical output, no slab in HBM). */
mf_status mf_reproject(mf_ctx* ctx, const float* d_stack, int bx, int by,
                       int nz, int tb, const float* tilts,
                       const unsigned char* skip, int target, int nproj,
                       int mode, float* d_reproj, float* d_slab);

/* ------------------------------------------------------------------- FFT */
/* Replace cuFFT as wrapped by cufft2D / cufft1D (method/Cufft1&2.cu:68-206).
 * In-place, padded layout: real rows of n values at a pitch of (n/2+1)*2
 * floats, spectra of n/2+1 complex values in the same rows.  Unnormalised
 * like cufftExecR2C/C2R; 'scale' multiplies the forward result (1/(nx*ny)
 * reproduces cufft2D::ForwardFFT with bNorm, Cufft1&2.cu:88-103). */
mf_status mf_fft_r2c_2d(mf_ctx* ctx, float* d_buf, int nx, int ny, int nimg,
                        float scale);
mf_status mf_fft_c2r_2d(mf_ctx* ctx, float* d_buf, int nx, int ny, int nimg);
/* Batched 1-D transforms of length n on nlines lines 'pitch' floats apart
 * (cufft1D::Forward/Inverse, Cufft1&2.cu:181-206). */
mf_status mf_fft_r2c_1d(mf_ctx* ctx, float* d_buf, int n, int nlines,
                        int pitch);
mf_status mf_fft_c2r_1d(mf_ctx* ctx, float* d_buf, int n, int nlines,
                        int pitch);

#ifdef __cplusplus
}
#endif
#endif /* MARKERFREE_MF_API_H */
