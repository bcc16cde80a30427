/*
 * C-ABI of the MI355X-native Markerfree alignment library
 * (libmarkerfree_hip.so).
 *
 * The reference exposes no plugin or FFI API: its boundary is the process
 * (Markerfree -i -o -a -g -p -s, main/main.cpp:6-43) and, inside it, the
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
/* Device bytes the context and its series hold now, and the most they have
 * held at once (since creation or the last reset_peak). */
mf_status mf_ctx_memory(mf_ctx* ctx, size_t* in_use, size_t* peak,
                        int reset_peak);
mf_status mf_dev_free(mf_ctx* ctx, void* d_ptr);
/* Page-locked host memory (hipHostMalloc): stacks uploaded from it and
 * aligned stacks downloaded into it (mf_series_create, mf_correct,
 * mf_fetch) move by DMA at the full PCIe rate instead of through staging
 * copies.  Independent of any context.  MF_ENOMEM when the allocation fails;
 * mf_host_free(NULL) is a no-op.  No counterpart in the reference (it reads
 * and writes files). */
mf_status mf_host_alloc(size_t bytes, void** h_out);
mf_status mf_host_free(void* h_ptr);
mf_status mf_h2d(mf_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
mf_status mf_d2h(mf_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
mf_status mf_d2d(mf_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---------------------------------------------------------- kernel timing */
/* When enabled, every launch of a named kernel is bracketed by HIP events on
 * the stream it runs on; mf_kernel_time sums the elapsed times (ms) and counts
 * launches.  Names: "bp_slab", "rp_slab", "bprp_fused", "fft_rows",
 * "fft_cols", "correct", "rweight", "correct_rweight", "radon", "cc" (the
 * launches of one correlation).  mf_timing_select restricts timing to the
 * launches of one name (NULL or "": all), so that timing one kernel adds no
 * events around the others; mf_timing_every(ctx, k) times only every k-th
 * launch of each name (each timed launch costs two stream markers, a few
 * microseconds of idle device), and mf_kernel_count returns all launches of
 * a name since the last reset, timed or not.  No counterpart in the
 * reference. */
mf_status mf_timing_enable(mf_ctx* ctx, int enable);
/* Launch-shape tunables (workgroup size, lines per workgroup) keyed by kernel
 * family and shape, e.g. "cc_rows/1024x1024/threads"; a key that is not set
 * takes the built-in heuristic.  Written by the autotuner
 * (markerfree_amd.autotune) and read back from its JSON cache.  No
 * counterpart in the reference.  mf_tune_set with value 0 removes the key;
 * mf_tune_get returns MF_EINVAL for a key that is not set. */
mf_status mf_tune_set(mf_ctx* ctx, const char* key, int value);
mf_status mf_tune_get(mf_ctx* ctx, const char* key, int* value);
/* Sets every tunable of a JSON object {"key": positive integer, ...} read
 * from 'path' (the autotuner's cache); *count receives how many were set.
 * MF_EIO if the file cannot be read, MF_EINVAL if it is not such an
 * object. */
mf_status mf_tune_load(mf_ctx* ctx, const char* path, int* count);
mf_status mf_timing_select(mf_ctx* ctx, const char* name);
mf_status mf_timing_every(mf_ctx* ctx, int k);
mf_status mf_kernel_count(mf_ctx* ctx, const char* name, long long* launches);
mf_status mf_timing_reset(mf_ctx* ctx);
mf_status mf_kernel_time(mf_ctx* ctx, const char* name, double* total_ms,
                         long long* launches);

/* ------------------------------------------------- reprojection (PA core) */

/* CalcReprojFBP::FindProjRange (method/CalcReproj.cu:206-259): the contiguous
 * range [range2[0], range2[1]] of reference projections for 'target'
 * (range2[0] = -1 when none qualifies).  Host-only. */
mf_status mf_find_proj_range(const float* tilts, int nz,
                             const unsigned char* skip, int target, int nproj,
                             int* range2);

/* CalcReprojFBP::DoIt (method/CalcReproj.cu:180-204) for one target:
 * backproject the in-range rows of the corrected stack d_stack [nz][by][bx]
 * into per-row xz slabs of 2bx x tb voxels (mGBackProj, :26-64) and ray-sum
 * them at tilts[target] (mGForProj, :68-124) into d_reproj [by][bx].
 * mode 0: materialised slab (d_slab [by][tb][2bx] if non-NULL, else a
 * library workspace); mode 1: fused backprojection+reprojection that samples
 * the slab on the fly (identical output, no slab in HBM), with as many image
 * rows per work unit (4, 2 or 1) as the sinogram rows allow in LDS; mode 2:
 * the one-row fused form without per-target tables (identical output; the
 * last fallback of mode 1); modes 3, 4 and 5 pin forms of mode 1 (one
 * row with tables, two rows with tables, two rows without; identical
 * output); modes 6 and 7: materialised slab of R = 4 or 2 image rows per
 * voxel, d_slab [cdiv(by, R)][tb][2bx][R] (image row R*g + r of group g at
 * position r; the same voxel values as mode 0; MF_EINVAL when the R rows'
 * sinogram rows do not fit in LDS); mode 8: mode 1 where, in the part of
 * each ray whose samples all have every neighbour in range (one weight sum
 * W), the ray adds the samples' weighted sums and divides by W once instead
 * of dividing every sample (a different rounding from the reference's
 * per-voxel division; within 1e-5 relative of the FP64 evaluation). */
mf_status mf_reproject(mf_ctx* ctx, const float* d_stack, int bx, int by,
                       int nz, int tb, const float* tilts,
                       const unsigned char* skip, int target, int nproj,
                       int mode, float* d_reproj, float* d_slab);

/* ------------------------------------------------------------------- FFT */
/* Replace cuFFT as wrapped by cufft2D / cufft1D (method/Cufft1&2.cu:17-165).
 * In-place, padded layout: real rows of n values at a pitch of (n/2+1)*2
 * floats, spectra of n/2+1 complex values in the same rows.  Unnormalised
 * like cufftExecR2C/C2R; 'scale' multiplies the forward result (1/(nx*ny)
 * reproduces cufft2D::ForwardFFT with bNorm, Cufft1&2.cu:47-62). */
mf_status mf_fft_r2c_2d(mf_ctx* ctx, float* d_buf, int nx, int ny, int nimg,
                        float scale);
mf_status mf_fft_c2r_2d(mf_ctx* ctx, float* d_buf, int nx, int ny, int nimg);
/* Batched 1-D transforms of length n on nlines lines 'pitch' floats apart
 * (cufft1D::Forward/Inverse, Cufft1&2.cu:140-165). */
mf_status mf_fft_r2c_1d(mf_ctx* ctx, float* d_buf, int n, int nlines,
                        int pitch);
mf_status mf_fft_c2r_1d(mf_ctx* ctx, float* d_buf, int n, int nlines,
                        int pitch);

/* ----------------------------------------------------- stage-level kernels */

/* CorrTomoStack::mCorrectProj (method/CorrStack.cu:370-397) on one raw
 * projection d_raw [ny][nx]: rotate by axis_deg (0 when shift_only) and
 * shift (nearest pixel), random fill or invalid outside, bin x bin sum
 * binning, optional R-weighting; d_out receives the unpadded [by][bx] image
 * with bx = nx/bin/2*2, by = ny/bin/2*2. */
mf_status mf_correct_proj(mf_ctx* ctx, const float* d_raw, int nx, int ny,
                          float shift_x, float shift_y, float axis_deg,
                          int shift_only, int random_fill, int bin,
                          int rweight, float* d_out);

/* CalcCC::DoIt + getshift (method/CalcCC.cu:237-271, 437-441) for one pair
 * of bx x by images: shift2 (host) receives the peak offset in binned
 * pixels. */
mf_status mf_stretch_cc(mf_ctx* ctx, const float* d_ref, const float* d_img,
                        int bx, int by, float ref_tilt, float tilt,
                        float axis_deg, float bfactor, float* shift2);

/* ProjAlign::MeaAlignProj (method/ProjAlign.cpp:243-278): correlation of a
 * reprojection with a corrected projection; shift2 (host) = ishift. */
mf_status mf_pa_measure(mf_ctx* ctx, const float* d_reproj,
                        const float* d_proj, int bx, int by, float* shift2);

/* CalcTIltAxis::Radon (method/Rotate.cu:397-428, GRadon :42-95) for the
 * candidate grid of CLsetup (:339-374) around axis0 over 'range' degrees:
 * the mean line values of every image of d_stack [nz][ny][nx] for every
 * candidate, as [nz][num][L] floats into d_lines (device), L = *line_len.
 * tilts and per-image shifts on the host; the common region follows the
 * zero-tilt image's shift.  d_lines NULL: only *line_len is set. */
mf_status mf_common_lines(mf_ctx* ctx, const float* d_stack, int nx, int ny,
                          int nz, const float* tilts, const float* shift_x,
                          const float* shift_y, float axis0, float range,
                          int num, float* d_lines, int* line_len);

/* CalcTIltAxis::Radon + FFT1d + findTiltAxis/CalcScore
 * (method/Rotate.cu:397-561) for the candidate grid of CLsetup (:339-374)
 * around axis0: the leave-one-out line-correlation score of every candidate
 * into scores (host, num floats), the best candidate before the wrap rules
 * of DoIt (:316-322) into *best_axis (the first maximum above 0, as
 * findTiltAxis), and, when d_spectra (device) is non-NULL, the line spectra
 * [nz][num][L/2+1] complex values (float pairs, unnormalised R2C of the
 * mean-removed, tapered lines).  Any output pointer may be NULL. */
mf_status mf_line_scores(mf_ctx* ctx, const float* d_stack, int nx, int ny,
                         int nz, const float* tilts, const float* shift_x,
                         const float* shift_y, float axis0, float range,
                         int num, float* scores, float* d_spectra,
                         int* line_len, float* best_axis);

/* FindOffset::Correlate (method/FindOffset.cu:176-238) of one w x h pair:
 * d_img stretched towards d_ref by cos(ref_tilt)/cos(tilt) (axis 0, no
 * fill), the six masked moments of GConv2D (:7-47) -- sums of v1, v2, v1^2,
 * v2^2, v1*v2 and the count of pixels valid in both -- into moments6 (host),
 * and the reference's correlation with quirk Q7 into *cc.  Either output
 * may be NULL. */
mf_status mf_ncc_moments(mf_ctx* ctx, const float* d_ref, const float* d_img,
                         int w, int h, float ref_tilt, float tilt,
                         double* moments6, float* cc);

/* FindOffset::SearchOffset + CalcAverangedCC (method/FindOffset.cu:129-174)
 * over a bin-4 corrected stack d_corr [nz][h][w] (device): the mean
 * correlation of each of the 31 one-degree steps into ccs (host, 31 floats,
 * may be NULL) and the best step (strict maximum above 0) into *best. */
mf_status mf_offset_scan(mf_ctx* ctx, const float* d_corr, int w, int h,
                         int nz, const float* tilts, float* ccs, float* best);

/* Peak search and parabolic sub-pixel fit on a w x h correlation map
 * d_map (device, row pitch 'pitch' floats): mode 0 CalcCC::FindPeak
 * (method/CalcCC.cu:381-441), mode 1 ProjAlign::PAFindPeak
 * (method/ProjAlign.cpp:325-375).  shift2 (host) = peak - (w/2, h/2). */
mf_status mf_peak2d(mf_ctx* ctx, const float* d_map, int w, int h, int pitch,
                    int mode, float* shift2);

/* mf_peak2d for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It allocates nothing: its partial results
 * live in a fixed device array of the library, one per GPU, held until the
 * process ends.  Maps of up to 32768 rows (h larger: MF_EINVAL).  Every
 * call on a GPU shares that array, so calls that overlap on one GPU race;
 * each call writes all it reads, so nothing carries over between calls.
 * Same results as mf_peak2d. */
mf_status mf_peak2d_noalloc(mf_ctx* ctx, const float* d_map, int w, int h,
                            int pitch, int mode, float* shift2);

/* ------------------------------------------------------------- a series */

/* Flags of the reference command line (main/opts.h:154-195, -g and -p). */
typedef struct {
  float offset;      /* -g field 1: tilt-angle offset, 0 = search it */
  float pitch_angle; /* -g field 2: tilt axis, 0 = search it */
  float zshift;      /* -g field 3 */
  int thickness;     /* -g field 4 (parsed, unused by the reference) */
  int align_z;       /* -g field 5: projection-matching thickness */
  int out_bin;       /* -g field 6 */
  int nproj;         /* -p: neighbouring projections per target */
  int verbose;       /* print the reference's stage trace to stdout */
  int reproj_mode;   /* mf_reproject mode 0..8: 0 slab, 1 fused, 8 fused with
                        one division per ray interval (default) */
  int pa_pairs;      /* 1 (default): a projection-matching sweep measures the
                        pair of targets z0+i, z0-i together when neither
                        reads the other's image; 0: one target at a time.
                        Both give identical results. */
  int graphs;        /* 1: the projection-matching sweeps after the first of
                        each setup replay a HIP graph captured from the
                        second (fewer launches; identical results); 0
                        (default): every sweep launches its kernels.  Sweeps
                        run eagerly while kernel timing is on. */
} mf_options;

/* Defaults of InitOpts (main/opts.h:67-81): AlignZ 600, OutBin 1, NPROJ 10;
 * reproj_mode 8, pa_pairs 1, graphs 0. */
void mf_options_default(mf_options* o);

typedef struct mf_series mf_series;

/* Copies a raw stack [nz][ny][nx] (file order, as read from the MRC) and its
 * tilt angles (file order, the .rawtlt) into HBM.  Replaces
 * Process::ReadStack/ReadAngles (main/Process.cpp:129-206) minus file I/O.
 * stack_zyx NULL: the raw stack is left to mf_series_upload. */
mf_status mf_series_create(mf_ctx* ctx, int nx, int ny, int nz,
                           const float* stack_zyx, const float* tilts,
                           mf_series** out);
/* Streaming ingest (no counterpart in the reference, which reads the whole
 * stack before any GPU work): copies raw images [z0, z0+count) from host
 * memory into a series created with stack_zyx NULL.  The copy is queued on
 * the series' stream (asynchronous when host is page-locked, from
 * mf_host_alloc), so later stages are ordered after it; the call returns
 * once the previous upload has finished, so a caller reading the next
 * images into the previous upload's buffer overlaps the reading with this
 * copy.  count 0 only waits for the last upload. */
mf_status mf_series_upload(mf_series* s, int z0, int count, const float* host);
/* Destroy every series before the context it was created on. */
void mf_series_destroy(mf_series* s);
const char* mf_series_last_error(const mf_series* s);
mf_status mf_series_set_verbose(mf_series* s, int verbose);
/* Graph replay of projection-matching sweeps (mf_options.graphs) for the
 * stage-by-stage calls; mf_run sets it from its options.  replays (may be
 * NULL) receives how many sweeps of this series ran as a graph replay.  No
 * counterpart in the reference. */
mf_status mf_series_set_graphs(mf_series* s, int on);
/* Releases the series' stage workspaces (coarse, axis, offset and
 * projection-matching buffers, the transposed stack), keeping the
 * preprocessed stack, tilts and parameters: what mf_correct and
 * mf_get_params need.  A later stage call needs its setup again.  No
 * counterpart in the reference. */
mf_status mf_series_trim(mf_series* s);
mf_status mf_series_graph_replays(const mf_series* s, long long* replays);

/* PreProcess::SetPositive + MassNormalization (method/Preprocess.cu:103-177);
 * reverses descending files into ascending order; zeroes the parameters. */
mf_status mf_preprocess(mf_series* s);
/* Preprocessed stack [nz][ny][nx] (ascending tilt order) to the host. */
mf_status mf_get_stack(mf_series* s, float* out);
/* Current tilts (ascending; including the offset once applied). */
mf_status mf_get_tilts(mf_series* s, float* tilts);
mf_status mf_set_tilts(mf_series* s, const float* tilts);
/* AlignParam (method/Util.h:14-21) in ascending tilt order. */
mf_status mf_get_params(mf_series* s, double* shift_x, double* shift_y,
                        float* rotate, float* angle_offset);
mf_status mf_set_params(mf_series* s, const double* shift_x,
                        const double* shift_y, const float* rotate,
                        float angle_offset);

/* Transform::Setup / DoIt (method/Transform.cpp:24-97): one coarse
 * stretched-CC sweep; maxerr = largest measured pair shift (px). */
mf_status mf_coarse_setup(mf_series* s);
mf_status mf_coarse_iter(mf_series* s, float* maxerr);
/* CalcTIltAxis::DoIt (method/Rotate.cu:292-337): common-line search over
 * 'num' candidates spanning 'range' degrees; sets every rotate[i]. */
mf_status mf_tiltaxis(mf_series* s, float range, int num, float* axis);
/* FindOffset::DoIt (method/FindOffset.cu:85-96): best tilt-angle offset
 * of the 31-step scan, added to angle_offset. */
mf_status mf_find_offset(mf_series* s, float* best);
/* Adds angle_offset to the tilts (main/Process.cpp:266-269). */
mf_status mf_apply_offset(mf_series* s);
/* ProjAlign::Setup (method/ProjAlign.cpp:48-125). */
mf_status mf_projalign_setup(mf_series* s, float zshift, int align_z,
                             int nproj, int reproj_mode);
/* ProjAlign::DoIt (method/ProjAlign.cpp:127-179): one projection-matching
 * sweep over all targets; maxerr = largest measured shift (px). */
mf_status mf_projalign_iter(mf_series* s, float* maxerr);
/* Process::ProjAlignOnce (main/Process.cpp:292-319). */
mf_status mf_projalign_once(mf_series* s);
/* Process::CorrectStack (main/Process.cpp:329-372): the aligned stack
 * [nz][by][bx] at OutBin 'out_bin' to the host (out may be NULL to query
 * bx, by). */
mf_status mf_correct(mf_series* s, int out_bin, float* out, int* bx, int* by);
/* mf_correct that returns once the correction and the download are queued
 * on the series' stream: with 'out' in page-locked memory (mf_host_alloc)
 * the download runs while the host goes on, for example to the next
 * series' mf_run, and 'out' holds the aligned stack once mf_series_wait
 * returns (other host memory is filled before the call returns).  'out'
 * must stay allocated until then.  No counterpart in the reference (its
 * CorrectStack writes the file synchronously). */
mf_status mf_correct_async(mf_series* s, int out_bin, float* out, int* bx,
                           int* by);
/* Waits for all queued work of the series, the download of
 * mf_correct_async included.  No counterpart in the reference. */
mf_status mf_series_wait(mf_series* s);
/* Process::DoIt (main/Process.cpp:34-92) minus file I/O: preprocess, coarse
 * alignment, axis and offset searches, projection matching. */
mf_status mf_run(mf_series* s, const mf_options* o);

/* The last projection-matching sweep's measured shift of every target (raw
 * pixels, [nz][2], before RotShift; the zero tilt holds 0) and its length
 * |shift| (ProjAlign::ClacAlignProj, method/ProjAlign.cpp:181-241, the values
 * the sweep's verbose table prints).  Either pointer may be NULL. */
mf_status mf_pa_measurements(mf_series* s, float* meas, float* err);
/* Scores of the last mf_tiltaxis (*num = candidates; up to cap written) and
 * of the last mf_find_offset (31 values); NULL pointers are skipped. */
mf_status mf_axis_scores(mf_series* s, float* scores, int cap, int* num);
mf_status mf_offset_scores(mf_series* s, float* ccs);

/* Stage trace: every stage call of a series appends one event (kind,
 * value) -- the stage sequence Process::DoIt drives (main/Process.cpp:34-92,
 * 238-319), with the values its stdout lines print.  mf_series_trace copies
 * up to cap events and sets *count to the number recorded. */
enum {
  MF_EV_PREPROCESS = 1,   /* value 0 */
  MF_EV_COARSE = 2,       /* Transform::DoIt; value maxerr */
  MF_EV_AXIS = 3,         /* CalcTIltAxis::DoIt; value the new axis */
  MF_EV_OFFSET = 4,       /* FindOffset::DoIt; value the scan's best step */
  MF_EV_APPLY_OFFSET = 5, /* value the total offset */
  MF_EV_PA_SETUP = 6,     /* ProjAlign::Setup; value the slab thickness */
  MF_EV_PA_ITER = 7,      /* ProjAlign::DoIt; value maxerr */
  MF_EV_ONCE_BEGIN = 8,   /* ProjAlignOnce entered; value 0 */
  MF_EV_ONCE_END = 9      /* value the number of DoIt calls it made */
};
mf_status mf_series_trace(mf_series* s, int* kinds, float* values, int cap,
                          int* count);
mf_status mf_series_trace_clear(mf_series* s);

/* ------------------------------------------- batched and served forms
 * No counterpart in the reference, which aligns one series per process;
 * these run many series on one GPU.  Each series' results equal what
 * mf_run computes for it alone. */

/* mf_run on n series of one context (sizes and tilt counts may differ),
 * 'concurrency' at a time: each series has its own stream, so the kernels
 * of different series overlap on the device.  Each series is trimmed
 * (mf_series_trim) after its run, so a finished series holds only its
 * stack until mf_correct. */
mf_status mf_batch_run(mf_series** series, int n, const mf_options* o,
                       int concurrency);

/* The streaming batched form: mf_run on each series with its own options
 * (opts[i], n entries), 'concurrency' series at a time on their own
 * streams, largest series (nx*ny*nz) first; each series is trimmed after
 * its run.  correct nonzero: each series is then corrected at its
 * opts[i].out_bin into a page-locked block of the worker (kept by the
 * context between calls) and trimmed again.  done (may be NULL) is called
 * from the worker once per series, with the series' index, status and,
 * when corrected, its aligned stack [nz][by][bx]; that memory is valid only
 * during the call.  Callbacks of different workers run concurrently.
 * Returns the first failing status in index order. */
typedef void (*mf_batch_done)(int index, mf_status status,
                              const float* aligned, int bx, int by, int nz,
                              void* user);
mf_status mf_batch_run_each(mf_series** series, int n, const mf_options* opts,
                            int concurrency, int correct, mf_batch_done done,
                            void* user);

/* Upper bound of the device bytes one series of this shape holds from
 * mf_series_create through mf_correct (the served form's admission
 * charge). */
size_t mf_series_hbm_bytes(int nx, int ny, int nz, const mf_options* o);

/* Request loop: series arrive with mf_submit, wait in arrival order, and
 * start when a worker is free and their byte bound fits next to the
 * running series within 'hbm_budget' (0: the context's budget, else the
 * device memory free at creation less 1/16).  Raw stacks are copied at
 * submission into page-locked host memory; results (parameters and the
 * OutBin aligned stack, also page-locked) stay in host memory until
 * released. */
typedef struct mf_server mf_server;
enum { MF_REQ_QUEUED = 0, MF_REQ_RUNNING = 1, MF_REQ_DONE = 2,
       MF_REQ_FAILED = 3 };
mf_status mf_server_create(mf_ctx* ctx, int workers, size_t hbm_budget,
                           int queue_cap, mf_server** out);
/* Waits for running series; queued ones fail with "server stopped". */
void mf_server_destroy(mf_server* sv);
/* MF_EBUSY: 'queue_cap' requests already wait; MF_ENOMEM: the series could
 * never fit the budget, or no host memory is left for its copy (the message
 * says which, mf_last_error of the server's context).  The copy goes to
 * page-locked memory where the system grants it, else to ordinary memory. */
mf_status mf_submit(mf_server* sv, int nx, int ny, int nz,
                    const float* stack_zyx, const float* tilts,
                    const mf_options* o, long long* id);
/* mf_submit without the copy: the server reads stack_zyx in place when the
 * request starts, so the caller keeps it valid and unchanged until the
 * request is done or failed (mf_poll / mf_wait).  A stack in page-locked
 * memory (mf_host_alloc) is then uploaded at the PCIe rate with no host
 * copy at all. */
mf_status mf_submit_borrowed(mf_server* sv, int nx, int ny, int nz,
                             const float* stack_zyx, const float* tilts,
                             const mf_options* o, long long* id);
/* Seconds since the server was created: the clock of mf_fetch's times3. */
mf_status mf_server_clock(mf_server* sv, double* seconds);
mf_status mf_poll(mf_server* sv, long long id, int* state);
mf_status mf_wait(mf_server* sv, long long id, int* state);
/* Results of a finished request (any pointer may be NULL); times3 =
 * submit, start, finish in seconds since the server started.  A failed
 * request returns MF_EDEVICE with its message in err. */
mf_status mf_fetch(mf_server* sv, long long id, double* shift_x,
                   double* shift_y, float* rotate, float* angle_offset,
                   float* aligned, int* bx, int* by, double* times3,
                   char* err, size_t err_size);
mf_status mf_release(mf_server* sv, long long id);
mf_status mf_server_stats(mf_server* sv, size_t* budget, size_t* reserved,
                          int* queued, int* running, long long* completed);

#ifdef __cplusplus
}
#endif
#endif /* MARKERFREE_MF_API_H */
