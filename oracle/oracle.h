// CPU restatement of the Markerfree alignment hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the timed
// CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the baseline: the
// product (libmarkerfree_hip.so, the markerfree CLI) never links or calls it.
//
// Every function restates the reference algorithm step by step, citing the
// reference file:line it follows (paths relative to the reference checkout).
// Bulk arithmetic is FP32 with the reference's expression order; reductions
// accumulate in FP64 (the reference's FP32 tree order is not reproducible
// without the reference and is documented as parity unpinned).  Transcendental
// values that decide pixel indices (cos/sin of tilt and axis angles) are taken
// from host tables computed in double and rounded to float, the same tables the
// HIP path uses, so that index selection is bit-identical.
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reprojection for projection matching (CalcReproj.cu) ---- */

/* FindProjRange (CalcReproj.cu:281-334).  Writes [istart, iend]; istart = -1
 * when no projection qualifies. */
void orc_find_proj_range(const float* tilts, int nz, const unsigned char* skip,
                         int target, int nproj, int* range2);

/* One target of CalcReprojFBP::DoIt (CalcReproj.cu:255-279): for every image
 * row, backproject the in-range rows into an xz slab of 2bx x tb voxels, then
 * ray-sum it at the target tilt.  acc64 != 0 accumulates values in FP64 (the
 * FP64 oracle); acc64 == 0 accumulates in FP32 in the reference's per-voxel
 * order.  slab_out (nullable) receives the full slab [by][tb][2bx]. */
void orc_reproject(const float* stack, int bx, int by, int nz, int tb,
                   const float* tilts, const unsigned char* skip, int target,
                   int nproj, int acc64, float* reproj, float* slab_out);

/* Per-target algorithmic traffic helpers used by the benchmark: in-bounds RP
 * samples for one image row (CalcReproj.cu:68-124 bounds), independent of
 * data. */
long long orc_rp_inbounds_samples(int bx, int tb, float target_tilt);

/* ---- FFT (third-party cuFFT in the reference; restated from the DFT) ---- */
/* In-place real-to-complex 2-D transform of an nx x ny image stored with row
 * pitch (nx/2+1)*2 floats, unnormalised (cufftExecR2C, Cufft1&2.cu:50). */
void orc_fft_r2c_2d(float* buf, int nx, int ny);
/* In-place complex-to-real, unnormalised (cufftExecC2R, Cufft1&2.cu:67). */
void orc_fft_c2r_2d(float* buf, int nx, int ny);
/* Batched in-place 1-D R2C / C2R of length n on lines of pitch (n/2+1)*2. */
void orc_fft_r2c_1d(float* buf, int n, int nlines);
void orc_fft_c2r_1d(float* buf, int n, int nlines);

/* ---- stack correction (CorrStack.cu) ---- */
/* mCorrectProj for one projection (CorrStack.cu:503-530): rotate+shift
 * (nearest), optional random fill, b x b sum binning, optional R-weighting,
 * unpadded output [by][bx].  Returns bx, by through out_size. */
void orc_correct_proj(const float* raw, int nx, int ny, float shift_x,
                      float shift_y, float axis_deg, int shift_only,
                      int random_fill, int bin, int rweight, float* out,
                      int* out_size);
void orc_binned_size(int nx, int ny, int bin, int* out_size);

/* ---- stretched cross-correlation (CalcCC.cu) ---- */
/* CalcCC::DoIt + getshift for one pair of bx x by images (unpadded input);
 * returns the shift in binned pixels (before the bin factor). */
void orc_stretch_cc(const float* ref, const float* img, int bx, int by,
                    float ref_tilt, float tilt, float axis_deg, float bfactor,
                    float* shift2);
/* ProjAlign::MeaAlignProj (ProjAlign.cpp:321-375) on a reprojection and a
 * corrected projection: returns ishift (binned pixels). */
void orc_pa_measure(const float* reproj, const float* proj, int bx, int by,
                    float* shift2);

/* ---- tilt-axis search (Rotate.cu) ---- */
/* CalcTIltAxis::DoIt (Rotate.cu:292-337) on the preprocessed stack; returns
 * the new axis and writes the 'num' candidate scores. */
float orc_tilt_axis(const float* stack, int nx, int ny, int nz,
                    const float* tilts, const float* shift_x,
                    const float* shift_y, const float* rotate, float range,
                    int num, float* scores, int* line_size);

/* CalcTIltAxis::Radon (Rotate.cu:397-428, GRadon :83-136) over the CLsetup
 * grid around axis0: lines [nz][num][L] of mean values; returns L (lines
 * NULL: L only). */
int orc_common_lines(const float* stack, int nx, int ny, int nz,
                     const float* tilts, const float* shift_x,
                     const float* shift_y, float axis0, float range, int num,
                     float* lines);

/* ---- tilt-angle offset (FindOffset.cu) ---- */
float orc_find_offset(const float* stack, int nx, int ny, int nz,
                      const float* tilts, const float* shift_x,
                      const float* shift_y, const float* rotate, float* ccs);
/* FindOffset::Correlate (FindOffset.cu:204-266) of one w x h pair. */
float orc_offset_correlate(const float* ref, const float* img, int w, int h,
                           float ref_tilt, float tilt);

/* ---- whole pipeline (Process.cpp) ---- */
typedef struct {
  float offset, pitch_angle, zshift;
  int thickness, align_z, out_bin, nproj, verbose;
} orc_options;

typedef struct orc_series orc_series;
orc_series* orc_series_create(int nx, int ny, int nz, const float* raw,
                              const float* tilts_file_order);
void orc_series_destroy(orc_series*);
void orc_preprocess(orc_series*);
void orc_get_stack(const orc_series*, float* out);
void orc_get_tilts(const orc_series*, float* out);
void orc_set_tilts(orc_series*, const float* in);
void orc_get_params(const orc_series*, float* sx, float* sy, float* rot,
                    float* offset);
void orc_set_params(orc_series*, const float* sx, const float* sy,
                    const float* rot, float offset);
void orc_coarse_setup(orc_series*);
float orc_coarse_iter(orc_series*);
float orc_series_tilt_axis(orc_series*, float range, int num);
float orc_series_find_offset(orc_series*);
void orc_apply_offset(orc_series*);
void orc_projalign_setup(orc_series*, float zshift, int align_z, int nproj);
float orc_projalign_iter(orc_series*);
/* Rotation-centre height from the last FitRotCenterZ (ProjAlign.cpp:459-486). */
float orc_pa_z0(const orc_series*);
void orc_projalign_once(orc_series*);
void orc_correct_final(orc_series*, int out_bin, float* out, int* out_size);
void orc_run(orc_series*, const orc_options*);
void orc_get_corrected(const orc_series*, float* out, int* out_size);

/* CPU threads the oracle uses (OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
