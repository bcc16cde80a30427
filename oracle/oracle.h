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

/* FindProjRange (CalcReproj.cu:206-259).  Writes [istart, iend]; istart = -1
 * when no projection qualifies. */
void orc_find_proj_range(const float* tilts, int nz, const unsigned char* skip,
                         int target, int nproj, int* range2);

/* One target of CalcReprojFBP::DoIt (CalcReproj.cu:180-204, GetSinogram
 * :261-274, Reproj :276-286, kernels mGBackProj :26-64 and mGForProj
 * :68-124): for every image
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
 * pitch (nx/2+1)*2 floats, unnormalised (cufft2D::ForwardFFT without bNorm,
 * Cufft1&2.cu:47-62). */
void orc_fft_r2c_2d(float* buf, int nx, int ny);
/* In-place complex-to-real, unnormalised (cufft2D::InverseFFT,
 * Cufft1&2.cu:64-69). */
void orc_fft_c2r_2d(float* buf, int nx, int ny);
/* Batched in-place 1-D R2C / C2R of length n on lines of pitch (n/2+1)*2
 * (cufft1D::Forward / Inverse, Cufft1&2.cu:140-165). */
void orc_fft_r2c_1d(float* buf, int n, int nlines);
void orc_fft_c2r_1d(float* buf, int n, int nlines);

/* ---- stack correction (CorrStack.cu) ---- */
/* mCorrectProj for one projection (CorrStack.cu:370-397; kernels mGCorrect
 * :142-179, mGRandom :116-140, mGBinImage :40-63, GRWeight :431-446): rotate+shift
 * (nearest), optional random fill, b x b sum binning, optional R-weighting,
 * unpadded output [by][bx].  Returns bx, by through out_size. */
void orc_correct_proj(const float* raw, int nx, int ny, float shift_x,
                      float shift_y, float axis_deg, int shift_only,
                      int random_fill, int bin, int rweight, float* out,
                      int* out_size);
void orc_binned_size(int nx, int ny, int bin, int* out_size);

/* ---- stretched cross-correlation (CalcCC.cu) ---- */
/* CalcCC::DoIt + getshift (CalcCC.cu:237-271, 437-441) for one pair of
 * bx x by images (unpadded input);
 * returns the shift in binned pixels (before the bin factor). */
void orc_stretch_cc(const float* ref, const float* img, int bx, int by,
                    float ref_tilt, float tilt, float axis_deg, float bfactor,
                    float* shift2);
/* ProjAlign::MeaAlignProj (ProjAlign.cpp:243-278; PANorm :295-312,
 * PAFindPeak :325-375) on a reprojection and a
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

/* CalcTIltAxis::Radon (Rotate.cu:397-428, GRadon :42-95) over the CLsetup
 * grid around axis0: lines [nz][num][L] of mean values; returns L (lines
 * NULL: L only). */
int orc_common_lines(const float* stack, int nx, int ny, int nz,
                     const float* tilts, const float* shift_x,
                     const float* shift_y, float axis0, float range, int num,
                     float* lines);

/* Radon + FFT1d (Rotate.cu:397-451; GCalcMean :97-142, GRemoveMean :144-157)
 * of every image: the line spectra [nz][num][L/2+1] complex (float pairs)
 * that CalcScore correlates; returns L (spec NULL: L only). */
int orc_line_spectra(const float* stack, int nx, int ny, int nz,
                     const float* tilts, const float* shift_x,
                     const float* shift_y, float axis0, float range, int num,
                     float* spec);

/* ---- peak search (CalcCC.cu, ProjAlign.cpp) ---- */
/* mode 0: CalcCC::FindPeak (CalcCC.cu:381-435); mode 1: PAFindPeak
 * (ProjAlign.cpp:325-375), on a w x h map with row pitch 'pitch' floats;
 * shift2 = sub-pixel peak - (w/2, h/2). */
void orc_peak2d(const float* map, int w, int h, int pitch, int mode,
                float* shift2);

/* ---- tilt-angle offset (FindOffset.cu) ---- */
/* FindOffset::DoIt with FOsetup (FindOffset.cu:85-127) on the preprocessed
 * stack: bin-4 correction, then the 31-step scan; ccs receives the mean
 * correlation of every step. */
float orc_find_offset(const float* stack, int nx, int ny, int nz,
                      const float* tilts, const float* shift_x,
                      const float* shift_y, const float* rotate, float* ccs);
/* FindOffset::Correlate (FindOffset.cu:176-238) of one w x h pair. */
float orc_offset_correlate(const float* ref, const float* img, int w, int h,
                           float ref_tilt, float tilt);
/* The six masked moments of Correlate (GConv2D, FindOffset.cu:7-47):
 * sums of v1, v2, v1^2, v2^2, v1 v2 and the count of pixels valid in both. */
void orc_offset_moments(const float* ref, const float* img, int w, int h,
                        float ref_tilt, float tilt, double* m6);
/* SearchOffset + CalcAverangedCC (FindOffset.cu:129-174) over a bin-4
 * corrected stack [nz][h][w]; returns the best offset, ccs[31]. */
float orc_offset_scan(const float* corr, int w, int h, int nz,
                      const float* tilts, float* ccs);

/* ---- whole pipeline (Process.cpp) ---- */
/* Stage trace events, one per stage call: (kind, value). */
enum {
  ORC_EV_PREPROCESS = 1,   /* value 0 */
  ORC_EV_COARSE = 2,       /* Transform::DoIt; value maxerr */
  ORC_EV_AXIS = 3,         /* CalcTIltAxis::DoIt; value the new axis */
  ORC_EV_OFFSET = 4,       /* FindOffset::DoIt; value the step's best */
  ORC_EV_APPLY_OFFSET = 5, /* value the total offset */
  ORC_EV_PA_SETUP = 6,     /* value the slab thickness T_b */
  ORC_EV_PA_ITER = 7,      /* ProjAlign::DoIt; value maxerr */
  ORC_EV_ONCE_BEGIN = 8,   /* ProjAlignOnce entered */
  ORC_EV_ONCE_END = 9      /* value the number of DoIt calls it made */
};
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
/* ProjAlign::DoIt (ProjAlign.cpp:127-179). */
float orc_projalign_iter(orc_series*);
/* The same sweep where forced_meas ([nz][2] raw pixels, nullable) replaces
 * each target's measured shift before it is applied (so that two
 * implementations can be kept on one trajectory), and own_meas ([nz][2],
 * nullable) receives the shift this sweep measured itself. */
float orc_projalign_iter_forced(orc_series*, const float* forced_meas,
                                float* own_meas);
/* Rotation-centre height from the last FitRotCenterZ (ProjAlign.cpp:459-486). */
float orc_pa_z0(const orc_series*);
/* Process::ProjAlignOnce (Process.cpp:292-319). */
void orc_projalign_once(orc_series*);
/* The DoIt calls ProjAlignOnce makes when its sweeps return errs[0..n)
 * in turn (its break rule alone; 10 at most). */
int orc_projalign_once_script(const float* errs, int n);
void orc_correct_final(orc_series*, int out_bin, float* out, int* out_size);
/* Process::DoIt (Process.cpp:34-92) minus file I/O. */
void orc_run(orc_series*, const orc_options*);
/* The stage trace since creation (or the last clear): up to cap events;
 * returns the number recorded. */
int orc_series_trace(const orc_series*, int* kinds, float* values, int cap);
void orc_series_trace_clear(orc_series*);
void orc_get_corrected(const orc_series*, float* out, int* out_size);

/* CPU threads the oracle uses (OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
