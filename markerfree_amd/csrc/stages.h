// Device stages of the alignment path (launchers).  Each mirrors one
// reference method class; the .hip files cite the kernels they replace.
#pragma once
#include <vector>

#include "context.h"

namespace mf {

// ---------------------------------------------------------------- sizes
// One CorrTomoStack configuration (method/CorrStack.cu:298-328).
struct CorrGeom {
  int nx, ny;   // raw image
  int cx;       // corrected width ((pad_x(nx)/2 - 1) * 2)
  int bx, by;   // binned size
  int binx, biny;
  static CorrGeom make(int nx, int ny, int bin);
};

// Per-projection correction parameters as the kernels take them.
struct CorrProj {
  float shift_x, shift_y;  // global shift (raw pixels)
  float cos_a, sin_a;      // of the axis angle in radians (0 when shift-only)
};

// ----------------------------------------------------------- preprocess
// Per-slice min and max over valid pixels (>= -1e10) of nz slices of 'plane'
// values (Preprocess.cu:7-85, 179-200); out [nz][2] on the host.
void slice_minmax(mf_ctx* ctx, hipStream_t st, const float* d, size_t plane,
                  int nz, DevBuf& tmp, float* h_out);
// stack[a] = raw[a], or stack[nz-1-a] = raw[a] when 'rev' (descending
// files), minus sub for valid pixels (>= -1e10) when sub != 0; one launch.
void reorder_subtract(mf_ctx* ctx, hipStream_t st, const float* d_raw,
                      float* d_stack, size_t plane, int nz, bool rev,
                      float sub);
// Mean of the central box of each slice, skipping values <= -1e20
// (Preprocess.cu:216-237); FP64 sums, host out [nz].
void box_means(mf_ctx* ctx, hipStream_t st, const float* d, int nx, int ny,
               int nz, int x0, int y0, int w, int h, DevBuf& tmp,
               float* h_out);
// Multiply valid (> -1e20) pixels of each slice by scale[z].
void scale_slices(mf_ctx* ctx, hipStream_t st, float* d, size_t plane, int nz,
                  DevBuf& tmp, const float* h_scale);

// ----------------------------------------------------------- correction
// mGCorrect + mGRandom + mGBinImage (CorrStack.cu:40-63, 116-179) for 'count'
// projections starting at index z0 of the raw stack; params from d_proj
// (device CorrProj[nz]).  Writes binned images with row pitch out_pitch and
// image stride out_stride (floats).
// d_raw_t (optional): d_raw transposed ([nz][nx][ny]); projections rotated
// nearer a quarter turn than none gather from it (same values).
void correct_bin(mf_ctx* ctx, hipStream_t st, const float* d_raw,
                 const CorrGeom& g, const CorrProj* d_proj, int z0, int count,
                 bool random_fill, float* d_out, int out_pitch,
                 size_t out_stride, const float* d_raw_t = nullptr);
// RWeight (CorrStack.cu:415-462) + Unpad: each padded binned row (pitch
// pad_x(bx)) is R2C'd, ramp-filtered, C2R'd and written unpadded to d_out.
void rweight_rows(mf_ctx* ctx, hipStream_t st, const float* d_pad, int bx,
                  int nrows, float* d_out);
// correct_bin with random fill choice, then rweight_rows, in one pass: the
// R-weighted binned projections z0 + k*zstep (k < count) go to d_out as
// planes of by x bx (unpadded), projection z0 + k*zstep at plane k*zstep
// (d_out holds projection z0 first).
void correct_rweight(mf_ctx* ctx, hipStream_t st, const float* d_raw,
                     const CorrGeom& g, const CorrProj* d_proj, int z0,
                     int count, bool random_fill, float* d_out, int zstep = 1,
                     const float* d_raw_t = nullptr);

// ------------------------------------------------------- correlation
// Per-image FP64 moments sum(v) and sum(v^2) over valid pixels of w x h
// images (pitch, stride in floats); out [nimg][2] (device).
void image_moments(mf_ctx* ctx, hipStream_t st, const float* d, int w, int h,
                   int pitch, size_t stride, int nimg, double* d_out);

}  // namespace mf
