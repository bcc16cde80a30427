// Cross-correlation stages: CalcCC (method/CalcCC.cu) for the stretched
// coarse alignment and ProjAlign::MeaAlignProj (method/ProjAlign.cpp:243-375)
// for projection matching.
#pragma once
#include <vector>

#include "context.h"
#include "stages.h"

namespace mf {

// One stretched pair: reference image and image indices into an image array,
// and the inverse stretch matrix of CalcCC::Stretch (CalcCC.cu:318-345).
struct PairDesc {
  int ref, img;
  float m[3];
};

// Host: the stretch matrix exactly as CalcCC::Stretch forms it.
void stretch_matrix(float stretch, float axis_deg, float m[3]);

// Masks of GRoundEdge (CalcCC.cu:173-195) as host-built tables: 'coarse'
// (centre w/2,h/2, size w,h) and 'pa' (centre w/2,h/2, size pad_x(w),h,
// ProjAlign.cpp:295-312).
const float* cc_mask(mf_ctx* ctx, int w, int h, bool pa);

// Batched CalcCC::DoIt for the pairs of w x h images from 'stack'
// (contiguous [n][h][w]).  Workspace: stretch_cc_work_floats() floats and
// cc_scratch_doubles(npair, h) doubles of reduction partials.  d_shift receives
// float2 shifts (binned pixels, before the bin factor); returns after the
// stream has drained.
void stretch_cc_pairs(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                      int w, int h, const std::vector<PairDesc>& pairs,
                      float bfactor, float* d_work, double* d_mom,
                      DevBuf& tab, float* d_shift);

// The per-target parameter update of a projection-matching step
// (ClacAlignProj, ProjAlign.cpp:200-221; RotShift, :448-457), applied by the correlation's fit for map
// j to target t[j]: meas = ishift * bin, err = |meas|, shift = RotShift(meas)
// + Z-induced shift, (sx, sy) += shift, and the target's correction
// parameters cp[t] refreshed.  sx == nullptr: no update.
struct PaUpdate {
  float* meas;        // [nz][2]
  float* err;         // [nz]
  const float* ind;   // [nz][2]
  const float* rcs;   // [nz][2] cos, sin of the axis
  double* sx;
  double* sy;
  CorrProj* cp;
  int t[2];
  float binx, biny;
};

// MeaAlignProj: d_reproj and d_proj are w x h; writes the float2 ishift.
// Workspace: 2*h*pad_x(w) floats + cc_scratch_doubles(1, h) doubles.
void pa_measure(mf_ctx* ctx, hipStream_t st, const float* d_reproj,
                const float* d_proj, int w, int h, float* d_work,
                double* d_mom, float* d_ishift,
                const PaUpdate* up = nullptr);

// One input image of a correlation: its pixels (row pitch in floats) and the
// padded work image [h][pad_x(w)] it is normalised and transformed in.
struct CcImg {
  const float* src;
  int pitch;
  float* dst;
};

// MeaAlignProj for npair independent pairs at once, images from the device
// table d_tab (entry 2k: reprojection k, entry 2k+1: projection k, each with
// its own work image); d_ishift receives npair float2.  Every pair's result
// equals pa_measure's on the same images.  Workspace: cc_scratch_doubles(
// npair, h) doubles.
void pa_measure_tab(mf_ctx* ctx, hipStream_t st, const CcImg* d_tab,
                    int npair, int w, int h, double* d_mom, float* d_ishift,
                    const PaUpdate* up = nullptr);

// Peak search and sub-pixel fit on a w x h correlation map (row pitch
// 'pitch' floats) in device memory: mode 0 CalcCC::FindPeak
// (CalcCC.cu:381-441), mode 1 PAFindPeak (ProjAlign.cpp:325-375); d_shift
// receives the float2 offset of the peak from the map centre.  The same
// window, tie rule and fit as the fused correlation's.  Asynchronous.
void peak2d(mf_ctx* ctx, hipStream_t st, const float* d_map, int w, int h,
            int pitch, int mode, DevBuf& ws, float* d_shift);

// peak2d for callers that cannot allocate: partials and result in a
// file-scope device array (maps of up to 32768 rows), shift2 to the host;
// synchronous.  peak_scratch_address(device) returns that array's address
// on 'device', which must be current; mf_ctx_create keeps it in the context.
void peak2d_noalloc(mf_ctx* ctx, hipStream_t st, const float* d_map, int w,
                    int h, int pitch, int mode, float* h_shift);
void* peak_scratch_address(int device);

size_t stretch_cc_work_floats(int w, int h, int npair);
size_t cc_scratch_doubles(int npair, int h);

}  // namespace mf
