// Tilt-axis (CalcTIltAxis, method/Rotate.cu) and tilt-angle offset
// (FindOffset, method/FindOffset.cu) searches.
#pragma once
#include <vector>

#include "context.h"

namespace mf {

// Candidate grid and line length of CLsetup (Rotate.cu:339-374).
struct AxisSetup {
  int num = 0, L = 0;
  std::vector<float> rot;  // candidate angles (degrees)
  std::vector<float> cs;   // cos, sin of rot * 3.1415926f / 180.0f
  static AxisSetup make(int nx, int ny, float axis0, float range, int num);
};

struct AxisResult {
  float axis = 0.0f;          // best candidate, before the wrap rules
  std::vector<float> scores;  // per candidate
  int L = 0;
};

struct AxisWork {
  DevBuf lines, reg, small;
  // Transposed copy of the stack for column-like candidates, and the stack
  // it was made from (reset this when that stack's pixels change).
  DevBuf tstack;
  const void* t_of = nullptr;
  // Per image, kNegParts flags: whether that part of the image holds a
  // negative value (a sample GRadon skips); made once per stack like tstack.
  DevBuf negf;
  const void* neg_of = nullptr;
};

struct OffsetWork {
  DevBuf buf;
};

// The stack transposed ([nz][nx][ny]) in ws.tstack, made unless ws already
// holds it for d_stack.  Asynchronous on 'st'.
const float* stack_transposed(mf_ctx* ctx, hipStream_t st,
                              const float* d_stack, int nx, int ny, int nz,
                              AxisWork& ws);

// GRadon (Rotate.cu:397-428, 42-95): the mean line values of every image
// for every candidate of 'a', in ws.lines as [nz][num] lines of a.L values at
// a pitch of a.L + 2 floats (returned).  Asynchronous on 'st'.
float* common_lines(mf_ctx* ctx, hipStream_t st, const float* d_stack, int nx,
                    int ny, int nz, const float* tilts, const double* sx,
                    const double* sy, const AxisSetup& a, AxisWork& ws);

// Radon lines of every image for every candidate, their spectra and the
// leave-one-out line correlation score (Rotate.cu:292-561).  Synchronous.
// The spectra stay in ws.lines ([nz][num] lines at a pitch of L + 2 floats)
// until the next search.
AxisResult axis_search(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                       int nx, int ny, int nz, const float* tilts,
                       const double* sx, const double* sy, float axis0,
                       float range, int num, AxisWork& ws);

// FindOffset::Correlate's NCC with quirk Q7 from the six masked moments
// (sums of v1, v2, v1^2, v2^2, v1 v2 and the count).
float offset_cc(const double m[6]);

// The six masked moments of FindOffset::Correlate (FindOffset.cu:176-238,
// GConv2D :7-47) of one w x h pair: d_img stretched (axis 0, no fill)
// towards d_ref.  Synchronous.
void ncc_moments(mf_ctx* ctx, hipStream_t st, const float* d_ref,
                 const float* d_img, int w, int h, float ref_tilt, float tilt,
                 OffsetWork& ws, double m6[6]);

// 31-step offset scan over a bin-4 corrected stack [nz][h][w]
// (FindOffset.cu:129-238).  Returns the best offset; ccs (optional) gets the
// mean correlation per step.  Synchronous.
float offset_search(mf_ctx* ctx, hipStream_t st, const float* d_corr, int w,
                    int h, int nz, const float* tilts, OffsetWork& ws,
                    std::vector<float>* ccs);

}  // namespace mf
