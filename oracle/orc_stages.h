// Internal stage interfaces of the CPU oracle (test infrastructure only).
#pragma once
#include <vector>

namespace orc {

// Sizes of one CorrTomoStack configuration (CorrStack.cu:298-328).
struct CorrSizes {
  int nx, ny;    // raw image
  int pad;       // padded row pitch of the corrected image
  int cx;        // corrected image width (pad/2 - 1) * 2
  int bx, by;    // binned size
  int padb;      // padded row pitch of the binned image
  int binx, biny;
};

void corr_sizes(int nx, int ny, int bin, CorrSizes* s);
void correct_one(const float* raw, const CorrSizes& s, float shift_x,
                 float shift_y, float axis_deg, bool shift_only,
                 bool random_fill, bool rweight, float* out);
void correct_stack(const float* stack, int nx, int ny, int nz, int bin,
                   const float* sx, const float* sy, const float* rot,
                   bool shift_only, bool random_fill, bool rweight,
                   std::vector<float>& out, CorrSizes* sizes);

// CalcCC::DoIt + FindPeak (CalcCC.cu:237-435); shift in binned pixels.
void stretch_cc(const float* ref, const float* img, int bx, int by,
                float ref_tilt, float tilt, float axis_deg, float bfactor,
                float* shift2);
// CalcCC::Stretch without padding and without random fill (FindOffset use).
void stretch_plain(const float* in, int w, int h, float stretch,
                   float axis_deg, float* out);
// ProjAlign::MeaAlignProj (ProjAlign.cpp:243-375).
void pa_measure(const float* reproj, const float* proj, int bx, int by,
                float* shift2);

}  // namespace orc
