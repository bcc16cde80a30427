// CPU restatement of the tilt-angle offset search (FindOffset.cu) -- test
// infrastructure.
#include "oracle.h"
#include "orc_internal.h"
#include "orc_stages.h"

namespace orc {

// FindOffset::Correlate (FindOffset.cu:176-238), first half: stretch the
// image towards its reference (axis 0, no fill, CalcCC::Stretch) and the six
// masked moments of GConv2D (FindOffset.cu:7-47) over pixels valid in both:
// sums of v1, v2, v1^2, v2^2, v1 v2 and the count.
void offset_moments(const float* ref, const float* img, int w, int h,
                    float ref_tilt, float tilt, double m[6]) {
  double ds = std::cos(kD2R * ref_tilt) / std::cos(kD2R * tilt);
  std::vector<float> st((size_t)w * h);
  stretch_plain(img, w, h, (float)ds, 0.0f, st.data());
  for (int i = 0; i < 6; i++) m[i] = 0.0;
  for (size_t i = 0; i < (size_t)w * h; i++) {
    float v1 = ref[i];
    if (v1 < -1e10f) continue;
    float v2 = st[i];
    if (v2 < -1e10f) continue;
    m[0] += v1;
    m[1] += v2;
    m[2] += (double)(v1 * v1);
    m[3] += (double)(v2 * v2);
    m[4] += (double)(v1 * v2);
    m[5] += 1.0;
  }
}

// Second half (FindOffset.cu:221-237): the reference's NCC with quirk Q7
// ("std" = sqrt(E[x^2] - E[x])), clamped to [-1, 1].
float offset_cc(const double m[6]) {
  float a[6];
  for (int i = 0; i < 6; i++) a[i] = (float)m[i];
  for (int i = 0; i < 5; i++) a[i] /= a[5];
  double s1 = a[2] - a[0];
  double s2 = a[3] - a[1];
  double cc = a[4] - a[0] * a[1];
  if (s1 < 0) s1 = 0;
  if (s2 < 0) s2 = 0;
  s1 = std::sqrt(s1);
  s2 = std::sqrt(s2);
  cc = cc / (s1 * s2 + 1e-30);
  if (cc < -1) cc = -1;
  else if (cc > 1) cc = 1;
  return (float)cc;
}

float offset_correlate(const float* ref, const float* img, int w, int h,
                       float ref_tilt, float tilt) {
  double m[6];
  offset_moments(ref, img, w, h, ref_tilt, tilt, m);
  return offset_cc(m);
}

// SearchOffset + CalcAverangedCC (FindOffset.cu:129-174) over a bin-4
// corrected stack [nz][h][w].  The tilt copy drifts by the float rounding of
// += offset / -= offset exactly as the reference's member copy does.
float offset_search(const float* corr, int w, int h, int nz,
                    const float* tilts_in, float* ccs) {
  std::vector<float> ang(tilts_in, tilts_in + nz);
  const int num = 31;
  const float step = 1.0f, init = 0.0f;
  float maxcc = 0.0f, best = 0.0f;
  const size_t plane = (size_t)w * h;
  for (int i = 0; i < num; i++) {
    const float off = init + (float)(i - num / 2) * step;
    for (int z = 0; z < nz; z++) ang[z] += off;
    const int z0 = zero_tilt_index(ang.data(), nz, 0.0f);
    std::vector<float> cc(nz, 0.0f);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nz; k++) {
      if (k == z0) continue;
      int r = (k < z0) ? k + 1 : k - 1;
      cc[k] = offset_correlate(corr + r * plane, corr + k * plane, w, h,
                               ang[r], ang[k]);
    }
    float sum = 0.0f;
    int count = 0;
    for (int k = 0; k < nz; k++) {
      if (k == z0) continue;
      sum += cc[k];
      count++;
    }
    const float mean = sum / (float)count;
    for (int z = 0; z < nz; z++) ang[z] -= off;
    if (ccs) ccs[i] = mean;
    if (mean > maxcc) {
      maxcc = mean;
      best = off;
    }
  }
  return best;
}

}  // namespace orc

using namespace orc;

extern "C" float orc_find_offset(const float* stack, int nx, int ny, int nz,
                                 const float* tilts, const float* shift_x,
                                 const float* shift_y, const float* rotate,
                                 float* ccs) {
  // FOsetup (FindOffset.cu:98-127): bin 4 (quirk Q8), rotation, no random
  // fill, no R-weighting.
  std::vector<float> corr;
  CorrSizes s;
  correct_stack(stack, nx, ny, nz, 4, shift_x, shift_y, rotate, false, false,
                false, corr, &s);
  return offset_search(corr.data(), s.bx, s.by, nz, tilts, ccs);
}

extern "C" float orc_offset_correlate(const float* ref, const float* img,
                                      int w, int h, float ref_tilt,
                                      float tilt) {
  return offset_correlate(ref, img, w, h, ref_tilt, tilt);
}

extern "C" void orc_offset_moments(const float* ref, const float* img, int w,
                                   int h, float ref_tilt, float tilt,
                                   double* m6) {
  offset_moments(ref, img, w, h, ref_tilt, tilt, m6);
}

extern "C" float orc_offset_scan(const float* corr, int w, int h, int nz,
                                 const float* tilts, float* ccs) {
  return offset_search(corr, w, h, nz, tilts, ccs);
}
