// CPU restatement of the stretched cross-correlation (CalcCC.cu) and of the
// projection-matching measurement (ProjAlign.cpp:243-375) -- test
// infrastructure.
#include "oracle.h"
#include "orc_internal.h"
#include "orc_stages.h"

namespace orc {

void stretch_matrix(float stretch, float axis_deg, float m[3]) {
  // CalcCC::Stretch (CalcCC.cu:318-345).
  double d2t = 2 * kD2R * axis_deg;
  double s2 = std::sin(d2t), c2 = std::cos(d2t);
  double P = 0.5 * ((double)stretch + 1), M = 0.5 * ((double)stretch - 1);
  float a0 = (float)(P + M * c2);
  float a1 = (float)(1 * M * s2);
  float a2 = (float)(P - M * c2);
  float det = a0 * a2 - a1 * a1;
  m[0] = a2 / det;
  m[1] = -a1 / det;
  m[2] = a0 / det;
}

namespace {

// Grandomfill (CalcCC.cu:23-46): window 31, next = 7 next mod 961; the
// offset uses next mod 961 (quirk Q9) and the row width is the padded pitch.
float random_fill_cc(float fx, float fy, int pitch, int h, int w,
                     const float* in) {
  int x = (int)std::fabs(fx);
  int y = (int)std::fabs(fy);
  if (x >= pitch) x = pitch - 1 - (x % (pitch - 1));
  if (y >= h) y = h - 1 - (y % (h - 1));
  const int win = 31, size = win * win;
  uint32_t next = (uint32_t)(y * pitch + x);
  for (int i = 0; i < size; i++) {
    next = (next * 7u) % (uint32_t)size;
    int ix = (int)(next % size) - win / 2 + x;
    if (ix < 0 || ix >= pitch) continue;
    int iy = (int)(next / win) - win / 2 + y;
    if (iy < 0 || iy >= h) continue;
    return in[(size_t)iy * pitch + ix];
  }
  return in[(size_t)(h / 2) * pitch + w / 2];
}

// GStretchRandom (CalcCC.cu:149-171) over a w x h image with row pitch
// 'pitch'; the output has the same pitch.
void stretch_image(const float* in, int w, int h, int pitch, const float m[3],
                   bool fill, float* out) {
  const float hw = 0.5f * (float)w, hh = 0.5f * (float)h;
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      float X = (float)x - hw + 0.5f;
      float Y = (float)((double)y - 0.5 * h + 0.5);
      float buf = X * m[0] + Y * m[1];
      float Y2 = X * m[1] + Y * m[2] + hh - 0.5f;
      float X2 = buf + hw - 0.5f;
      float v;
      if (X2 >= 0 && X2 < (float)(w - 1) && Y2 >= 0 && Y2 < (float)(h - 1)) {
        // mGBilinear (CalcCC.cu:8-21).
        int ix = (int)X2, iy = (int)Y2;
        size_t i = (size_t)iy * pitch + ix;
        float fx = X2 - (float)ix, fy = Y2 - (float)iy;
        float f2 = 1.0f - fx, f3 = 1.0f - fy;
        v = in[i] * f2 * f3 + in[i + 1] * fx * f3 + in[i + pitch] * f2 * fy +
            in[i + pitch + 1] * fx * fy;
      } else if (fill) {
        v = random_fill_cc(X2, Y2, pitch, h, w, in);
      } else {
        v = invalid();
      }
      out[(size_t)y * pitch + x] = v;
    }
  }
}

// CalcMoment (CalcCC.cu:48-77, 295-308): sum of v^e over valid pixels
// divided by w*h (quirk Q23: invalid pixels count in the divisor).  FP64
// accumulation.
float moment(const float* img, int w, int h, int pitch, int e) {
  double s = 0.0;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float v = img[(size_t)y * pitch + x];
      if (v < -1e10f) continue;
      float ev = v;
      for (int k = 1; k < e; k++) ev *= v;
      s += ev;
    }
  return (float)(s / ((double)w * h));
}

// GNormalize (CalcCC.cu:137-147).
void normalize(float* img, int w, int h, int pitch, float mean, float sd) {
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float& v = img[(size_t)y * pitch + x];
      if (v < -1e10f) continue;
      v = (v - mean) / sd;
    }
}

// CalcCC::mNormalize (CalcCC.cu:284-293).
void mnormalize(float* img, int w, int h, int pitch) {
  float m = moment(img, w, h, pitch, 1);
  float m2 = moment(img, w, h, pitch, 2);
  float var = m2 - m * m;
  float sd = var <= 0 ? 0.0f : std::sqrt(var);
  normalize(img, w, h, pitch, m, sd);
}

// GRoundEdge (CalcCC.cu:173-195).
void round_edge(float* img, int w, int h, int pitch, float cx, float cy,
                float sx, float sy, float power) {
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float& v = img[(size_t)y * pitch + x];
      if (v < -1e10f) {
        v = 0.0f;
        continue;
      }
      float fx = 2 * std::fabs((float)x - cx) / sx;
      float fy = 2 * std::fabs((float)y - cy) / sy;
      float fr = std::sqrt(fx * fx + fy * fy);
      if (fr >= 1.0f) {
        v = 0.0f;
        continue;
      }
      fr = 0.5f * (1 - fcos(3.1415926f * fr));
      fr = 1.0f - (float)std::pow((double)fr, (double)power);
      v = v * fr;
    }
}

// GConv + GWiener + GCenterOrigin (CalcCC.cu:95-135, 356-364) on spectra
// of (w/2+1) x h complex values stored as interleaved floats.
void spectral_cc(const float* ref, float* img, int w, int h, float bfactor) {
  const int nc = w / 2 + 1;
  const float fnx = 2.0f * (float)(nc - 1);
  const float filt0 = -2.0f * bfactor / (fnx * fnx + (float)(h * h));
  for (int y = 0; y < h; y++) {
    int yy = (y > h / 2) ? y - h : y;
    for (int k = 0; k < nc; k++) {
      size_t i = 2 * ((size_t)y * nc + k);
      float re = ref[i] * img[i] + ref[i + 1] * img[i + 1];
      float im = ref[i] * img[i + 1] - ref[i + 1] * img[i];
      float f = (float)std::exp((double)(filt0 * (float)(unsigned)(k * k + yy * yy)));
      float amp = std::sqrt(re * re + im * im);
      f /= std::sqrt(amp + 0.01f);
      re *= f;
      im *= f;
      if ((k + y) % 2 != 0) {
        re = -re;
        im = -im;
      }
      img[i] = re;
      img[i + 1] = im;
    }
  }
}

// Spectrum of a padded real image (in place, the reference's cuFFT R2C).
void r2c(std::vector<float>& pad, int w, int h, float scale) {
  orc_fft_r2c_2d(pad.data(), w, h);
  if (scale != 1.0f)
    for (int y = 0; y < h; y++)
      for (int k = 0; k < 2 * (w / 2 + 1); k++) pad[(size_t)y * pad_x(w) + k] *= scale;
}

}  // namespace

void stretch_plain(const float* in, int w, int h, float stretch,
                   float axis_deg, float* out) {
  float m[3];
  stretch_matrix(stretch, axis_deg, m);
  stretch_image(in, w, h, w, m, false, out);
}

// Peak search and parabolic sub-pixel fit on a w x h correlation map with
// row pitch 'pitch'.  mode 0: CalcCC::FindPeak (CalcCC.cu:381-435), the
// central window of w*8/20*2 x h*8/20*2 pixels, FP32 fit terms, the integer
// peak clamped to [1, n-2].  mode 1: ProjAlign::PAFindPeak
// (ProjAlign.cpp:325-375), the window [max(n/20, 3), n - that), FP64 fit
// terms.  In both the first maximum in row-major order wins.  shift2 = peak
// - (w/2, h/2).
void find_peak(const float* cc, int w, int h, int pitch, int mode,
               float* shift2) {
  int x0, y0, x1, y1;
  float best;
  if (mode == 0) {
    const int sx = w * 8 / 20 * 2, sy = h * 8 / 20 * 2;
    x0 = (w - sx) / 2;
    y0 = (h - sy) / 2;
    x1 = x0 + sx;
    y1 = y0 + sy;
    best = -1e30f;
  } else {
    x0 = std::max(w / 20, 3);
    y0 = std::max(h / 20, 3);
    x1 = w - x0;
    y1 = h - y0;
    best = -1e20f;
  }
  int px = 0, py = 0;
  for (int y = y0; y < y1; y++)
    for (int x = x0; x < x1; x++) {
      const float v = cc[(size_t)y * pitch + x];
      if (best >= v) continue;
      px = x;
      py = y;
      best = v;
    }
  auto at = [&](int x, int y) { return cc[(size_t)y * pitch + x]; };
  double A, B, C, D;
  if (mode == 0) {
    if (px < 1) px = 1;
    else if (px > w - 2) px = w - 2;
    if (py < 1) py = 1;
    else if (py > h - 2) py = h - 2;
    A = (at(px + 1, py) + at(px - 1, py)) * 0.5f - at(px, py);
    B = (at(px + 1, py) - at(px - 1, py)) * 0.5f;
    C = (at(px, py + 1) + at(px, py - 1)) * 0.5f - at(px, py);
    D = (at(px, py + 1) - at(px, py - 1)) * 0.5f;
  } else {
    // The reference reads outside the map when nothing in the window beats
    // -1e20 (peak left at 0, 0); the window's corner is taken instead.
    if (best == -1e20f) {
      px = x0;
      py = y0;
    }
    A = (at(px + 1, py) + at(px - 1, py)) * 0.5 - at(px, py);
    B = (at(px + 1, py) - at(px - 1, py)) * 0.5;
    C = (at(px, py + 1) + at(px, py - 1)) * 0.5 - at(px, py);
    D = (at(px, py + 1) - at(px, py - 1)) * 0.5;
  }
  double dx = -B / (2 * A + 1e-30), dy = -D / (2 * C + 1e-30);
  if (std::fabs(dx) > 1) dx = 0;
  if (std::fabs(dy) > 1) dy = 0;
  const float fpx = (float)(px + dx), fpy = (float)(py + dy);
  shift2[0] = fpx - (float)(w / 2);
  shift2[1] = fpy - (float)(h / 2);
}

void stretch_cc(const float* ref, const float* img, int w, int h,
                float ref_tilt, float tilt, float axis_deg, float bfactor,
                float* shift2) {
  const int pitch = pad_x(w);
  std::vector<float> pref((size_t)pitch * h, 0.0f), pimg((size_t)pitch * h, 0.0f),
      pst((size_t)pitch * h, 0.0f);
  for (int y = 0; y < h; y++) {
    std::memcpy(&pref[(size_t)y * pitch], ref + (size_t)y * w, 4 * w);
    std::memcpy(&pimg[(size_t)y * pitch], img + (size_t)y * w, 4 * w);
  }
  mnormalize(pref.data(), w, h, pitch);
  double ds = std::cos(kD2R * ref_tilt) / std::cos(kD2R * tilt);
  float m[3];
  stretch_matrix((float)ds, axis_deg, m);
  stretch_image(pimg.data(), w, h, pitch, m, true, pst.data());
  mnormalize(pst.data(), w, h, pitch);
  round_edge(pref.data(), w, h, pitch, w * 0.5f, h * 0.5f, w * 1.0f, h * 1.0f, 4);
  round_edge(pst.data(), w, h, pitch, w * 0.5f, h * 0.5f, w * 1.0f, h * 1.0f, 4);
  const float norm = (float)(1.0 / w / h);
  r2c(pref, w, h, norm);
  r2c(pst, w, h, norm);
  spectral_cc(pref.data(), pst.data(), w, h, bfactor);
  orc_fft_c2r_2d(pst.data(), w, h);
  find_peak(pst.data(), w, h, pitch, 0, shift2);
}

void pa_measure(const float* reproj, const float* proj, int w, int h,
                float* shift2) {
  const int pitch = pad_x(w);
  std::vector<float> a((size_t)pitch * h, 0.0f), b((size_t)pitch * h, 0.0f);
  // PANorm (ProjAlign.cpp:280-312): centre copy, subtract the mean (std
  // fixed to 1), full-size elliptical mask of pitch x h (quirk Q14).
  auto panorm = [&](const float* src, std::vector<float>& dst) {
    for (int y = 0; y < h; y++)
      std::memcpy(&dst[(size_t)y * pitch], src + (size_t)y * w, 4 * w);
    float mean = moment(dst.data(), w, h, pitch, 1);
    normalize(dst.data(), w, h, pitch, mean, 1.0f);
    round_edge(dst.data(), w, h, pitch, (float)w * 0.5f, (float)h * 0.5f,
               (float)pitch, (float)h, 4.0f);
  };
  panorm(reproj, a);
  panorm(proj, b);
  r2c(a, w, h, 1.0f);
  r2c(b, w, h, 1.0f);
  spectral_cc(a.data(), b.data(), w, h, 500.0f);
  orc_fft_c2r_2d(b.data(), w, h);
  find_peak(b.data(), w, h, pitch, 1, shift2);
}

}  // namespace orc

using namespace orc;

extern "C" void orc_stretch_cc(const float* ref, const float* img, int bx,
                               int by, float ref_tilt, float tilt,
                               float axis_deg, float bfactor, float* shift2) {
  stretch_cc(ref, img, bx, by, ref_tilt, tilt, axis_deg, bfactor, shift2);
}

extern "C" void orc_pa_measure(const float* reproj, const float* proj, int bx,
                               int by, float* shift2) {
  pa_measure(reproj, proj, bx, by, shift2);
}

extern "C" void orc_peak2d(const float* map, int w, int h, int pitch, int mode,
                           float* shift2) {
  find_peak(map, w, h, pitch, mode, shift2);
}
