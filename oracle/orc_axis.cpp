// CPU restatement of the common-line tilt-axis search (Rotate.cu) -- test
// infrastructure.
#include "oracle.h"
#include "orc_internal.h"

namespace orc {

struct AxisSetup {
  std::vector<float> rot;   // candidate angles (deg)
  std::vector<float> cosr, sinr;  // cos/sin of rot*3.1415926f/180.0f
  int L;
};

// CLsetup (Rotate.cu:339-374).
AxisSetup axis_setup(int nx, int ny, float axis0, float range, int num) {
  AxisSetup a;
  float step = range / (float)(num - 1);
  a.rot.resize(num);
  a.rot[0] = axis0 - step * (float)num / 2.0f;
  for (int i = 1; i < num; i++) a.rot[i] = a.rot[i - 1] + step;
  float fmin = 1e20f;
  for (int i = 0; i < num; i++) {
    float fs = (float)std::fabs(std::sin(kD2R * a.rot[i]));
    float fc = (float)std::fabs(std::cos(kD2R * a.rot[i]));
    float s1 = (float)nx / (fs + 0.000001f);
    float s2 = (float)ny / (fc + 0.000001f);
    if (fmin > s1) fmin = s1;
    if (fmin > s2) fmin = s2;
  }
  int L = (int)fmin;
  a.L = L / 2 * 2 - 100;
  a.cosr.resize(num);
  a.sinr.resize(num);
  for (int i = 0; i < num; i++) {
    float r = a.rot[i] * 3.1415926f / 180.0f;
    a.cosr[i] = fcos(r);
    a.sinr[i] = fsin(r);
  }
  return a;
}

// GCalcCommonRegion (Rotate.cu:9-40): last in-image offset along the line
// normal on each side of the zero-tilt centre; entries a candidate never
// reaches stay 0 (the reference leaves them uninitialised).
void common_region(const AxisSetup& a, int nx, int ny, float sx, float sy,
                   std::vector<int>& reg) {
  const int num = (int)a.rot.size(), L = a.L, half = L / 2;
  reg.assign((size_t)num * L * 2, 0);
  const float ox = (float)nx * 0.5f + sx, oy = (float)ny * 0.5f + sy;
  for (int k = 0; k < num; k++) {
    const float c = a.cosr[k], s = a.sinr[k];
    for (int y = 0; y < L; y++) {
      const float fy = (float)y - (float)L * 0.5f;
      int* r = &reg[2 * ((size_t)k * L + y)];
      for (int x = 1; x < half; x++) {
        float X = (float)(-x) * c - fy * s + ox;
        float Y = (float)(-x) * s + fy * c + oy;
        if (X < 0 || X >= (float)nx) break;
        if (Y < 0 || Y >= (float)ny) break;
        r[0] = -x;
      }
      for (int x = 1; x < half; x++) {
        float X = (float)x * c - fy * s + ox;
        float Y = (float)x * s + fy * c + oy;
        if (X < 0 || X >= (float)nx) break;
        if (Y < 0 || Y >= (float)ny) break;
        r[1] = x;
      }
    }
  }
}

// Taper of GRemoveMean (Rotate.cu:144-157): (1 - (0.5(1-cos(pi r)))^100)^2.
std::vector<float> line_taper(int L) {
  std::vector<float> t(L);
  const float half = 0.5f * (float)L;
  for (int i = 0; i < L; i++) {
    float r = std::fabs((float)i - half) / half;
    r = 0.5f * (1 - fcos((float)(3.14159 * (double)r)));
    r = 1.0f - (float)std::pow((double)r, 100.0);
    t[i] = r * r;
  }
  return t;
}

// GRadon (Rotate.cu:42-95) of one image: the mean line values of candidate
// k, L values into line.
void radon_lines(const AxisSetup& a, const std::vector<int>& reg,
                 const float* img, int nx, int ny, float tilt, float sx,
                 float sy, int k, float* line) {
  const int L = a.L;
  const float ct = (float)std::cos(tilt * kD2R);
  const float ox = (float)nx * 0.5f + sx, oy = (float)ny * 0.5f + sy;
  {
    const float c = a.cosr[k], s = a.sinr[k];
    for (int y = 0; y < L; y++) {
      const float fy = (float)y - (float)L * 0.5f;
      const int* r = &reg[2 * ((size_t)k * L + y)];
      float sum = 0.0f;
      int cnt = 0;
      auto sample = [&](int x) {
        float X = (float)x * c - fy * s + ox;
        float Y = (float)x * s + fy * c + oy;
        if (X < 0 || X >= (float)nx) return;
        if (Y < 0 || Y >= (float)ny) return;
        float v = img[(size_t)(int)Y * nx + (int)X];
        if (v < 0) return;
        sum += v;
        cnt++;
      };
      const int left = (int)((float)r[0] * ct + 0.5f);
      for (int x = left; x < 0; x++) sample(x);
      const int right = (int)((float)r[1] * ct + 0.5f);
      for (int x = 0; x < right; x++) sample(x);
      line[y] = cnt == 0 ? invalid() : sum / (float)cnt;
    }
  }
}

// Radon + FFT1d for one image (Rotate.cu:397-451): one spectrum of L/2+1
// complex values per candidate.
void image_lines(const AxisSetup& a, const std::vector<int>& reg,
                 const float* img, int nx, int ny, float tilt, float sx,
                 float sy, const std::vector<float>& taper, float* spec) {
  const int num = (int)a.rot.size(), L = a.L, nc = L / 2 + 1;
  std::vector<float> line(L);
  for (int k = 0; k < num; k++) {
    radon_lines(a, reg, img, nx, ny, tilt, sx, sy, k, line.data());
    // GCalcMean + GRemoveMean (Rotate.cu:97-157).
    double ms = 0.0;
    int mc = 0;
    for (int y = 0; y < L; y++)
      if (line[y] > -1e10f) {
        ms += line[y];
        mc++;
      }
    const float mean = mc == 0 ? 0.0f : (float)(ms / mc);
    for (int y = 0; y < L; y++) {
      float v = line[y] < 0 ? 0.0f : line[y] - mean;
      line[y] = v * taper[y];
    }
    fft_r2c_line(line.data(), L, spec + (size_t)2 * k * nc);
  }
}

// Radon + FFT1d of every image (Rotate.cu:397-451): spec [nz][num][L/2+1]
// complex values; returns L.
int line_spectra(const float* stack, int nx, int ny, int nz,
                 const float* tilts, const float* shift_x,
                 const float* shift_y, float axis0, float range, int num,
                 std::vector<float>& spec) {
  const int z0 = zero_tilt_index(tilts, nz, 0.0f);
  AxisSetup a = axis_setup(nx, ny, axis0, range, num);
  const int L = a.L, nc = L / 2 + 1;
  std::vector<int> reg;
  common_region(a, nx, ny, shift_x[z0], shift_y[z0], reg);
  const std::vector<float> taper = line_taper(L);
  spec.assign((size_t)nz * num * nc * 2, 0.0f);
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < nz; z++)
    image_lines(a, reg, stack + (size_t)z * nx * ny, nx, ny, tilts[z],
                shift_x[z], shift_y[z], taper,
                spec.data() + (size_t)z * num * nc * 2);
  return L;
}

}  // namespace orc

using namespace orc;

extern "C" int orc_common_lines(const float* stack, int nx, int ny, int nz,
                                const float* tilts, const float* shift_x,
                                const float* shift_y, float axis0, float range,
                                int num, float* lines) {
  const int z0 = zero_tilt_index(tilts, nz, 0.0f);
  AxisSetup a = axis_setup(nx, ny, axis0, range, num);
  if (!lines) return a.L;
  std::vector<int> reg;
  common_region(a, nx, ny, shift_x[z0], shift_y[z0], reg);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int z = 0; z < nz; z++)
    for (int k = 0; k < num; k++)
      radon_lines(a, reg, stack + (size_t)z * nx * ny, nx, ny, tilts[z],
                  shift_x[z], shift_y[z], k,
                  lines + ((size_t)z * num + k) * a.L);
  return a.L;
}

extern "C" int orc_line_spectra(const float* stack, int nx, int ny, int nz,
                                const float* tilts, const float* shift_x,
                                const float* shift_y, float axis0, float range,
                                int num, float* spec) {
  if (!spec) return axis_setup(nx, ny, axis0, range, num).L;
  std::vector<float> v;
  const int L = line_spectra(stack, nx, ny, nz, tilts, shift_x, shift_y,
                              axis0, range, num, v);
  std::memcpy(spec, v.data(), v.size() * sizeof(float));
  return L;
}

extern "C" float orc_tilt_axis(const float* stack, int nx, int ny, int nz,
                               const float* tilts, const float* shift_x,
                               const float* shift_y, const float* rotate,
                               float range, int num, float* scores,
                               int* line_size) {
  const int z0 = zero_tilt_index(tilts, nz, 0.0f);
  AxisSetup a = axis_setup(nx, ny, rotate[z0], range, num);
  std::vector<float> spec;
  const int L = line_spectra(stack, nx, ny, nz, tilts, shift_x, shift_y,
                             rotate[z0], range, num, spec);
  const int nc = L / 2 + 1;
  if (line_size) *line_size = L;
  // CalcScore / GCC1d (Rotate.cu:489-561); filter exp(-20 (j/(2(n-1)))^2).
  std::vector<float> filt(nc);
  for (int j = 1; j < nc; j++) {
    float f = (float)j / (2.0f * (float)(nc - 1));
    filt[j] = (float)std::exp((double)(-2.0f * 10.0f * f * f));
  }
  double fsum = 0.0;
  for (int j = 1; j < nc; j++) fsum += filt[j];
  std::vector<float> sc(num);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < num; k++) {
    std::vector<float> S(2 * nc, 0.0f);
    for (int z = 0; z < nz; z++) {
      const float* l = &spec[((size_t)z * num + k) * nc * 2];
      for (int j = 0; j < 2 * nc; j++) S[j] = S[j] + l[j];
    }
    float ccsum = 0.0f;
    for (int z = 0; z < nz; z++) {
      const float* l = &spec[((size_t)z * num + k) * nc * 2];
      double cc = 0.0;
      for (int j = 1; j < nc; j++) {
        float rx = S[2 * j] - l[2 * j], ry = S[2 * j + 1] - l[2 * j + 1];
        float a1 = std::sqrt(rx * rx + ry * ry);
        float a2 = std::sqrt(l[2 * j] * l[2 * j] + l[2 * j + 1] * l[2 * j + 1]);
        float v = (l[2 * j] * rx + l[2 * j + 1] * ry) * filt[j] /
                  (a1 * a2 + 1e-20f);
        cc += v;
      }
      ccsum += (float)(cc / fsum);
    }
    sc[k] = ccsum / (float)nz;
  }
  // findTiltAxis (Rotate.cu:468-487): strict max from an initial 0.
  int best = 0;
  float bs = 0.0f;
  for (int k = 0; k < num; k++) {
    if (scores) scores[k] = sc[k];
    if (bs < sc[k]) {
      bs = sc[k];
      best = k;
    }
  }
  float axis = a.rot[best];
  // DoIt wrap rules (Rotate.cu:316-322).
  if (axis < -45 && rotate[0] == 0) axis += 180.0f;
  if (axis > 135) axis -= 180.0f;
  return axis;
}
