// CPU restatement of CalcReprojFBP (CalcReproj.cu) -- test infrastructure.
#include "oracle.h"
#include "orc_internal.h"
#include <algorithm>
#include <omp.h>

using namespace orc;

namespace orc {

int zero_tilt_index(const float* tilts, int nz, float tilt) {
  // Util.cpp:8-20: smallest |tilt - a_i|, first one wins.
  int idx = 0;
  float fmin = std::fabs(tilt - tilts[0]);
  for (int i = 1; i < nz; i++) {
    float d = std::fabs(tilt - tilts[i]);
    if (d >= fmin) continue;
    fmin = d;
    idx = i;
  }
  return idx;
}

int find_ref_index(const float* tilts, int nz, int z) {
  // Util.cpp:22-35: neighbour with smaller |tilt|, z-1 checked first.
  int i0 = z - 1, i2 = z + 1;
  if (i0 < 0) i0 = z;
  if (i2 >= nz) i2 = z;
  double t0 = std::fabs((double)tilts[i0]);
  double t = std::fabs((double)tilts[z]);
  double t2 = std::fabs((double)tilts[i2]);
  if (t0 < t) return i0;
  if (t2 < t) return i2;
  return z;
}

}  // namespace orc

extern "C" void orc_find_proj_range(const float* tilts, int nz,
                                    const unsigned char* skip, int target,
                                    int nproj, int* range2) {
  // CalcReproj.cu:206-259.
  const float ref_range = 20.5f, ref_stretch = 1.2f;
  float tilt = tilts[target];
  float fc = (float)std::cos(tilt * kD2R);
  int istart = -1;
  for (int i = 0; i < nz; i++) {
    if (skip[i]) continue;
    float ta = tilts[i];
    float d = tilt - ta;
    if (std::fabs(d) > ref_range) continue;
    float st = (float)(std::cos(ta * kD2R) / fc);
    if (st > ref_stretch) continue;
    istart = i;
    break;
  }
  int iend = -1;
  // The reference starts this loop at istart even when it is -1 (an
  // out-of-bounds read); no projection qualifies then, so the range is empty.
  for (int i = std::max(istart, 0); istart >= 0 && i < nz; i++) {
    if (skip[i]) continue;
    float ta = tilts[i];
    float d = tilt - ta;
    if (std::fabs(d) > ref_range) continue;
    float st = (float)(std::cos(ta * kD2R) / fc);
    if (st > ref_stretch) continue;
    iend = i;
  }
  int np = nproj - 1;
  if (istart >= 0 && (iend - istart) > np) {
    int z0 = zero_tilt_index(tilts, nz, 0.0f);
    if (target < z0)
      iend = istart + np;
    else
      istart = iend - np;
  }
  range2[0] = istart;
  range2[1] = iend;
}

namespace {

// Backprojection of one image row into the xz slab (CalcReproj.cu:26-64).
// rho = fma(x, cos, fma(z, sin, bx/2)) is the FP32 evaluation order shared
// with the HIP kernel; the reference leaves the contraction to nvcc
// (unpinned).
template <typename Acc>
void backproject_row(const float* stack, int bx, int by, int tb, int y,
                     int i0, int i1, const float* tw, const float* tc,
                     const float* ts, float* vol) {
  const int vx = 2 * bx;
  const float cx = bx * 0.5f;
  const float iend = (float)(bx - 1);
  const size_t plane = (size_t)bx * by;
  for (int iz = 0; iz < tb; iz++) {
    float fz = (float)iz + 0.5f - (float)tb * 0.5f;
    for (int ix = 0; ix < vx; ix++) {
      float fx = (float)ix + 0.5f - (float)vx * 0.5f;
      Acc sum = 0, cnt = 0;
      for (int i = i0; i <= i1; i++) {
        float c = tc[i - i0], s = ts[i - i0], w = tw[i - i0];
        float rho = std::fmaf(fx, c, std::fmaf(fz, s, cx));
        if (rho < 0.0f || rho > iend) continue;
        float v = stack[(size_t)i * plane + (size_t)y * bx + (int)rho];
        if (v < -1e10f) continue;  // mDIntProj -> -1e30 -> skipped
        if (sizeof(Acc) == sizeof(float))
          // fInt += v*cos*w as nvcc contracts it: fma(v*cos, w, fInt).
          sum = (Acc)std::fmaf(v * c, w, (float)sum);
        else
          sum += (Acc)v * (Acc)c * (Acc)w;
        cnt += (Acc)w;
      }
      vol[(size_t)iz * vx + ix] =
          (cnt < (Acc)0.001f) ? invalid() : (float)(sum / cnt);
    }
  }
}

// Ray-sum reprojection of one slab (CalcReproj.cu:68-124).
template <typename Acc>
void reproject_row(const float* vol, int bx, int tb, float c, float s, int L,
                   float* line) {
  const int vx = 2 * bx;
  const float iendx = (float)(vx - 1), iendz = (float)(tb - 1);
  for (int u = 0; u < bx; u++) {
    float xp = (float)u + 0.5f - (float)bx * 0.5f;
    float tx = std::fmaf(xp, c, (float)vx * 0.5f);
    float tz = std::fmaf(xp, s, (float)tb * 0.5f);
    float z0 = (-xp * s) / c - 0.5f * (float)L;
    Acc sum = 0;
    int cnt = 0;
    for (int t = 0; t < L; t++) {
      float fz = (float)t + z0;
      float X = std::fmaf(-fz, s, tx);
      float Z = std::fmaf(fz, c, tz);
      if (X >= 0.0f && X < iendx && Z >= 0.0f && Z < iendz) {
        float v = vol[(size_t)vx * (int)Z + (int)X];
        if (v >= -1e10f) {
          sum += (Acc)v;
          cnt++;
        }
      }
    }
    if ((float)cnt < 0.8f * (float)L)
      line[u] = invalid();
    else
      line[u] = (float)(sum / (Acc)cnt);
  }
}

}  // namespace

extern "C" void orc_reproject(const float* stack, int bx, int by, int nz,
                              int tb, const float* tilts,
                              const unsigned char* skip, int target, int nproj,
                              int acc64, float* reproj, float* slab_out) {
  int range[2];
  orc_find_proj_range(tilts, nz, skip, target, nproj, range);
  const float proj_angle = tilts[target];
  int n = (range[0] >= 0) ? range[1] - range[0] + 1 : 0;
  std::vector<float> tw(std::max(n, 1)), tc(std::max(n, 1)),
      ts(std::max(n, 1));
  for (int k = 0; k < n; k++) {
    float a = tilts[range[0] + k];
    tw[k] = fcos((proj_angle - a) * kBPD2R);
    float v = a * kBPD2R;
    tc[k] = fcos(v);
    ts[k] = fsin(v);
  }
  // ForwardProj (CalcReproj.cu:299-311): exact D2R, host cos/sin.
  float c = (float)std::cos(kD2R * proj_angle);
  float s = (float)std::sin(kD2R * proj_angle);
  int L = (int)((float)tb / c + 0.5f);
  const size_t vsz = (size_t)2 * bx * tb;
#pragma omp parallel
  {
    std::vector<float> vol(vsz);
#pragma omp for schedule(dynamic, 4)
    for (int y = 0; y < by; y++) {
      if (n == 0) {
        std::fill(vol.begin(), vol.end(), invalid());
      } else if (acc64) {
        backproject_row<double>(stack, bx, by, tb, y, range[0], range[1],
                                tw.data(), tc.data(), ts.data(), vol.data());
      } else {
        backproject_row<float>(stack, bx, by, tb, y, range[0], range[1],
                               tw.data(), tc.data(), ts.data(), vol.data());
      }
      if (slab_out)
        std::memcpy(slab_out + (size_t)y * vsz, vol.data(), vsz * 4);
      // Ray sums: FP32 in increasing sample order in the FP32 mode (the
      // reference sums in FP32 over a 512-thread tree, CalcReproj.cu:94-117,
      // whose order is unpinned), FP64 in the FP64 mode.
      if (acc64)
        reproject_row<double>(vol.data(), bx, tb, c, s, L,
                              reproj + (size_t)y * bx);
      else
        reproject_row<float>(vol.data(), bx, tb, c, s, L,
                             reproj + (size_t)y * bx);
    }
  }
}

extern "C" long long orc_rp_inbounds_samples(int bx, int tb,
                                             float target_tilt) {
  float c = (float)std::cos(kD2R * target_tilt);
  float s = (float)std::sin(kD2R * target_tilt);
  int L = (int)((float)tb / c + 0.5f);
  const int vx = 2 * bx;
  const float iendx = (float)(vx - 1), iendz = (float)(tb - 1);
  long long total = 0;
  for (int u = 0; u < bx; u++) {
    float xp = (float)u + 0.5f - (float)bx * 0.5f;
    float tx = std::fmaf(xp, c, (float)vx * 0.5f);
    float tz = std::fmaf(xp, s, (float)tb * 0.5f);
    float z0 = (-xp * s) / c - 0.5f * (float)L;
    for (int t = 0; t < L; t++) {
      float fz = (float)t + z0;
      float X = std::fmaf(-fz, s, tx);
      float Z = std::fmaf(fz, c, tz);
      if (X >= 0.0f && X < iendx && Z >= 0.0f && Z < iendz) total++;
    }
  }
  return total;
}

extern "C" int orc_num_threads(void) { return omp_get_max_threads(); }
extern "C" void orc_set_num_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
