// Fiducial residual evaluator: how well an alignment explains the tracks of
// point-like markers (host C++, part of libmf_io).
//
// Counterpart of the reference's accuracy tool cal_res
// (tool/src/main/main.cpp:35-237), which reads a Markerfree .txt alignment,
// detects gold markers in the stack, tracks them and triangulates each track
// under the alignment (PPBundleApp::Triangulate,
// tool/src/nbundle/bundle_core.cpp:149-216: per view, the rotated and
// shifted observation equals the projection of the marker's 3-D position at
// the tilt angle, solved by least squares; the error is the RMS reprojection
// distance).  Marker detection and matching (OpenCV templates) are outside
// this build: the tracks come in as coordinates.
//
// Model, in the frame the aligned stack is written in: an observation q (raw
// pixel position, pixel k at position k) of section j maps to
//   p = R(-rot_j) (q - c - s_j),   c = (nx/2, ny/2),
// the inverse of the stack correction (mGCorrect, method/CorrStack.cu:142-179).
// The marker at (X, Y, Z) projects at tilt b_j to
//   p_x = X cos b_j + Z sin b_j + d,   p_y = Y,
// the geometry of the backprojection (method/CalcReproj.cu:37-50), where d is
// one offset common to all sections and markers: the alignment leaves the
// rotation centre's x free (a constant x offset of every section moves the
// axis, not the markers), so d is fitted with the markers.  The tool's
// Triangulate writes the same equations with Z of opposite sign.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "markerfree/mf_io.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Obs {
  int sec;
  double px, py;
};

// Least squares of p_x = X c + Z s + d over one marker's views for fixed d:
// returns the residual vector's squared norm and (X, Z); also the two
// quantities the closed-form d needs: 1'(I - H)p and 1'(I - H)1.
struct Fit {
  bool ok = false;
  double X = 0, Z = 0, Y = 0;
  double a = 0, b = 0;  // 1'(I-H)p, 1'(I-H)1
};

Fit fit_marker(const std::vector<Obs>& v, const std::vector<double>& cb,
               const std::vector<double>& sb) {
  Fit f;
  double scc = 0, scs = 0, sss = 0, scp = 0, ssp = 0, sc = 0, ss = 0, sp = 0,
         sy = 0;
  const double n = (double)v.size();
  for (const Obs& o : v) {
    const double c = cb[o.sec], s = sb[o.sec];
    scc += c * c;
    scs += c * s;
    sss += s * s;
    scp += c * o.px;
    ssp += s * o.px;
    sc += c;
    ss += s;
    sp += o.px;
    sy += o.py;
  }
  const double det = scc * sss - scs * scs;
  if (v.size() < 2 || std::fabs(det) < 1e-12 * (scc + sss) * (scc + sss))
    return f;
  // (X, Z) = G^-1 M'p with G = M'M; H = M G^-1 M'.
  f.X = (sss * scp - scs * ssp) / det;
  f.Z = (scc * ssp - scs * scp) / det;
  f.Y = sy / n;
  // 1'H p = [sc ss] G^-1 [scp ssp]', 1'H 1 = [sc ss] G^-1 [sc ss]'.
  const double gx = (sss * sc - scs * ss) / det, gz = (scc * ss - scs * sc) / det;
  f.a = sp - (gx * scp + gz * ssp);
  f.b = n - (gx * sc + gz * ss);
  f.ok = true;
  return f;
}

}  // namespace

extern "C" int mfio_fiducial_residuals(
    int nx, int ny, int nz, const float* rotate, const double* shift_x,
    const double* shift_y, const float* tilts, int nmarkers,
    const float* raw_xy, int min_views, int fit_centre, double* per_section,
    double* marker_xyz, double* summary) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nmarkers < 0 || !rotate || !shift_x ||
      !shift_y || !tilts || (nmarkers > 0 && !raw_xy) || !summary)
    return -1;
  if (min_views < 2) min_views = 2;
  std::vector<double> cb(nz), sb(nz), cr(nz), sr(nz);
  for (int j = 0; j < nz; j++) {
    const double b = tilts[j] * kPi / 180.0, r = rotate[j] * kPi / 180.0;
    cb[j] = std::cos(b);
    sb[j] = std::sin(b);
    cr[j] = std::cos(r);
    sr[j] = std::sin(r);
  }
  const double cx = nx * 0.5, cy = ny * 0.5;
  // Observations in the aligned frame, per marker.
  std::vector<std::vector<Obs>> obs(nmarkers);
  for (int i = 0; i < nmarkers; i++) {
    for (int j = 0; j < nz; j++) {
      const float* q = raw_xy + 2 * ((size_t)i * nz + j);
      if (!(std::isfinite(q[0]) && std::isfinite(q[1]))) continue;
      const double ux = q[0] - cx - shift_x[j], uy = q[1] - cy - shift_y[j];
      // R(-rot): (c, s; -s, c).
      obs[i].push_back({j, cr[j] * ux + sr[j] * uy, -sr[j] * ux + cr[j] * uy});
    }
  }
  std::vector<Fit> fits(nmarkers);
  double num = 0, den = 0;
  for (int i = 0; i < nmarkers; i++) {
    if ((int)obs[i].size() < min_views) continue;
    fits[i] = fit_marker(obs[i], cb, sb);
    if (!fits[i].ok) continue;
    num += fits[i].a;
    den += fits[i].b;
  }
  const double d = (fit_centre && den > 1e-12) ? num / den : 0.0;
  std::vector<double> sec_sum(nz, 0.0);
  std::vector<int> sec_n(nz, 0);
  double sum = 0, sum2 = 0, mx = 0;
  long long n = 0;
  int used = 0;
  for (int i = 0; i < nmarkers; i++) {
    double* xyz = marker_xyz ? marker_xyz + 3 * (size_t)i : nullptr;
    if (!fits[i].ok) {
      if (xyz) xyz[0] = xyz[1] = xyz[2] = NAN;
      continue;
    }
    // Refit (X, Z) with the common offset removed from the observations.
    std::vector<Obs> v = obs[i];
    for (Obs& o : v) o.px -= d;
    const Fit f = fit_marker(v, cb, sb);
    if (xyz) {
      xyz[0] = f.X;
      xyz[1] = f.Y;
      xyz[2] = f.Z;
    }
    used++;
    for (const Obs& o : v) {
      const double ex = o.px - (f.X * cb[o.sec] + f.Z * sb[o.sec]);
      const double ey = o.py - f.Y;
      const double e2 = ex * ex + ey * ey;
      sec_sum[o.sec] += e2;
      sec_n[o.sec]++;
      sum += std::sqrt(e2);
      sum2 += e2;
      if (std::sqrt(e2) > mx) mx = std::sqrt(e2);
      n++;
    }
  }
  if (per_section)
    for (int j = 0; j < nz; j++)
      per_section[j] = sec_n[j] ? std::sqrt(sec_sum[j] / sec_n[j]) : NAN;
  summary[0] = n ? sum / n : NAN;           // mean residual (px)
  summary[1] = n ? std::sqrt(sum2 / n) : NAN;  // RMS residual (px)
  summary[2] = mx;                           // largest residual (px)
  summary[3] = (double)n;                    // observations used
  summary[4] = (double)used;                 // markers used
  summary[5] = d;                            // common x offset (px)
  return 0;
}

extern "C" int mfio_read_txt(const char* path, int cap, float* rotate,
                             double* shift_x, double* shift_y, float* tilts,
                             float* angle_offset, int* out_bin) {
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  char line[512];
  int n = 0;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    float off;
    int b;
    if (std::sscanf(line, "angleoffset: %f", &off) == 1) {
      if (angle_offset) *angle_offset = off;
      continue;
    }
    if (std::sscanf(line, "Output Bin: %d", &b) == 1) {
      if (out_bin) *out_bin = b;
      continue;
    }
    int sec;
    float rot, tilt;
    double tx, ty;
    if (std::sscanf(line, "%d %f %lf %lf %f", &sec, &rot, &tx, &ty, &tilt) != 5)
      continue;
    if (n < cap) {
      if (rotate) rotate[n] = rot;
      if (shift_x) shift_x[n] = tx;
      if (shift_y) shift_y[n] = ty;
      if (tilts) tilts[n] = tilt;
    }
    n++;
  }
  std::fclose(f);
  return n;
}
