// CPU restatement of the pipeline driver (main/Process.cpp) and of the
// per-stage host logic: PreProcess (Preprocess.cu), Transform
// (Transform.cpp), ProjAlign (ProjAlign.cpp) -- test infrastructure.
//
// Alignment parameters are kept in FP32 as the reference keeps them
// (AlignParam, Util.h:14-21); the HIP path accumulates them in FP64.
#include <algorithm>
#include <cstdio>

#include "oracle.h"
#include "orc_internal.h"
#include "orc_stages.h"

using namespace orc;

struct orc_series {
  int nx, ny, nz;
  std::vector<float> raw;          // file order
  std::vector<float> file_tilts;   // file order
  std::vector<float> stack;        // preprocessed, ascending tilt
  std::vector<float> tilts;        // ascending, + offset once applied
  std::vector<float> sx, sy, rot;  // AlignParam
  float offset = 0.0f;
  int verbose = 0;
  // Transform state (Transform.cpp:24-58).
  int tr_bin = 1;
  std::vector<float> psx, psy, tr_angles;
  // ProjAlign state (ProjAlign.cpp:48-125).
  int pa_bin = 1, pa_tb = 0, pa_nproj = 10;
  float pa_z0 = 0.0f;
  std::vector<float> pa_angles, pa_corr;
  CorrSizes pa_sizes{};
  bool pa_ready = false;
  // Stage trace: one (kind, value) event per stage call (ORC_EV_*).
  std::vector<std::pair<int, float>> trace;
  void ev(int kind, float value) { trace.emplace_back(kind, value); }
};

namespace {

void rot_shift(const float* in, float angle, float* out) {
  // ProjAlign::RotShift (ProjAlign.cpp:448-457).
  float c = (float)std::cos(angle * kD2R), s = (float)std::sin(angle * kD2R);
  float x = in[0] * c - in[1] * s;
  float y = in[0] * s + in[1] * c;
  out[0] = x;
  out[1] = y;
}

void z_induced(const orc_series* S, int i, float* out) {
  // ProjAlign::CalcZInducedShift (ProjAlign.cpp:553-561).
  out[0] = out[1] = 0.0f;
  if (S->pa_z0 == 0) return;
  float t = (float)(S->pa_angles[i] * kD2R);
  out[0] = -S->pa_z0 * fsin(t);
  rot_shift(out, S->rot[i], out);
}

void unstretch(const orc_series* S, int ref, int f, float* sum) {
  // Transform::UnStretch (Transform.cpp:133-157).
  double ds = std::cos(S->tr_angles[ref] * kD2R) / std::cos(S->tr_angles[f] * kD2R);
  double d2t = 2 * kD2R * S->rot[f];
  double s2 = std::sin(d2t), c2 = std::cos(d2t);
  double P = 0.5 * (ds + 1), M = 0.5 * (ds - 1);
  float a0 = (float)(P + M * c2), a1 = (float)(1 * M * s2), a2 = (float)(P - M * c2);
  float det = a0 * a2 - a1 * a1;
  float m0 = a2 / det, m1 = -a1 / det, m2 = a0 / det;
  float buf = sum[0] * m0 + sum[1] * m1;
  sum[1] = sum[0] * m1 + sum[1] * m2;
  sum[0] = buf;
}

}  // namespace

extern "C" orc_series* orc_series_create(int nx, int ny, int nz,
                                         const float* raw,
                                         const float* tilts_file_order) {
  orc_series* S = new orc_series;
  S->nx = nx;
  S->ny = ny;
  S->nz = nz;
  S->raw.assign(raw, raw + (size_t)nx * ny * nz);
  S->file_tilts.assign(tilts_file_order, tilts_file_order + nz);
  S->sx.assign(nz, 0.0f);
  S->sy.assign(nz, 0.0f);
  S->rot.assign(nz, 0.0f);
  return S;
}

extern "C" void orc_series_destroy(orc_series* S) { delete S; }

extern "C" void orc_preprocess(orc_series* S) {
  const int nz = S->nz;
  const size_t plane = (size_t)S->nx * S->ny;
  S->stack.assign(plane * nz, 0.0f);
  // SetPositive (Preprocess.cu:103-153): reverse descending files into
  // ascending order; per-slice min/max skip values < -1e10.
  const bool rev = nz > 1 && S->file_tilts[0] > S->file_tilts[1];
  float stack_min = 1e30f;
  if (S->verbose) printf("Set positivity ...\n  num   residue     min       max\n");
  for (int a = 0; a < nz; a++) {
    int i = rev ? nz - 1 - a : a;
    const float* src = S->raw.data() + a * plane;
    std::memcpy(S->stack.data() + i * plane, src, plane * 4);
    float mn = 1e20f, mx = -1e20f;
    for (size_t k = 0; k < plane; k++) {
      float v = src[k];
      if (v < -1e10f) continue;
      mn = std::fmin(mn, v);
      mx = std::fmax(mx, v);
    }
    if (stack_min > mn) stack_min = mn;
    if (S->verbose) printf("%4d      %4d    %8.2f  %8.2f\n", i, nz - i, mn, mx);
  }
  if (S->verbose) printf("StackMin:%f\n", stack_min);
  if (stack_min < 0) {
    for (size_t k = 0; k < plane * nz; k++) {
      if (S->stack[k] < -1e10f) continue;
      S->stack[k] -= stack_min;
    }
  }
  if (S->verbose) printf("Positivity set.\n\n");
  // MassNormalization (Preprocess.cu:155-177, 216-250); the zero-tilt index
  // comes from the file-order angles (quirk Q3).
  const int x0 = S->nx * 1 / 6, y0 = S->ny * 1 / 6;
  const int mw = S->nx * 4 / 6, mh = S->ny * 4 / 6;
  std::vector<float> mean(nz);
  for (int i = 0; i < nz; i++) {
    const float* f = S->stack.data() + i * plane;
    double s = 0.0;
    int c = 0;
    for (int y = 0; y < mh; y++)
      for (int x = 0; x < mw; x++) {
        float v = f[(size_t)(y + y0) * S->nx + x + x0];
        if (v <= -1e20f) continue;
        s += v;
        c++;
      }
    mean[i] = c == 0 ? 0.0f : (float)(s / c);
  }
  const int z0 = zero_tilt_index(S->file_tilts.data(), nz, 0.0f);
  float ref = mean[z0];
  if (ref > 1000.0f) ref = 1000.0f;
  for (int i = 0; i < nz; i++) {
    float scale = ref / (mean[i] + 0.00001f);
    float* f = S->stack.data() + i * plane;
    for (size_t k = 0; k < plane; k++) {
      if (f[k] <= -1e20f) continue;
      f[k] *= scale;
    }
  }
  S->tilts = S->file_tilts;
  if (rev) std::reverse(S->tilts.begin(), S->tilts.end());
  S->ev(ORC_EV_PREPROCESS, 0.0f);
}

extern "C" void orc_get_stack(const orc_series* S, float* out) {
  std::memcpy(out, S->stack.data(), S->stack.size() * 4);
}
extern "C" void orc_get_tilts(const orc_series* S, float* out) {
  std::memcpy(out, S->tilts.data(), S->nz * 4);
}
extern "C" void orc_set_tilts(orc_series* S, const float* in) {
  S->tilts.assign(in, in + S->nz);
}
extern "C" void orc_get_params(const orc_series* S, float* sx, float* sy,
                               float* rot, float* offset) {
  std::memcpy(sx, S->sx.data(), S->nz * 4);
  std::memcpy(sy, S->sy.data(), S->nz * 4);
  std::memcpy(rot, S->rot.data(), S->nz * 4);
  if (offset) *offset = S->offset;
}
extern "C" void orc_set_params(orc_series* S, const float* sx, const float* sy,
                               const float* rot, float offset) {
  S->sx.assign(sx, sx + S->nz);
  S->sy.assign(sy, sy + S->nz);
  S->rot.assign(rot, rot + S->nz);
  S->offset = offset;
}

extern "C" void orc_coarse_setup(orc_series* S) {
  // Transform::Setup (Transform.cpp:24-58).
  int bx = (int)(S->nx / 1024.0f + 0.1f), by = (int)(S->ny / 1024.0f + 0.1f);
  S->tr_bin = std::max(std::max(bx, by), 1);
  S->psx.assign(S->nz, 0.0f);
  S->psy.assign(S->nz, 0.0f);
  S->tr_angles = S->tilts;
  if (S->verbose) printf("Pre-align tilt series\n");
}

extern "C" float orc_coarse_iter(orc_series* S) {
  // Transform::DoIt / Measure / UpdateShift (Transform.cpp:60-131).
  const int nz = S->nz;
  std::vector<float> corr;
  CorrSizes s;
  correct_stack(S->stack.data(), S->nx, S->ny, nz, S->tr_bin, S->sx.data(),
                S->sy.data(), S->rot.data(), true, true, false, corr, &s);
  const size_t plane = (size_t)s.bx * s.by;
  const float binx = S->nx * 1.0f / s.bx, biny = S->ny * 1.0f / s.by;
  std::vector<int> refs(nz);
  for (int z = 0; z < nz; z++)
    refs[z] = find_ref_index(S->tr_angles.data(), nz, z);
#pragma omp parallel for schedule(dynamic, 1)
  for (int z = 0; z < nz; z++) {
    if (refs[z] == z) continue;
    float sh[2];
    stretch_cc(corr.data() + refs[z] * plane, corr.data() + z * plane, s.bx,
               s.by, S->tr_angles[refs[z]], S->tr_angles[z], S->rot[z], 400.0f,
               sh);
    S->psx[z] = sh[0] * binx;
    S->psy[z] = sh[1] * biny;
  }
  float maxerr = 0.0f;
  for (int z = 0; z < nz; z++) {
    if (refs[z] == z) continue;
    if (S->verbose)
      printf("Image %4d shift: %8.2f, %8.2f px\n", z + 1, S->psx[z], S->psy[z]);
    float e = (float)std::sqrt(S->psx[z] * S->psx[z] + S->psy[z] * S->psy[z]);
    if (e > maxerr) maxerr = e;
  }
  const int z0 = zero_tilt_index(S->tr_angles.data(), nz, 0.0f);
  for (int i = 0; i < nz; i++) {
    S->psx[i] -= S->psx[z0];
    S->psy[i] -= S->psy[z0];
  }
  float sum[2] = {0.0f, 0.0f};
  for (int i = z0 - 1; i >= 0; i--) {
    sum[0] += S->psx[i];
    sum[1] += S->psy[i];
    unstretch(S, i + 1, i, sum);
    S->sx[i] += sum[0];
    S->sy[i] += sum[1];
  }
  sum[0] = sum[1] = 0.0f;
  for (int i = z0 + 1; i < nz; i++) {
    sum[0] += S->psx[i];
    sum[1] += S->psy[i];
    unstretch(S, i - 1, i, sum);
    S->sx[i] += sum[0];
    S->sy[i] += sum[1];
  }
  if (S->verbose) printf("maxerr: %8.2f\n\n", maxerr);
  S->ev(ORC_EV_COARSE, maxerr);
  return maxerr;
}

extern "C" float orc_series_tilt_axis(orc_series* S, float range, int num) {
  float axis = orc_tilt_axis(S->stack.data(), S->nx, S->ny, S->nz,
                             S->tilts.data(), S->sx.data(), S->sy.data(),
                             S->rot.data(), range, num, nullptr, nullptr);
  for (int i = 0; i < S->nz; i++) S->rot[i] = axis;
  if (S->verbose)
    printf("Initial estimate of tilt axes:\nNew tilt axis: %.2f\n\n", axis);
  S->ev(ORC_EV_AXIS, axis);
  return axis;
}

extern "C" float orc_series_find_offset(orc_series* S) {
  std::vector<float> ccs(31);
  float best = orc_find_offset(S->stack.data(), S->nx, S->ny, S->nz,
                               S->tilts.data(), S->sx.data(), S->sy.data(),
                               S->rot.data(), ccs.data());
  if (S->verbose) {
    printf("Determine tilt angle offset.\n");
    float mx = 0.0f;
    for (int i = 0; i < 31; i++) {
      printf("...... %8.2f  %.4e\n", (float)(i - 15), ccs[i]);
      if (ccs[i] > mx) mx = ccs[i];
    }
    printf("Tilt offset: %8.2f,  CC: %.4f\n\n", best, mx);
  }
  S->offset += best;
  S->ev(ORC_EV_OFFSET, best);
  return best;
}

extern "C" void orc_apply_offset(orc_series* S) {
  // Process::mCoarseAlign (Process.cpp:266-269).
  for (int z = 0; z < S->nz; z++) S->tilts[z] += S->offset;
  S->ev(ORC_EV_APPLY_OFFSET, S->offset);
}

extern "C" void orc_projalign_setup(orc_series* S, float zshift, int align_z,
                                    int nproj) {
  // ProjAlign::Setup / Setsize (ProjAlign.cpp:48-125).
  S->pa_z0 = zshift;
  int bx = (int)(S->nx / 2048.0f + 0.5f), by = (int)(S->ny / 2048.0f + 0.5f);
  S->pa_bin = std::max(std::max(bx, by), 1);
  corr_sizes(S->nx, S->ny, S->pa_bin, &S->pa_sizes);
  S->pa_tb = align_z / S->pa_bin / 2 * 2;
  S->pa_nproj = nproj;
  S->pa_angles = S->tilts;
  S->pa_ready = true;
  S->ev(ORC_EV_PA_SETUP, (float)S->pa_tb);
  if (S->verbose)
    printf("bin:%d   cmpsize:%d,%d\nbinsize:%d, %d\n", S->pa_bin,
           S->pa_sizes.bx / 2 + 1, S->pa_sizes.by, S->pa_sizes.bx,
           S->pa_sizes.by);
}

extern "C" float orc_pa_z0(const orc_series* S) { return S->pa_z0; }

extern "C" float orc_projalign_iter(orc_series* S) {
  return orc_projalign_iter_forced(S, nullptr, nullptr);
}

extern "C" float orc_projalign_iter_forced(orc_series* S,
                                           const float* forced_meas,
                                           float* own_meas) {
  // ProjAlign::DoIt / DoItBin (ProjAlign.cpp:127-179).  forced_meas (raw
  // pixels, [nz][2]) replaces each target's measured shift before it is
  // applied; own_meas receives the measurement this sweep made itself.
  const int nz = S->nz;
  const CorrSizes& s = S->pa_sizes;
  correct_stack(S->stack.data(), S->nx, S->ny, nz, S->pa_bin, S->sx.data(),
                S->sy.data(), S->rot.data(), false, true, true, S->pa_corr,
                nullptr);
  const size_t plane = (size_t)s.bx * s.by;
  // FitRotCenterZ (ProjAlign.cpp:459-486); dA starts at 0 (quirk Q13).
  const float* ang = S->pa_angles.data();
  const int z0 = zero_tilt_index(ang, nz, 0.0f);
  float s0[2] = {S->sx[z0], S->sy[z0]};
  rot_shift(s0, -S->rot[z0], s0);
  float cent0 = (float)(s0[0] / std::cos(ang[z0] * kD2R));
  double dA = 0.0, dB = 0.0;
  for (int i = 0; i < nz; i++) {
    float r[2] = {S->sx[i], S->sy[i]};
    rot_shift(r, -S->rot[i], r);
    double sn = std::sin(ang[i] * kD2R), cs = std::cos(ang[i] * kD2R);
    dA += (cent0 * cs - r[0]) * sn;
    dB += sn * sn;
  }
  float old = S->pa_z0;
  S->pa_z0 = (float)(dA / (dB + 1e-30));
  if (S->verbose)
    printf("Rot center Z: %8.2f  %8.2f  %8.2f\n\n", old, S->pa_z0,
           S->pa_z0 - old);
  // RemoveOffsetZ(-1) (ProjAlign.cpp:488-501).
  if (S->pa_z0 != 0) {
    for (int i = 0; i < nz; i++) {
      if (i == z0) continue;
      float ind[2];
      z_induced(S, i, ind);
      S->sx[i] += ind[0] * -1.0f;
      S->sy[i] += ind[1] * -1.0f;
    }
  }
  std::vector<unsigned char> skip(nz, 1);
  skip[z0] = 0;
  if (S->verbose)
    printf("# Projection matching measurements\n# tilt angle  x shift  y shift\n");
  float maxerr = -1e20f;
  std::vector<float> rp(plane);
  const float binx = S->nx / (float)s.bx, biny = S->ny / (float)s.by;
  auto align = [&](int t) {
    // ClacAlignProj (ProjAlign.cpp:181-241).
    orc_reproject(S->pa_corr.data(), s.bx, s.by, nz, S->pa_tb, ang,
                  skip.data(), t, S->pa_nproj, 0, rp.data(), nullptr);
    float ish[2];
    pa_measure(rp.data(), S->pa_corr.data() + t * plane, s.bx, s.by, ish);
    float sh[2] = {ish[0] * binx, ish[1] * biny};
    if (S->verbose) printf("  %6.2f  %8.2f  %8.2f\n", ang[t], sh[0], sh[1]);
    float err = (float)std::sqrt(sh[0] * sh[0] + sh[1] * sh[1]);
    if (own_meas) {
      own_meas[2 * t] = sh[0];
      own_meas[2 * t + 1] = sh[1];
    }
    if (forced_meas) {
      sh[0] = forced_meas[2 * t];
      sh[1] = forced_meas[2 * t + 1];
    }
    rot_shift(sh, S->rot[t], sh);
    float ind[2];
    z_induced(S, t, ind);
    sh[0] += ind[0];
    sh[1] += ind[1];
    S->sx[t] += sh[0];
    S->sy[t] += sh[1];
    correct_one(S->stack.data() + (size_t)t * S->nx * S->ny, s, S->sx[t],
                S->sy[t], S->rot[t], false, true, true,
                S->pa_corr.data() + t * plane);
    skip[t] = 0;
    if (err > maxerr) maxerr = err;
  };
  for (int i = 1; i < nz; i++) {
    if (z0 + i < nz) align(z0 + i);
    if (z0 - i >= 0) align(z0 - i);
  }
  S->ev(ORC_EV_PA_ITER, maxerr);
  return maxerr;
}

namespace {

// Process::ProjAlignOnce (Process.cpp:292-319): one DoIt, then up to nine
// more; a sweep whose error does not exceed the last kept one ends the loop
// when it is below 2 px, else becomes the last kept error.  The parameters
// saved at that point are never restored (quirk Q15), so none are saved.
// Returns the number of DoIt calls.
template <class Iter>
int projalign_once_loop(Iter iter, int verbose) {
  float last = iter();
  int calls = 1;
  if (verbose) printf("fLastErr: %8.2f\n", last);
  for (int i = 1; i < 10; i++) {
    float err = iter();
    calls++;
    if (verbose) printf("fErr: %8.2f\n", err);
    if (err <= last) {
      if (err < 2.0f) break;
      last = err;
    }
  }
  return calls;
}

}  // namespace

extern "C" void orc_projalign_once(orc_series* S) {
  S->ev(ORC_EV_ONCE_BEGIN, 0.0f);
  const int calls =
      projalign_once_loop([&] { return orc_projalign_iter(S); }, S->verbose);
  S->ev(ORC_EV_ONCE_END, (float)calls);
}

extern "C" int orc_projalign_once_script(const float* errs, int n) {
  int k = 0;
  return projalign_once_loop(
      [&] {
        if (k >= n) return 1e30f;  // past the script: never breaks
        return errs[k++];
      },
      0);
}

extern "C" int orc_series_trace(const orc_series* S, int* kinds,
                                float* values, int cap) {
  const int n = (int)S->trace.size();
  for (int i = 0; i < n && i < cap; i++) {
    if (kinds) kinds[i] = S->trace[i].first;
    if (values) values[i] = S->trace[i].second;
  }
  return n;
}

extern "C" void orc_series_trace_clear(orc_series* S) { S->trace.clear(); }

extern "C" void orc_correct_final(orc_series* S, int out_bin, float* out,
                                  int* out_size) {
  // Process::CorrectStack (Process.cpp:329-372): rotation + shift, random
  // fill, no R-weighting, OutBin sum binning.
  std::vector<float> corr;
  CorrSizes s;
  correct_stack(S->stack.data(), S->nx, S->ny, S->nz, out_bin, S->sx.data(),
                S->sy.data(), S->rot.data(), false, true, false, corr, &s);
  if (out) std::memcpy(out, corr.data(), corr.size() * 4);
  if (out_size) {
    out_size[0] = s.bx;
    out_size[1] = s.by;
  }
}

extern "C" void orc_get_corrected(const orc_series* S, float* out,
                                  int* out_size) {
  if (out) std::memcpy(out, S->pa_corr.data(), S->pa_corr.size() * 4);
  if (out_size) {
    out_size[0] = S->pa_sizes.bx;
    out_size[1] = S->pa_sizes.by;
  }
}

extern "C" void orc_run(orc_series* S, const orc_options* o) {
  // Process::DoIt (Process.cpp:34-92) minus file I/O.
  S->verbose = o->verbose;
  orc_preprocess(S);
  if (o->align_z == 0) return;
  std::fill(S->sx.begin(), S->sx.end(), 0.0f);
  std::fill(S->sy.begin(), S->sy.end(), 0.0f);
  std::fill(S->rot.begin(), S->rot.end(), 0.0f);
  // mCoarseAlign (Process.cpp:238-270).
  if (o->pitch_angle == 0) {
    orc_coarse_setup(S);
    for (int i = 1; i <= 3; i++) {
      orc_coarse_iter(S);
      orc_series_tilt_axis(S, 180.0f / i, 100);
    }
  } else {
    for (int i = 0; i < S->nz; i++) S->rot[i] = o->pitch_angle;
  }
  if (o->offset == 0)
    orc_series_find_offset(S);
  else
    S->offset = o->offset;
  orc_apply_offset(S);
  // mProjAlign (Process.cpp:272-290).
  std::fill(S->sx.begin(), S->sx.end(), 0.0f);
  std::fill(S->sy.begin(), S->sy.end(), 0.0f);
  orc_projalign_setup(S, o->zshift, o->align_z, o->nproj);
  orc_projalign_once(S);
  for (int i = 1; i <= 4; i++) {
    orc_series_tilt_axis(S, 20.0f / i, 100);
    if (i == 1) orc_projalign_once(S);
  }
  orc_projalign_once(S);
}
