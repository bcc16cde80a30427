// Per-series pipeline: the Process::DoIt stage sequence (main/Process.cpp)
// over HBM-resident data and the device stages.
//
// Differences in mechanism from the reference, none in result:
//  - the preprocessed stack stays in HBM (the reference writes it to buf.mrc
//    and re-reads one slice per correction, Process.cpp:161-170,
//    CorrStack.cu:382);
//  - during a projection-matching sweep the alignment parameters live on the
//    device (FP64) and each target is measured, applied and re-corrected
//    without a host round trip; the host reads them back once per sweep;
//  - per-tilt shifts are accumulated in FP64 (the reference uses FP32).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "bprp.h"
#include "cc.h"
#include "series.h"

mf_series::~mf_series() {
  if (st) (void)hipStreamSynchronize(st);
  for (mf::DevBuf* b : {&raw, &stack, &tmp, &tr_corr, &tr_work, &tr_mom,
                        &tr_shift, &tr_tab, &cproj, &off_corr, &pa_corr,
                        &pa_reproj, &pa_work, &pa_mom, &pa_dev,
                        &pa_slab, &pa_tab, &axw.lines, &axw.reg, &axw.small,
                        &axw.tstack, &offw.buf})
    b->release(true);  // the series' stream is synchronised
  for (hipEvent_t e : up_ev)
    if (e) (void)hipEventDestroy(e);
  if (pa_graph) (void)hipGraphExecDestroy(pa_graph);
  if (st) (void)hipStreamDestroy(st);
}

namespace mf {

namespace {

int zero_index(const std::vector<float>& t) {
  // GetFrameIdxFromTilt(0) (Util.cpp:8-20): first minimum of |0 - a_i|.
  int idx = 0;
  float m = std::fabs(0.0f - t[0]);
  for (size_t i = 1; i < t.size(); i++) {
    float d = std::fabs(0.0f - t[i]);
    if (d >= m) continue;
    m = d;
    idx = (int)i;
  }
  return idx;
}

int ref_index(const std::vector<float>& t, int z) {
  // FindRefIndex (Util.cpp:22-35).
  const int nz = (int)t.size();
  int i0 = z - 1, i2 = z + 1;
  if (i0 < 0) i0 = z;
  if (i2 >= nz) i2 = z;
  double a0 = std::fabs((double)t[i0]), a = std::fabs((double)t[z]),
         a2 = std::fabs((double)t[i2]);
  if (a0 < a) return i0;
  if (a2 < a) return i2;
  return z;
}

// Correction parameters for every projection; shift_only uses angle 0.
std::vector<CorrProj> corr_params(const mf_series* s, bool shift_only) {
  std::vector<CorrProj> p(s->nz);
  for (int z = 0; z < s->nz; z++) {
    float th = shift_only ? 0.0f : (float)(s->rot[z] * kD2R);
    p[z].shift_x = (float)s->sx[z];
    p[z].shift_y = (float)s->sy[z];
    p[z].cos_a = hcos(th);
    p[z].sin_a = hsin(th);
  }
  return p;
}

// The preprocessed stack transposed (made once per stack, shared with the
// common-line search): the corrections of projections rotated nearer a
// quarter turn than none gather their rows from it.
const float* stack_t(mf_series* s) {
  return stack_transposed(s->ctx, s->st, (const float*)s->stack.get(), s->nx,
                          s->ny, s->nz, s->axw);
}

CorrProj* upload_corr(mf_series* s, const std::vector<CorrProj>& p) {
  CorrProj* d = (CorrProj*)s->cproj.ensure(s->ctx, p.size() * sizeof(CorrProj),
                                           s->st);
  MF_HIP(hipMemcpyAsync(d, p.data(), p.size() * sizeof(CorrProj),
                        hipMemcpyHostToDevice, s->st));
  return d;
}

// Device-side parameter block of a projection-matching sweep.
struct PaDev {
  float2* ishift;  // [2] measured shift of the step's one or two targets
  float2* meas;    // [nz] measured shift (raw pixels) per target
  float* err;      // [nz]
  float2* ind;     // [nz] Z-induced shift (CalcZInducedShift)
  float2* rcs;     // [nz] cos, sin of the axis for RotShift
  double* sx;      // [nz]
  double* sy;      // [nz]
};

PaDev pa_layout(mf_series* s) {
  const int nz = s->nz;
  const size_t bytes = 16 + (size_t)nz * (8 + 4 + 8 + 8 + 8 + 8) + 64;
  char* p = (char*)s->pa_dev.ensure(s->ctx, bytes, s->st);
  PaDev d;
  d.sx = (double*)p;
  d.sy = d.sx + nz;
  d.ishift = (float2*)(d.sy + nz);
  d.meas = d.ishift + 2;
  d.ind = d.meas + nz;
  d.rcs = d.ind + nz;
  d.err = (float*)(d.rcs + nz);
  return d;
}

void rot_shift(const float* in, float angle, float* out) {
  // ProjAlign::RotShift (ProjAlign.cpp:448-457).
  float c = (float)std::cos(angle * kD2R), s = (float)std::sin(angle * kD2R);
  float x = in[0] * c - in[1] * s;
  float y = in[0] * s + in[1] * c;
  out[0] = x;
  out[1] = y;
}

}  // namespace

void series_preprocess(mf_series* s) {
  // PreProcess::SetPositive + MassNormalization (Preprocess.cu:103-177).
  const int nz = s->nz;
  const size_t plane = (size_t)s->nx * s->ny;
  const bool rev = nz > 1 && s->file_tilts[0] > s->file_tilts[1];
  std::vector<float> mm(2 * nz);
  slice_minmax(s->ctx, s->st, (const float*)s->raw.get(), plane, nz, s->tmp,
               mm.data());
  float stack_min = 1e30f;
  if (s->verbose) printf("Set positivity ...\n  num   residue     min       max\n");
  for (int a = 0; a < nz; a++) {
    const int i = rev ? nz - 1 - a : a;
    if (stack_min > mm[2 * a]) stack_min = mm[2 * a];
    if (s->verbose)
      printf("%4d      %4d    %8.2f  %8.2f\n", i, nz - i, mm[2 * a], mm[2 * a + 1]);
  }
  if (s->verbose) printf("StackMin:%f\n", stack_min);
  float* stk = (float*)s->stack.ensure(s->ctx, plane * nz * sizeof(float),
                                       s->st);
  reorder_subtract(s->ctx, s->st, (const float*)s->raw.get(), stk, plane, nz,
                   rev, stack_min < 0 ? stack_min : 0.0f);
  if (s->verbose) printf("Positivity set.\n\nLinear mass normalization...\n");
  std::vector<float> means(nz);
  box_means(s->ctx, s->st, stk, s->nx, s->ny, nz, s->nx * 1 / 6, s->ny * 1 / 6,
            s->nx * 4 / 6, s->ny * 4 / 6, s->tmp, means.data());
  // The zero-tilt index is taken from the file-order angles (quirk Q3).
  const int z0 = zero_index(s->file_tilts);
  float ref = means[z0];
  if (ref > 1000.0f) ref = 1000.0f;
  std::vector<float> scale(nz);
  for (int i = 0; i < nz; i++) scale[i] = ref / (means[i] + 0.00001f);
  scale_slices(s->ctx, s->st, stk, plane, nz, s->tmp, scale.data());
  if (s->verbose) printf("Linear mass normalization: done.\n\n");
  // The raw stack's block goes to the context's cache once this series'
  // stream is done with it: freeing it would wait for every stream of the
  // device, another series' download included.
  MF_HIP(hipStreamSynchronize(s->st));
  s->raw.release(true);
  s->tilts = s->file_tilts;
  if (rev) std::reverse(s->tilts.begin(), s->tilts.end());
  s->sx.assign(nz, 0.0);
  s->sy.assign(nz, 0.0);
  s->rot.assign(nz, 0.0f);
  s->offset = 0.0f;
  s->axw.t_of = nullptr;  // the stack's pixels are new
  s->axw.neg_of = nullptr;
  s->preprocessed = true;
  s->ev(MF_EV_PREPROCESS, 0.0f);
}

void series_coarse_setup(mf_series* s) {
  // Transform::Setup (Transform.cpp:24-58).
  MF_REQUIRE(s->preprocessed, "preprocess the series first");
  int bx = (int)(s->nx / 1024.0f + 0.1f), by = (int)(s->ny / 1024.0f + 0.1f);
  s->tr_bin = std::max(std::max(bx, by), 1);
  s->psx.assign(s->nz, 0.0);
  s->psy.assign(s->nz, 0.0);
  s->tr_angles = s->tilts;
  s->tr_ready = true;
  if (s->verbose) printf("Pre-align tilt series\n");
}

float series_coarse_iter(mf_series* s) {
  // Transform::DoIt / Measure / UpdateShift (Transform.cpp:60-131).
  if (!s->tr_ready) throw Error(MF_ESTATE, "coarse alignment not set up");
  const int nz = s->nz;
  const CorrGeom g = CorrGeom::make(s->nx, s->ny, s->tr_bin);
  const size_t plane = (size_t)g.bx * g.by;
  float* corr = (float*)s->tr_corr.ensure(s->ctx, plane * nz * sizeof(float),
                                          s->st);
  CorrProj* cp = upload_corr(s, corr_params(s, true));
  correct_bin(s->ctx, s->st, (const float*)s->stack.get(), g, cp, 0, nz, true,
              corr, g.bx, plane, stack_t(s));
  std::vector<PairDesc> pairs;
  std::vector<int> pz;
  for (int z = 0; z < nz; z++) {
    const int r = ref_index(s->tr_angles, z);
    if (r == z) continue;
    PairDesc p;
    p.ref = r;
    p.img = z;
    double ds = std::cos(kD2R * s->tr_angles[r]) / std::cos(kD2R * s->tr_angles[z]);
    stretch_matrix((float)ds, s->rot[z], p.m);
    pairs.push_back(p);
    pz.push_back(z);
  }
  const int np = (int)pairs.size();
  float* work = (float*)s->tr_work.ensure(
      s->ctx,
      stretch_cc_work_floats(g.bx, g.by, std::max(np, 1)) * sizeof(float),
      s->st);
  double* mom = (double*)s->tr_mom.ensure(
      s->ctx, cc_scratch_doubles(std::max(np, 1), g.by) * sizeof(double),
      s->st);
  float* dsh = (float*)s->tr_shift.ensure(s->ctx, 2 * (np + 1) * sizeof(float),
                                          s->st);
  stretch_cc_pairs(s->ctx, s->st, corr, g.bx, g.by, pairs, 400.0f, work, mom,
                   s->tr_tab, dsh);
  std::vector<float> sh(2 * np);
  MF_HIP(hipMemcpyAsync(sh.data(), dsh, 2 * np * sizeof(float),
                        hipMemcpyDeviceToHost, s->st));
  MF_HIP(hipStreamSynchronize(s->st));
  const float binx = s->nx * 1.0f / g.bx, biny = s->ny * 1.0f / g.by;
  float maxerr = 0.0f;
  for (int p = 0; p < np; p++) {
    const int z = pz[p];
    const float px = sh[2 * p] * binx, py = sh[2 * p + 1] * biny;
    s->psx[z] = px;
    s->psy[z] = py;
    if (s->verbose) printf("Image %4d shift: %8.2f, %8.2f px\n", z + 1, px, py);
    float e = (float)std::sqrt(px * px + py * py);
    if (e > maxerr) maxerr = e;
  }
  const int z0 = zero_index(s->tr_angles);
  for (int i = 0; i < nz; i++) {
    s->psx[i] -= s->psx[z0];
    s->psy[i] -= s->psy[z0];
  }
  // Accumulate outward from the zero tilt, un-stretching the running sum
  // into each image's frame (UnStretch, Transform.cpp:133-157), in FP64.
  auto unstretch = [&](int ref, int f, double* sum) {
    double ds = std::cos(s->tr_angles[ref] * kD2R) / std::cos(s->tr_angles[f] * kD2R);
    double d2t = 2 * kD2R * s->rot[f];
    double P = 0.5 * (ds + 1), M = 0.5 * (ds - 1);
    float a0 = (float)(P + M * std::cos(d2t)), a1 = (float)(1 * M * std::sin(d2t)),
          a2 = (float)(P - M * std::cos(d2t));
    float det = a0 * a2 - a1 * a1;
    double m0 = a2 / det, m1 = -a1 / det, m2 = a0 / det;
    double b = sum[0] * m0 + sum[1] * m1;
    sum[1] = sum[0] * m1 + sum[1] * m2;
    sum[0] = b;
  };
  double sum[2] = {0.0, 0.0};
  for (int i = z0 - 1; i >= 0; i--) {
    sum[0] += s->psx[i];
    sum[1] += s->psy[i];
    unstretch(i + 1, i, sum);
    s->sx[i] += sum[0];
    s->sy[i] += sum[1];
  }
  sum[0] = sum[1] = 0.0;
  for (int i = z0 + 1; i < nz; i++) {
    sum[0] += s->psx[i];
    sum[1] += s->psy[i];
    unstretch(i - 1, i, sum);
    s->sx[i] += sum[0];
    s->sy[i] += sum[1];
  }
  if (s->verbose) printf("maxerr: %8.2f\n\n", maxerr);
  s->ev(MF_EV_COARSE, maxerr);
  return maxerr;
}

float series_tilt_axis(mf_series* s, float range, int num) {
  // CalcTIltAxis::DoIt (Rotate.cu:292-337).
  MF_REQUIRE(s->preprocessed, "preprocess the series first");
  MF_REQUIRE(num >= 2, "need at least two axis candidates");
  const int z0 = zero_index(s->tilts);
  AxisResult r = axis_search(s->ctx, s->st, (const float*)s->stack.get(), s->nx,
                             s->ny, s->nz, s->tilts.data(), s->sx.data(),
                             s->sy.data(), s->rot[z0], range, num, s->axw);
  s->axis_scores = r.scores;
  float axis = r.axis;
  if (axis < -45 && s->rot[0] == 0) axis += 180.0f;
  if (axis > 135) axis -= 180.0f;
  for (int i = 0; i < s->nz; i++) s->rot[i] = axis;
  if (s->verbose)
    printf("Initial estimate of tilt axes:\nNew tilt axis: %.2f\n\n", axis);
  s->ev(MF_EV_AXIS, axis);
  return axis;
}

float series_find_offset(mf_series* s) {
  // FindOffset::DoIt / FOsetup (FindOffset.cu:85-127): bin 4 (quirk Q8),
  // rotation, no random fill, no R-weighting.
  MF_REQUIRE(s->preprocessed, "preprocess the series first");
  const CorrGeom g = CorrGeom::make(s->nx, s->ny, 4);
  const size_t plane = (size_t)g.bx * g.by;
  float* corr = (float*)s->off_corr.ensure(
      s->ctx, plane * s->nz * sizeof(float), s->st);
  CorrProj* cp = upload_corr(s, corr_params(s, false));
  correct_bin(s->ctx, s->st, (const float*)s->stack.get(), g, cp, 0, s->nz,
              false, corr, g.bx, plane, stack_t(s));
  std::vector<float> ccs;
  const float best = offset_search(s->ctx, s->st, corr, g.bx, g.by, s->nz,
                                   s->tilts.data(), s->offw, &ccs);
  if (s->verbose) {
    printf("Determine tilt angle offset.\n");
    float mx = 0.0f;
    for (int i = 0; i < 31; i++) {
      printf("...... %8.2f  %.4e\n", (float)(i - 15), ccs[i]);
      if (ccs[i] > mx) mx = ccs[i];
    }
    printf("Tilt offset: %8.2f,  CC: %.4f\n\n", best, mx);
  }
  s->offset_scores = ccs;
  s->offset += best;
  if (s->verbose)
    printf("Tilt axis angle offset is: %f\n", s->offset);
  s->ev(MF_EV_OFFSET, best);
  return best;
}

void series_apply_offset(mf_series* s) {
  // Process::mCoarseAlign (Process.cpp:266-269).
  for (int z = 0; z < s->nz; z++) s->tilts[z] += s->offset;
  s->ev(MF_EV_APPLY_OFFSET, s->offset);
}

void series_projalign_setup(mf_series* s, float zshift, int align_z,
                            int nproj) {
  // ProjAlign::Setup / Setsize (ProjAlign.cpp:48-125).
  MF_REQUIRE(s->preprocessed, "preprocess the series first");
  MF_REQUIRE(align_z > 0, "AlignZ must be > 0");
  MF_REQUIRE(nproj > 0 && nproj <= kMaxRefProj, "NPROJ out of range");
  s->pa_z0 = zshift;
  int bx = (int)(s->nx / 2048.0f + 0.5f), by = (int)(s->ny / 2048.0f + 0.5f);
  s->pa_bin = std::max(std::max(bx, by), 1);
  s->pa_geom = CorrGeom::make(s->nx, s->ny, s->pa_bin);
  s->pa_tb = align_z / s->pa_bin / 2 * 2;
  MF_REQUIRE(s->pa_tb >= 2, "projection-matching thickness too small");
  s->pa_nproj = nproj;
  s->pa_angles = s->tilts;
  const CorrGeom& g = s->pa_geom;
  const size_t plane = (size_t)g.bx * g.by, padb = pad_x(g.bx);
  s->pa_corr.ensure(s->ctx, plane * s->nz * sizeof(float), s->st);
  // Two targets per sweep step at most: two reprojections, four work images.
  s->pa_reproj.ensure(s->ctx, 2 * plane * sizeof(float), s->st);
  s->pa_work.ensure(s->ctx, 4 * (size_t)g.by * padb * sizeof(float), s->st);
  s->pa_mom.ensure(s->ctx, cc_scratch_doubles(2, g.by) * sizeof(double), s->st);
  // New angles and thickness: a captured sweep no longer applies.
  s->pa_gen++;
  s->pa_ekey.clear();
  if (s->pa_graph) {
    MF_HIP(hipStreamSynchronize(s->st));
    MF_HIP(hipGraphExecDestroy(s->pa_graph));
    s->pa_graph = nullptr;
  }
  s->pa_gkey.clear();
  s->pa_ready = true;
  s->ev(MF_EV_PA_SETUP, (float)s->pa_tb);
  if (s->verbose)
    printf("bin:%d   cmpsize:%d,%d\nbinsize:%d, %d\n", s->pa_bin, g.bx / 2 + 1,
           g.by, g.bx, g.by);
}

namespace {

// Runs one sweep's launches on the series' stream: eagerly, or, with graphs
// enabled and kernel timing off, by replaying the graph captured from an
// earlier sweep with the same key.  The first sweep of a key runs eagerly
// (it also creates the tables, transform plans and launch attributes the
// launches use), the second is captured, instantiated and launched, later
// ones are replayed.  The captured launches hold no memory set or copy.
template <class F>
void sweep_launches(mf_series* s, const std::vector<unsigned long long>& key,
                    F&& launch) {
  const bool can = s->graphs && !s->ctx->timer.enabled();
  if (can && s->pa_graph && key == s->pa_gkey) {
    MF_HIP(hipGraphLaunch(s->pa_graph, s->st));
    s->pa_replays++;
    return;
  }
  if (!can || key != s->pa_ekey) {
    launch();
    s->pa_ekey = key;
    return;
  }
  if (s->pa_graph) {
    MF_HIP(hipStreamSynchronize(s->st));
    MF_HIP(hipGraphExecDestroy(s->pa_graph));
    s->pa_graph = nullptr;
    s->pa_gkey.clear();
  }
  hipGraph_t graph = nullptr;
  MF_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  try {
    launch();
  } catch (...) {
    (void)hipStreamEndCapture(s->st, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  MF_HIP(hipStreamEndCapture(s->st, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  MF_HIP(e);
  s->pa_graph = exec;
  s->pa_gkey = key;
  MF_HIP(hipGraphLaunch(exec, s->st));
  s->pa_replays++;
}

}  // namespace

float series_projalign_iter(mf_series* s) {
  // ProjAlign::DoIt / DoItBin (ProjAlign.cpp:127-179).
  if (!s->pa_ready) throw Error(MF_ESTATE, "projection matching not set up");
  mf_ctx* ctx = s->ctx;
  const int nz = s->nz;
  const CorrGeom& g = s->pa_geom;
  const size_t plane = (size_t)g.bx * g.by;
  float* corr = (float*)s->pa_corr.get();
  CorrProj* cp = upload_corr(s, corr_params(s, false));
  // FitRotCenterZ (ProjAlign.cpp:459-486) with dA starting at 0 (quirk Q13),
  // then RemoveOffsetZ(-1) (:488-501).
  const std::vector<float>& ang = s->pa_angles;
  const int z0 = zero_index(ang);
  float s0[2] = {(float)s->sx[z0], (float)s->sy[z0]};
  rot_shift(s0, -s->rot[z0], s0);
  const double cent0 = s0[0] / std::cos(ang[z0] * kD2R);
  double dA = 0.0, dB = 0.0;
  for (int i = 0; i < nz; i++) {
    float r[2] = {(float)s->sx[i], (float)s->sy[i]};
    rot_shift(r, -s->rot[i], r);
    double sn = std::sin(ang[i] * kD2R), cs = std::cos(ang[i] * kD2R);
    dA += (cent0 * cs - r[0]) * sn;
    dB += sn * sn;
  }
  const float old = s->pa_z0;
  s->pa_z0 = (float)(dA / (dB + 1e-30));
  if (s->verbose)
    printf("Rot center Z: %8.2f  %8.2f  %8.2f\n\n", old, s->pa_z0, s->pa_z0 - old);
  std::vector<float> ind(2 * nz, 0.0f), rcs(2 * nz);
  for (int i = 0; i < nz; i++) {
    // CalcZInducedShift (ProjAlign.cpp:553-561).
    if (s->pa_z0 != 0) {
      float t = (float)(ang[i] * kD2R);
      float v[2] = {-s->pa_z0 * hsin(t), 0.0f};
      rot_shift(v, s->rot[i], v);
      ind[2 * i] = v[0];
      ind[2 * i + 1] = v[1];
    }
    rcs[2 * i] = (float)std::cos(s->rot[i] * kD2R);
    rcs[2 * i + 1] = (float)std::sin(s->rot[i] * kD2R);
  }
  if (s->pa_z0 != 0)
    for (int i = 0; i < nz; i++) {
      if (i == z0) continue;
      s->sx[i] += (double)(ind[2 * i] * -1.0f);
      s->sy[i] += (double)(ind[2 * i + 1] * -1.0f);
    }
  PaDev d = pa_layout(s);
  MF_HIP(hipMemcpyAsync(d.sx, s->sx.data(), nz * 8, hipMemcpyHostToDevice, s->st));
  MF_HIP(hipMemcpyAsync(d.sy, s->sy.data(), nz * 8, hipMemcpyHostToDevice, s->st));
  MF_HIP(hipMemcpyAsync(d.ind, ind.data(), nz * 8, hipMemcpyHostToDevice, s->st));
  MF_HIP(hipMemcpyAsync(d.rcs, rcs.data(), nz * 8, hipMemcpyHostToDevice, s->st));
  std::vector<unsigned char> skip(nz, 1);
  skip[z0] = 0;
  const float binx = s->nx / (float)g.bx, biny = s->ny / (float)g.by;
  float* reproj = (float*)s->pa_reproj.get();
  float* work = (float*)s->pa_work.get();
  double* mom = (double*)s->pa_mom.get();
  float* slab = nullptr;
  if (s->reproj_mode == 0 || s->reproj_mode == 6 || s->reproj_mode == 7)
    slab = (float*)s->pa_slab.ensure(
        ctx, slab_elems(g.bx, g.by, s->pa_tb, s->reproj_mode) * sizeof(float),
        s->st);
  // The sweep visits z0+1, z0-1, z0+2, z0-2, ...  A step takes the pair
  // (z0+i, z0-i) together when the second target's reference range, formed
  // exactly as it would be after the first is aligned, does not contain the
  // first: then neither target reads the other's corrected image, both see
  // the same parameters as in the one-by-one order, and their measurements,
  // updates and re-corrections can share launches.  The schedule depends
  // only on the tilts, so it is fixed before the sweep.
  struct Step {
    int a, b;  // b < 0: one target
  };
  std::vector<Step> steps;
  {
    std::vector<unsigned char> sk = skip;
    for (int i = 1; i < nz; i++) {
      const int a = z0 + i < nz ? z0 + i : -1, b = z0 - i >= 0 ? z0 - i : -1;
      if (a >= 0 && b >= 0) {
        sk[a] = 0;
        int r[2];
        find_proj_range(ang.data(), nz, sk.data(), b, s->pa_nproj, r);
        sk[b] = 0;
        if (s->pa_pairs && (r[0] < 0 || a < r[0] || a > r[1])) {
          steps.push_back({a, b});
        } else {
          steps.push_back({a, -1});
          steps.push_back({b, -1});
        }
      } else if (a >= 0 || b >= 0) {
        const int t = a >= 0 ? a : b;
        sk[t] = 0;
        steps.push_back({t, -1});
      }
    }
  }
  // Correlation image tables of the paired steps: reprojection k and
  // projection k (k = 0, 1) with their work images.
  const size_t wimg = (size_t)g.by * pad_x(g.bx);
  std::vector<CcImg> tab;
  for (const Step& st : steps)
    if (st.b >= 0) {
      const int ts[2] = {st.a, st.b};
      for (int k = 0; k < 2; k++) {
        tab.push_back(CcImg{reproj + k * plane, g.bx, work + (2 * k) * wimg});
        tab.push_back(CcImg{corr + ts[k] * plane, g.bx, work + (2 * k + 1) * wimg});
      }
    }
  CcImg* d_tab = nullptr;
  if (!tab.empty()) {
    d_tab = (CcImg*)s->pa_tab.ensure(ctx, tab.size() * sizeof(CcImg), s->st);
    MF_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(CcImg),
                          hipMemcpyHostToDevice, s->st));
  }
  std::vector<int> order;
  for (const Step& st : steps) {
    order.push_back(st.a);
    if (st.b >= 0) order.push_back(st.b);
  }
  // The sweep's launches.  Their arguments depend only on the tilts, the
  // schedule and the buffers in 'key'; the parameters they read and update
  // (cp, d.*) are device-resident and uploaded above.
  auto launch = [&] {
    // Correct + bin + R-weight the whole stack with the current parameters.
    correct_rweight(ctx, s->st, (const float*)s->stack.get(), g, cp, 0, nz,
                    true, corr, 1, stack_t(s));
    std::vector<unsigned char> sk = skip;
    int paired = 0;
    for (const Step& st : steps) {
      // ClacAlignProj (ProjAlign.cpp:181-241), for one target or a pair.
      const int n = st.b >= 0 ? 2 : 1;
      const int ts[2] = {st.a, st.b};
      for (int k = 0; k < n; k++) {
        reproject_target(ctx, s->st, corr, g.bx, g.by, nz, s->pa_tb,
                         ang.data(), sk.data(), ts[k], s->pa_nproj,
                         s->reproj_mode, reproj + k * plane, slab);
        sk[ts[k]] = 0;  // the next target's range sees this one aligned
      }
      // The correlation's fit also applies each target's parameter update.
      PaUpdate up{};
      up.meas = (float*)d.meas;
      up.err = d.err;
      up.ind = (const float*)d.ind;
      up.rcs = (const float*)d.rcs;
      up.sx = d.sx;
      up.sy = d.sy;
      up.cp = cp;
      up.t[0] = st.a;
      up.t[1] = st.b;
      up.binx = binx;
      up.biny = biny;
      if (n == 2) {
        pa_measure_tab(ctx, s->st, d_tab + 4 * paired, 2, g.bx, g.by, mom,
                       (float*)d.ishift, &up);
        paired++;
      } else {
        pa_measure(ctx, s->st, reproj, corr + st.a * plane, g.bx, g.by, work,
                   mom, (float*)d.ishift, &up);
      }
      // Re-correct the step's targets (one launch for a pair).
      correct_rweight(ctx, s->st, (const float*)s->stack.get(), g, cp, st.a, n,
                      true, corr + st.a * plane, n == 2 ? st.b - st.a : 1,
                      stack_t(s));
    }
  };
  const std::vector<unsigned long long> key = {
      s->pa_gen, ctx->tune_gen.load(), (unsigned long long)s->reproj_mode,
      (unsigned long long)s->pa_pairs, (unsigned long long)s->pa_nproj,
      (unsigned long long)(uintptr_t)s->stack.get(),
      (unsigned long long)(uintptr_t)stack_t(s),
      (unsigned long long)(uintptr_t)corr, (unsigned long long)(uintptr_t)cp,
      (unsigned long long)(uintptr_t)reproj,
      (unsigned long long)(uintptr_t)work, (unsigned long long)(uintptr_t)mom,
      (unsigned long long)(uintptr_t)slab,
      (unsigned long long)(uintptr_t)d_tab,
      (unsigned long long)(uintptr_t)d.meas};
  sweep_launches(s, key, launch);
  std::vector<float> err(nz), meas(2 * nz);
  MF_HIP(hipMemcpyAsync(s->sx.data(), d.sx, nz * 8, hipMemcpyDeviceToHost, s->st));
  MF_HIP(hipMemcpyAsync(s->sy.data(), d.sy, nz * 8, hipMemcpyDeviceToHost, s->st));
  MF_HIP(hipMemcpyAsync(err.data(), d.err, nz * 4, hipMemcpyDeviceToHost, s->st));
  MF_HIP(hipMemcpyAsync(meas.data(), d.meas, nz * 8, hipMemcpyDeviceToHost, s->st));
  MF_HIP(hipStreamSynchronize(s->st));
  if (s->verbose)
    printf("# Projection matching measurements\n# tilt angle  x shift  y shift\n");
  float maxerr = -1e20f;
  s->pa_meas.assign(2 * nz, 0.0f);
  s->pa_err.assign(nz, 0.0f);
  for (int t : order) {
    if (s->verbose)
      printf("  %6.2f  %8.2f  %8.2f\n", ang[t], meas[2 * t], meas[2 * t + 1]);
    if (err[t] > maxerr) maxerr = err[t];
    s->pa_meas[2 * t] = meas[2 * t];
    s->pa_meas[2 * t + 1] = meas[2 * t + 1];
    s->pa_err[t] = err[t];
  }
  s->ev(MF_EV_PA_ITER, maxerr);
  return maxerr;
}

void series_projalign_once(mf_series* s) {
  // Process::ProjAlignOnce (Process.cpp:292-319); the saved parameters are
  // never restored (quirk Q15), so none are saved here.
  s->ev(MF_EV_ONCE_BEGIN, 0.0f);
  int calls = 1;
  float last = series_projalign_iter(s);
  if (s->verbose) printf("fLastErr为: %8.2f\n", last);
  for (int i = 1; i < 10; i++) {
    float err = series_projalign_iter(s);
    calls++;
    if (s->verbose) printf("fErr为: %8.2f\n", err);
    if (err <= last) {
      if (err < 2.0f) break;
      last = err;
    }
  }
  s->ev(MF_EV_ONCE_END, (float)calls);
}

void series_correct(mf_series* s, int out_bin, float* out, int* bx, int* by,
                    bool wait) {
  // Process::CorrectStack (Process.cpp:329-372): rotation + shift, random
  // fill, no R-weighting, OutBin sum binning.  Without 'wait' the download
  // into page-locked 'out' is left running on the series' stream (any other
  // host memory is copied before the call returns).
  MF_REQUIRE(s->preprocessed, "preprocess the series first");
  MF_REQUIRE(out_bin >= 1, "OutBin must be >= 1");
  const CorrGeom g = CorrGeom::make(s->nx, s->ny, out_bin);
  if (bx) *bx = g.bx;
  if (by) *by = g.by;
  if (!out) return;
  const size_t plane = (size_t)g.bx * g.by;
  float* d = (float*)s->tmp.ensure(s->ctx, plane * s->nz * sizeof(float),
                                   s->st);
  CorrProj* cp = upload_corr(s, corr_params(s, false));
  correct_bin(s->ctx, s->st, (const float*)s->stack.get(), g, cp, 0, s->nz,
              true, d, g.bx, plane, stack_t(s));
  MF_HIP(hipMemcpyAsync(out, d, plane * s->nz * sizeof(float),
                        hipMemcpyDeviceToHost, s->st));
  if (wait || !host_pinned(out)) MF_HIP(hipStreamSynchronize(s->st));
}

void series_wait(mf_series* s) { MF_HIP(hipStreamSynchronize(s->st)); }

size_t series_hbm_bytes(int nx, int ny, int nz, const mf_options* o) {
  const size_t raw = (size_t)4 * nx * ny * nz;
  // Coarse alignment: corrected stack at the Transform bin plus the CC work
  // of all adjacent pairs.
  const int tb = std::max(std::max((int)(nx / 1024.0f + 0.1f),
                                   (int)(ny / 1024.0f + 0.1f)), 1);
  const CorrGeom gc = CorrGeom::make(nx, ny, tb);
  const size_t coarse = (size_t)4 * gc.bx * gc.by * nz +
                        stretch_cc_work_floats(gc.bx, gc.by, nz) * 4;
  // Projection matching: the corrected, R-weighted stack, two
  // reprojections and four work images (a sweep step measures up to two
  // targets); the slab if materialised.
  const int pb = std::max(std::max((int)(nx / 2048.0f + 0.5f),
                                   (int)(ny / 2048.0f + 0.5f)), 1);
  const CorrGeom gp = CorrGeom::make(nx, ny, pb);
  size_t pa = (size_t)4 * nz * gp.by * gp.bx +
              (size_t)4 * 6 * gp.by * pad_x(gp.bx);
  if (o->reproj_mode == 0 || o->reproj_mode == 6 || o->reproj_mode == 7)
    pa += 4 * slab_elems(gp.bx, gp.by, o->align_z / pb / 2 * 2, o->reproj_mode);
  // Common lines (lines and spectra of every image and candidate, the
  // transposed stack), offset scan at bin 4, final correction, small tables.
  const size_t axis = (size_t)8 * nz * 100 * (std::min(nx, ny) + 2) + raw;
  const CorrGeom go = CorrGeom::make(nx, ny, std::max(o->out_bin, 1));
  const size_t out = (size_t)4 * go.bx * go.by * nz;
  // Stage buffers stay allocated for the life of the series: sum them.
  return 2 * raw + coarse + pa + axis + out + raw / 8 + ((size_t)64 << 20);
}

void series_trim(mf_series* s) {
  MF_HIP(hipStreamSynchronize(s->st));
  if (s->pa_graph) {
    MF_HIP(hipGraphExecDestroy(s->pa_graph));
    s->pa_graph = nullptr;
  }
  s->pa_gkey.clear();
  s->pa_ekey.clear();
  for (DevBuf* b : {&s->tmp, &s->tr_corr, &s->tr_work, &s->tr_mom,
                    &s->tr_shift, &s->tr_tab, &s->cproj, &s->off_corr,
                    &s->pa_corr, &s->pa_reproj, &s->pa_work, &s->pa_mom,
                    &s->pa_dev, &s->pa_slab, &s->pa_tab, &s->axw.lines,
                    &s->axw.reg, &s->axw.small, &s->axw.tstack, &s->axw.negf,
                    &s->offw.buf})
    b->release(true);  // the series' stream is synchronised
  s->axw.t_of = nullptr;
  s->axw.neg_of = nullptr;
  s->tr_ready = false;
  s->pa_ready = false;
}

void series_run(mf_series* s, const mf_options* o) {
  // Process::DoIt (Process.cpp:34-92) minus file I/O.
  s->verbose = o->verbose;
  s->reproj_mode = o->reproj_mode;
  s->pa_pairs = o->pa_pairs != 0;
  s->graphs = o->graphs != 0;
  if (!s->preprocessed) series_preprocess(s);
  if (o->align_z == 0) return;
  // SetParam (Process.cpp:208-221).
  std::fill(s->sx.begin(), s->sx.end(), 0.0);
  std::fill(s->sy.begin(), s->sy.end(), 0.0);
  std::fill(s->rot.begin(), s->rot.end(), 0.0f);
  s->offset = 0.0f;
  // mCoarseAlign (Process.cpp:238-270).
  if (o->pitch_angle == 0) {
    series_coarse_setup(s);
    for (int i = 1; i <= 3; i++) {
      series_coarse_iter(s);
      series_tilt_axis(s, 180.0f / i, 100);
    }
  } else {
    std::fill(s->rot.begin(), s->rot.end(), o->pitch_angle);
  }
  if (o->offset == 0)
    series_find_offset(s);
  else
    s->offset = o->offset;
  series_apply_offset(s);
  // mProjAlign (Process.cpp:272-290); ResetShift first (quirk Q16).
  std::fill(s->sx.begin(), s->sx.end(), 0.0);
  std::fill(s->sy.begin(), s->sy.end(), 0.0);
  series_projalign_setup(s, o->zshift, o->align_z, o->nproj);
  series_projalign_once(s);
  for (int i = 1; i <= 4; i++) {
    series_tilt_axis(s, 20.0f / i, 100);
    if (i == 1) series_projalign_once(s);
  }
  series_projalign_once(s);
}

}  // namespace mf
