// One tilt series resident in HBM and the pipeline that aligns it
// (the counterpart of Process, main/Process.cpp, and of the stage classes it
// drives).
#pragma once
#include <string>
#include <vector>

#include "axis.h"
#include "context.h"
#include "stages.h"

struct mf_series {
  mf_ctx* ctx = nullptr;
  hipStream_t st = nullptr;  // each series owns a stream (batched form)
  int nx = 0, ny = 0, nz = 0;
  int verbose = 0;
  int reproj_mode = 8;       // 0 materialised slab, 1 and 8 fused
  bool pa_pairs = true;      // sweep steps may measure two targets at once
  hipEvent_t up_ev[2] = {nullptr, nullptr};  // streaming uploads (alternate)
  long long up_n = 0;                        // uploads recorded
  std::vector<float> file_tilts;  // as read (file order)
  std::vector<float> tilts;       // ascending; + offset once applied
  // AlignParam (Util.h:14-21), shifts in FP64.
  std::vector<double> sx, sy;
  std::vector<float> rot;
  float offset = 0.0f;
  bool preprocessed = false;
  mf::DevBuf raw, stack, tmp;
  // Transform (coarse alignment) state.
  bool tr_ready = false;
  int tr_bin = 1;
  std::vector<float> tr_angles;
  std::vector<double> psx, psy;
  mf::DevBuf tr_corr, tr_work, tr_mom, tr_shift, tr_tab, cproj;
  // Tilt axis / offset workspaces.
  mf::AxisWork axw;
  mf::OffsetWork offw;
  mf::DevBuf off_corr;
  // ProjAlign state.
  bool pa_ready = false;
  int pa_bin = 1, pa_tb = 0, pa_nproj = 10;
  float pa_z0 = 0.0f;
  std::vector<float> pa_angles;
  mf::CorrGeom pa_geom{};
  mf::DevBuf pa_corr, pa_reproj, pa_work, pa_mom, pa_dev, pa_slab, pa_tab;
  // Captured sweep (mf_options.graphs): within one setup every sweep makes
  // the same launches with the same arguments (only device-resident
  // parameters change), so the second sweep is captured into a HIP graph
  // and later sweeps replay it.  pa_ekey: what the last eager sweep's
  // launches depended on; pa_gkey: the same for the captured graph.
  bool graphs = false;
  unsigned long long pa_gen = 0;  // setups so far
  std::vector<unsigned long long> pa_ekey, pa_gkey;
  hipGraphExec_t pa_graph = nullptr;
  long long pa_replays = 0;       // sweeps run by graph replay
  // The last sweep's measured shift (raw pixels, [nz][2]) and its length
  // per target; targets the sweep did not measure hold 0.
  std::vector<float> pa_meas, pa_err;
  // Scores of the last axis search (per candidate) and offset scan (31).
  std::vector<float> axis_scores, offset_scores;
  // Stage trace: one (kind, value) event per stage call (MF_EV_*).
  std::vector<std::pair<int, float>> trace;
  void ev(int kind, float value) { trace.emplace_back(kind, value); }
  std::string last_error;
  ~mf_series();
};

namespace mf {

void series_preprocess(mf_series* s);
void series_coarse_setup(mf_series* s);
float series_coarse_iter(mf_series* s);
float series_tilt_axis(mf_series* s, float range, int num);
float series_find_offset(mf_series* s);
void series_apply_offset(mf_series* s);
void series_projalign_setup(mf_series* s, float zshift, int align_z,
                            int nproj);
float series_projalign_iter(mf_series* s);
void series_projalign_once(mf_series* s);
// Final correction (Process::CorrectStack); out [nz][by][bx] host, or
// nullptr to query the size only.
void series_correct(mf_series* s, int out_bin, float* out, int* bx, int* by,
                    bool wait = true);
void series_wait(mf_series* s);
void series_run(mf_series* s, const mf_options* o);
// Releases the stage workspaces (coarse, axis, offset and projection-matching
// buffers, the transposed stack); keeps the preprocessed stack, tilts and
// parameters.  A later stage call needs its setup again.
void series_trim(mf_series* s);

// Upper bound of the device bytes one series holds through series_run and
// series_correct (admission control of the served form).
size_t series_hbm_bytes(int nx, int ny, int nz, const mf_options* o);

}  // namespace mf
