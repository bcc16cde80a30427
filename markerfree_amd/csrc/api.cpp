// extern "C" entry points: argument checks, exception-to-status translation.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <new>

#include "bprp.h"
#include "cc.h"
#include "context.h"
#include "fft.h"
#include "series.h"
#include "stages.h"

using namespace mf;

namespace {

// Runs f; a thrown error becomes a status and its message.  HIP keeps the
// last failed call's error per thread and hands it to the next
// hipGetLastError, so it is cleared here: the failure has been reported, and
// the next call on this thread (another request of a server worker, for
// example) must not inherit it.
template <typename F>
mf_status guard_msg(std::string* msg, F&& f) {
  mf_status st = MF_OK;
  try {
    f();
    return MF_OK;
  } catch (const Error& e) {
    *msg = e.what();
    st = e.code;
  } catch (const std::bad_alloc&) {
    *msg = "host allocation failed";
    st = MF_ENOMEM;
  } catch (const std::exception& e) {
    *msg = e.what();
    st = MF_EDEVICE;
  }
  (void)hipGetLastError();
  return st;
}

void set_ctx_error(mf_ctx* ctx, const std::string& msg) {
  if (!ctx) return;
  std::lock_guard<std::mutex> g(ctx->err_mu);
  ctx->last_error = msg;
}

template <typename F>
mf_status guard(mf_ctx* ctx, F&& f) {
  std::string msg;
  const mf_status st = guard_msg(&msg, f);
  if (st != MF_OK) set_ctx_error(ctx, msg);
  return st;
}

struct DeviceScope {
  explicit DeviceScope(mf_ctx* ctx) { MF_HIP(hipSetDevice(ctx->device)); }
};

}  // namespace

extern "C" {

const char* mf_version(void) { return "markerfree-mi355x 0.1"; }

int mf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

mf_status mf_ctx_create(int device, size_t budget, mf_ctx** out) {
  if (!out) return MF_EINVAL;
  *out = nullptr;
  mf_ctx* ctx = nullptr;
  try {
    ctx = new mf_ctx;
  } catch (...) {
    return MF_ENOMEM;
  }
  mf_status st = guard(ctx, [&] {
    int n = 0;
    MF_HIP(hipGetDeviceCount(&n));
    MF_REQUIRE(device >= 0 && device < n, "no such HIP device");
    ctx->device = device;
    MF_HIP(hipSetDevice(device));
    MF_HIP(hipDeviceGetAttribute(&ctx->num_cu,
                                 hipDeviceAttributeMultiprocessorCount, device));
    MF_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->budget = budget;
    ctx->fft.reset(new FftCache(ctx));
    ctx->peak_scratch = peak_scratch_address(device);
  });
  if (st != MF_OK) {
    delete ctx;
    return st;
  }
  *out = ctx;
  return MF_OK;
}

void mf_ctx_destroy(mf_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  delete ctx;
}

const char* mf_last_error(const mf_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

mf_status mf_ctx_synchronize(mf_ctx* ctx) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_ctx_memory(mf_ctx* ctx, size_t* in_use, size_t* peak,
                        int reset_peak) {
  if (!ctx) return MF_EINVAL;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (in_use) *in_use = ctx->in_use;
  if (peak) *peak = ctx->peak;
  if (reset_peak) ctx->peak = ctx->in_use;
  return MF_OK;
}

mf_status mf_dev_malloc(mf_ctx* ctx, size_t bytes, void** d_out) {
  if (!ctx || !d_out) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    void* p = dev_alloc(ctx, bytes);
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->user_allocs[p] = bytes;
    *d_out = p;
  });
}

mf_status mf_dev_free(mf_ctx* ctx, void* p) {
  if (!ctx) return MF_EINVAL;
  if (!p) return MF_OK;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    size_t bytes = 0;
    {
      std::lock_guard<std::mutex> g(ctx->mu);
      auto it = ctx->user_allocs.find(p);
      MF_REQUIRE(it != ctx->user_allocs.end(), "pointer not from mf_dev_malloc");
      bytes = it->second;
      ctx->user_allocs.erase(it);
    }
    MF_HIP(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, p, bytes);
  });
}

mf_status mf_host_alloc(size_t bytes, void** h_out) {
  if (!h_out) return MF_EINVAL;
  *h_out = nullptr;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();  // reported here; later checks must not see it
    return MF_ENOMEM;
  }
  *h_out = p;
  return MF_OK;
}

mf_status mf_host_free(void* p) {
  if (!p) return MF_OK;
  return hipHostFree(p) == hipSuccess ? MF_OK : MF_EINVAL;
}

mf_status mf_h2d(mf_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                          ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_d2h(mf_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                          ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_d2d(mf_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    MF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                          ctx->stream));
  });
}

mf_status mf_timing_enable(mf_ctx* ctx, int enable) {
  if (!ctx) return MF_EINVAL;
  ctx->timer.enable(enable != 0);
  return MF_OK;
}

mf_status mf_tune_set(mf_ctx* ctx, const char* key, int value) {
  if (!ctx || !key) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(value >= 0, "tunable values are positive (0 removes the key)");
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    if (value == 0)
      ctx->tune.erase(key);
    else
      ctx->tune[key] = value;
    ctx->tune_gen++;
  });
}

mf_status mf_tune_get(mf_ctx* ctx, const char* key, int* value) {
  if (!ctx || !key || !value) return MF_EINVAL;
  std::lock_guard<std::mutex> g(ctx->tune_mu);
  auto it = ctx->tune.find(key);
  if (it == ctx->tune.end()) return MF_EINVAL;
  *value = it->second;
  return MF_OK;
}

mf_status mf_tune_load(mf_ctx* ctx, const char* path, int* count) {
  if (!ctx || !path) return MF_EINVAL;
  return guard(ctx, [&] {
    FILE* f = fopen(path, "rb");
    if (!f) throw Error(MF_EIO, std::string("cannot read ") + path);
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    fclose(f);
    // A flat object of string keys and positive integer values.
    std::map<std::string, int> kv;
    size_t i = text.find('{');
    MF_REQUIRE(i != std::string::npos, "tuning file is not a JSON object");
    auto skip_ws = [&] {
      while (i < text.size() && isspace((unsigned char)text[i])) i++;
    };
    i++;
    skip_ws();
    while (i < text.size() && text[i] != '}') {
      MF_REQUIRE(text[i] == '"', "tuning file: expected a string key");
      const size_t e = text.find('"', i + 1);
      MF_REQUIRE(e != std::string::npos, "tuning file: unterminated key");
      const std::string key = text.substr(i + 1, e - i - 1);
      i = e + 1;
      skip_ws();
      MF_REQUIRE(i < text.size() && text[i] == ':', "tuning file: expected ':'");
      i++;
      skip_ws();
      size_t used = 0;
      long v = 0;
      try {
        v = std::stol(text.substr(i, 24), &used);
      } catch (...) {
        used = 0;
      }
      MF_REQUIRE(used > 0 && v > 0 && v < (1L << 30),
                 "tuning file: values are positive integers");
      kv[key] = (int)v;
      i += used;
      skip_ws();
      if (i < text.size() && text[i] == ',') {
        i++;
        skip_ws();
      }
    }
    MF_REQUIRE(i < text.size(), "tuning file: unterminated object");
    std::lock_guard<std::mutex> g(ctx->tune_mu);
    for (auto& p : kv) ctx->tune[p.first] = p.second;
    ctx->tune_gen++;
    if (count) *count = (int)kv.size();
  });
}

mf_status mf_timing_select(mf_ctx* ctx, const char* name) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] { ctx->timer.select(name ? name : ""); });
}

mf_status mf_timing_every(mf_ctx* ctx, int k) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(k >= 1, "timing stride must be at least 1");
    ctx->timer.every(k);
  });
}

mf_status mf_kernel_count(mf_ctx* ctx, const char* name, long long* launches) {
  if (!ctx || !name || !launches) return MF_EINVAL;
  return guard(ctx, [&] { *launches = ctx->timer.count(name); });
}

mf_status mf_timing_reset(mf_ctx* ctx) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    ctx->timer.reset();
  });
}

mf_status mf_kernel_time(mf_ctx* ctx, const char* name, double* total_ms,
                         long long* launches) {
  if (!ctx || !name || !total_ms || !launches) return MF_EINVAL;
  return guard(ctx, [&] {
    DeviceScope d(ctx);
    ctx->timer.query(name, total_ms, launches);
  });
}

mf_status mf_find_proj_range(const float* tilts, int nz,
                             const unsigned char* skip, int target, int nproj,
                             int* range2) {
  if (!tilts || !skip || !range2 || nz <= 0 || target < 0 || target >= nz ||
      nproj <= 0)
    return MF_EINVAL;
  find_proj_range(tilts, nz, skip, target, nproj, range2);
  return MF_OK;
}

mf_status mf_reproject(mf_ctx* ctx, const float* d_stack, int bx, int by,
                       int nz, int tb, const float* tilts,
                       const unsigned char* skip, int target, int nproj,
                       int mode, float* d_reproj, float* d_slab) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_stack && tilts && skip && d_reproj, "null argument");
    MF_REQUIRE(mode >= 0 && mode <= 8, "mode must be 0..8");
    DeviceScope d(ctx);
    reproject_target(ctx, ctx->stream, d_stack, bx, by, nz, tb, tilts, skip,
                     target, nproj, mode, d_reproj, d_slab);
  });
}

mf_status mf_fft_r2c_2d(mf_ctx* ctx, float* d_buf, int nx, int ny, int nimg,
                        float scale) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_buf && nx >= 2 && nx % 2 == 0 && ny >= 1 && nimg >= 1,
               "bad 2-D FFT arguments");
    DeviceScope d(ctx);
    fft_r2c_2d(ctx, ctx->stream, d_buf, nx, ny, nimg,
               (size_t)pad_x(nx) * ny, scale);
  });
}

mf_status mf_fft_c2r_2d(mf_ctx* ctx, float* d_buf, int nx, int ny, int nimg) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_buf && nx >= 2 && nx % 2 == 0 && ny >= 1 && nimg >= 1,
               "bad 2-D FFT arguments");
    DeviceScope d(ctx);
    fft_c2r_2d(ctx, ctx->stream, d_buf, nx, ny, nimg, (size_t)pad_x(nx) * ny);
  });
}

mf_status mf_fft_r2c_1d(mf_ctx* ctx, float* d_buf, int n, int nlines,
                        int pitch) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_buf && n >= 2 && n % 2 == 0 && nlines >= 1 &&
                   pitch >= pad_x(n) && pitch % 2 == 0,
               "bad 1-D FFT arguments");
    DeviceScope d(ctx);
    fft_r2c_rows(ctx, ctx->stream, d_buf, n, nlines, pitch, 1.0f);
  });
}

mf_status mf_fft_c2r_1d(mf_ctx* ctx, float* d_buf, int n, int nlines,
                        int pitch) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_buf && n >= 2 && n % 2 == 0 && nlines >= 1 &&
                   pitch >= pad_x(n) && pitch % 2 == 0,
               "bad 1-D FFT arguments");
    DeviceScope d(ctx);
    fft_c2r_rows(ctx, ctx->stream, d_buf, n, nlines, pitch, 1.0f);
  });
}

mf_status mf_correct_proj(mf_ctx* ctx, const float* d_raw, int nx, int ny,
                          float shift_x, float shift_y, float axis_deg,
                          int shift_only, int random_fill, int bin,
                          int rweight, float* d_out) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_raw && d_out && nx > 1 && ny > 1 && bin >= 1, "bad arguments");
    DeviceScope d(ctx);
    const CorrGeom g = CorrGeom::make(nx, ny, bin);
    CorrProj p;
    const float th = shift_only ? 0.0f : (float)(axis_deg * kD2R);
    p.shift_x = shift_x;
    p.shift_y = shift_y;
    p.cos_a = hcos(th);
    p.sin_a = hsin(th);
    DevBuf pb;
    CorrProj* dp = (CorrProj*)pb.ensure(ctx, sizeof(CorrProj));
    MF_HIP(hipMemcpyAsync(dp, &p, sizeof(p), hipMemcpyHostToDevice, ctx->stream));
    if (rweight) {
      correct_rweight(ctx, ctx->stream, d_raw, g, dp, 0, 1, random_fill != 0,
                      d_out);
    } else {
      correct_bin(ctx, ctx->stream, d_raw, g, dp, 0, 1, random_fill != 0,
                  d_out, g.bx, (size_t)g.bx * g.by);
    }
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_stretch_cc(mf_ctx* ctx, const float* d_ref, const float* d_img,
                        int bx, int by, float ref_tilt, float tilt,
                        float axis_deg, float bfactor, float* shift2) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_ref && d_img && shift2 && bx >= 4 && by >= 4 && bx % 2 == 0,
               "bad arguments");
    DeviceScope d(ctx);
    const size_t plane = (size_t)bx * by;
    DevBuf stk, work, mom, tab, sh;
    float* s2 = (float*)stk.ensure(ctx, 2 * plane * sizeof(float));
    MF_HIP(hipMemcpyAsync(s2, d_ref, plane * 4, hipMemcpyDeviceToDevice, ctx->stream));
    MF_HIP(hipMemcpyAsync(s2 + plane, d_img, plane * 4, hipMemcpyDeviceToDevice,
                          ctx->stream));
    PairDesc p;
    p.ref = 0;
    p.img = 1;
    double ds = std::cos(kD2R * ref_tilt) / std::cos(kD2R * tilt);
    stretch_matrix((float)ds, axis_deg, p.m);
    float* w = (float*)work.ensure(ctx, stretch_cc_work_floats(bx, by, 1) * 4);
    double* m = (double*)mom.ensure(ctx, cc_scratch_doubles(1, by) * sizeof(double));
    float* dsh = (float*)sh.ensure(ctx, 2 * sizeof(float));
    stretch_cc_pairs(ctx, ctx->stream, s2, bx, by, {p}, bfactor, w, m, tab, dsh);
    MF_HIP(hipMemcpy(shift2, dsh, 2 * sizeof(float), hipMemcpyDeviceToHost));
  });
}

mf_status mf_pa_measure(mf_ctx* ctx, const float* d_reproj,
                        const float* d_proj, int bx, int by, float* shift2) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_reproj && d_proj && shift2 && bx >= 8 && by >= 8 &&
                   bx % 2 == 0,
               "bad arguments");
    DeviceScope d(ctx);
    DevBuf work, mom, sh;
    float* w = (float*)work.ensure(ctx, 2 * (size_t)by * pad_x(bx) * 4);
    double* m = (double*)mom.ensure(ctx, cc_scratch_doubles(1, by) * sizeof(double));
    float* dsh = (float*)sh.ensure(ctx, 2 * sizeof(float));
    pa_measure(ctx, ctx->stream, d_reproj, d_proj, bx, by, w, m, dsh);
    MF_HIP(hipMemcpyAsync(shift2, dsh, 2 * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_common_lines(mf_ctx* ctx, const float* d_stack, int nx, int ny,
                          int nz, const float* tilts, const float* shift_x,
                          const float* shift_y, float axis0, float range,
                          int num, float* d_lines, int* line_len) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_stack && tilts && shift_x && shift_y && line_len && nx > 0 &&
                   ny > 0 && nz > 0 && num >= 2,
               "bad arguments");
    DeviceScope d(ctx);
    const AxisSetup a = AxisSetup::make(nx, ny, axis0, range, num);
    MF_REQUIRE(a.L >= 4, "common-line length too small for this image size");
    *line_len = a.L;
    if (!d_lines) return;
    std::vector<double> sx(shift_x, shift_x + nz), sy(shift_y, shift_y + nz);
    AxisWork ws;
    float* lines = common_lines(ctx, ctx->stream, d_stack, nx, ny, nz, tilts,
                                sx.data(), sy.data(), a, ws);
    MF_HIP(hipMemcpy2DAsync(d_lines, (size_t)a.L * 4, lines,
                            (size_t)(a.L + 2) * 4, (size_t)a.L * 4,
                            (size_t)nz * num, hipMemcpyDeviceToDevice,
                            ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_line_scores(mf_ctx* ctx, const float* d_stack, int nx, int ny,
                         int nz, const float* tilts, const float* shift_x,
                         const float* shift_y, float axis0, float range,
                         int num, float* scores, float* d_spectra,
                         int* line_len, float* best_axis) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_stack && tilts && shift_x && shift_y && nx > 0 && ny > 0 &&
                   nz > 1 && num >= 2,
               "bad arguments");
    DeviceScope d(ctx);
    const AxisSetup a = AxisSetup::make(nx, ny, axis0, range, num);
    MF_REQUIRE(a.L >= 4, "common-line length too small for this image size");
    std::vector<double> sx(shift_x, shift_x + nz), sy(shift_y, shift_y + nz);
    AxisWork ws;
    const AxisResult r = axis_search(ctx, ctx->stream, d_stack, nx, ny, nz,
                                     tilts, sx.data(), sy.data(), axis0, range,
                                     num, ws);
    if (scores) std::copy(r.scores.begin(), r.scores.end(), scores);
    if (line_len) *line_len = r.L;
    if (best_axis) *best_axis = r.axis;
    if (d_spectra) {
      // The spectra sit in the lines' buffer at the lines' pitch (L + 2
      // floats), which is the spectra's own (L/2 + 1 complex values).
      MF_HIP(hipMemcpyAsync(d_spectra, ws.lines.get(),
                            (size_t)nz * num * (r.L + 2) * sizeof(float),
                            hipMemcpyDeviceToDevice, ctx->stream));
      MF_HIP(hipStreamSynchronize(ctx->stream));
    }
  });
}

mf_status mf_ncc_moments(mf_ctx* ctx, const float* d_ref, const float* d_img,
                         int w, int h, float ref_tilt, float tilt,
                         double* moments6, float* cc) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_ref && d_img && w >= 2 && h >= 2, "bad arguments");
    DeviceScope d(ctx);
    OffsetWork ws;
    double m[6];
    ncc_moments(ctx, ctx->stream, d_ref, d_img, w, h, ref_tilt, tilt, ws, m);
    if (moments6) std::copy(m, m + 6, moments6);
    if (cc) *cc = offset_cc(m);
  });
}

mf_status mf_offset_scan(mf_ctx* ctx, const float* d_corr, int w, int h,
                         int nz, const float* tilts, float* ccs, float* best) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_corr && tilts && best && w >= 2 && h >= 2 && nz >= 2,
               "bad arguments");
    DeviceScope d(ctx);
    OffsetWork ws;
    std::vector<float> v;
    *best = offset_search(ctx, ctx->stream, d_corr, w, h, nz, tilts, ws, &v);
    if (ccs) std::copy(v.begin(), v.end(), ccs);
  });
}

mf_status mf_peak2d(mf_ctx* ctx, const float* d_map, int w, int h, int pitch,
                    int mode, float* shift2) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_map && shift2, "bad arguments");
    DeviceScope d(ctx);
    DevBuf ws, out;
    float* dsh = (float*)out.ensure(ctx, 2 * sizeof(float));
    peak2d(ctx, ctx->stream, d_map, w, h, pitch, mode, ws, dsh);
    MF_HIP(hipMemcpyAsync(shift2, dsh, 2 * sizeof(float),
                          hipMemcpyDeviceToHost, ctx->stream));
    MF_HIP(hipStreamSynchronize(ctx->stream));
  });
}

mf_status mf_peak2d_noalloc(mf_ctx* ctx, const float* d_map, int w, int h,
                            int pitch, int mode, float* shift2) {
  if (!ctx) return MF_EINVAL;
  return guard(ctx, [&] {
    MF_REQUIRE(d_map && shift2, "bad arguments");
    DeviceScope d(ctx);
    peak2d_noalloc(ctx, ctx->stream, d_map, w, h, pitch, mode, shift2);
  });
}

void mf_options_default(mf_options* o) {
  if (!o) return;
  o->offset = 0.0f;
  o->pitch_angle = 0.0f;
  o->zshift = 0.0f;
  o->thickness = 0;
  o->align_z = 600;
  o->out_bin = 1;
  o->nproj = 10;
  o->verbose = 0;
  o->reproj_mode = 8;
  o->pa_pairs = 1;
  o->graphs = 0;
}

mf_status mf_series_create(mf_ctx* ctx, int nx, int ny, int nz,
                           const float* stack, const float* tilts,
                           mf_series** out) {
  if (!ctx || !out) return MF_EINVAL;
  *out = nullptr;
  mf_series* s = nullptr;
  mf_status st = guard(ctx, [&] {
    MF_REQUIRE(tilts && nx >= 16 && ny >= 16 && nz >= 2,
               "series needs at least 2 images of 16x16");
    DeviceScope d(ctx);
    s = new mf_series;
    s->ctx = ctx;
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    s->file_tilts.assign(tilts, tilts + nz);
    MF_HIP(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    const size_t bytes = (size_t)nx * ny * nz * sizeof(float);
    float* raw = (float*)s->raw.ensure(ctx, bytes);
    if (stack) {
      MF_HIP(hipMemcpyAsync(raw, stack, bytes, hipMemcpyHostToDevice, s->st));
      MF_HIP(hipStreamSynchronize(s->st));
    }
    s->sx.assign(nz, 0.0);
    s->sy.assign(nz, 0.0);
    s->rot.assign(nz, 0.0f);
  });
  if (st != MF_OK) {
    delete s;
    return st;
  }
  *out = s;
  return MF_OK;
}

void mf_series_destroy(mf_series* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  delete s;
}

const char* mf_series_last_error(const mf_series* s) {
  return s ? s->last_error.c_str() : "null series";
}

}  // extern "C"

namespace {

template <typename F>
mf_status sguard(mf_series* s, F&& f) {
  if (!s) return MF_EINVAL;
  std::string msg;
  const mf_status st = guard_msg(&msg, [&] {
    DeviceScope d(s->ctx);
    f();
  });
  if (st != MF_OK) {
    s->last_error = msg;
    set_ctx_error(s->ctx, msg);
  }
  return st;
}

}  // namespace

extern "C" {

mf_status mf_series_upload(mf_series* s, int z0, int count, const float* host) {
  if (!s) return MF_EINVAL;
  return sguard(s, [&] {
    MF_REQUIRE(count >= 0 && z0 >= 0 && z0 + count <= s->nz,
               "image range out of bounds");
    MF_REQUIRE(count == 0 || host, "null host images");
    MF_REQUIRE(!s->preprocessed, "the raw stack is gone after preprocessing");
    DeviceScope d(s->ctx);
    for (hipEvent_t& e : s->up_ev)
      if (!e) MF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // Upload k records event k mod 2, then waits for upload k-1.
    const long long k = s->up_n;
    if (count > 0) {
      const size_t plane = (size_t)s->nx * s->ny;
      float* raw = (float*)s->raw.get() + (size_t)z0 * plane;
      MF_HIP(hipMemcpyAsync(raw, host, (size_t)count * plane * sizeof(float),
                            hipMemcpyHostToDevice, s->st));
      MF_HIP(hipEventRecord(s->up_ev[k & 1], s->st));
      s->up_n = k + 1;
    }
    if (k > 0) MF_HIP(hipEventSynchronize(s->up_ev[(k - 1) & 1]));
  });
}


mf_status mf_series_set_verbose(mf_series* s, int verbose) {
  return sguard(s, [&] { s->verbose = verbose; });
}

mf_status mf_series_trim(mf_series* s) {
  return sguard(s, [&] { series_trim(s); });
}

mf_status mf_series_set_graphs(mf_series* s, int on) {
  return sguard(s, [&] { s->graphs = on != 0; });
}

mf_status mf_series_graph_replays(const mf_series* s, long long* replays) {
  if (!s) return MF_EINVAL;
  if (replays) *replays = s->pa_replays;
  return MF_OK;
}

mf_status mf_preprocess(mf_series* s) {
  return sguard(s, [&] {
    if (!s->preprocessed) series_preprocess(s);
  });
}

mf_status mf_get_stack(mf_series* s, float* out) {
  return sguard(s, [&] {
    MF_REQUIRE(out, "null output");
    MF_REQUIRE(s->preprocessed, "preprocess the series first");
    MF_HIP(hipMemcpyAsync(out, s->stack.get(),
                          (size_t)s->nx * s->ny * s->nz * sizeof(float),
                          hipMemcpyDeviceToHost, s->st));
    MF_HIP(hipStreamSynchronize(s->st));
  });
}

mf_status mf_get_tilts(mf_series* s, float* tilts) {
  return sguard(s, [&] {
    MF_REQUIRE(tilts, "null output");
    const auto& t = s->preprocessed ? s->tilts : s->file_tilts;
    std::copy(t.begin(), t.end(), tilts);
  });
}

mf_status mf_set_tilts(mf_series* s, const float* tilts) {
  return sguard(s, [&] {
    MF_REQUIRE(tilts, "null input");
    s->tilts.assign(tilts, tilts + s->nz);
  });
}

mf_status mf_get_params(mf_series* s, double* shift_x, double* shift_y,
                        float* rotate, float* angle_offset) {
  return sguard(s, [&] {
    if (shift_x) std::copy(s->sx.begin(), s->sx.end(), shift_x);
    if (shift_y) std::copy(s->sy.begin(), s->sy.end(), shift_y);
    if (rotate) std::copy(s->rot.begin(), s->rot.end(), rotate);
    if (angle_offset) *angle_offset = s->offset;
  });
}

mf_status mf_set_params(mf_series* s, const double* shift_x,
                        const double* shift_y, const float* rotate,
                        float angle_offset) {
  return sguard(s, [&] {
    if (shift_x) s->sx.assign(shift_x, shift_x + s->nz);
    if (shift_y) s->sy.assign(shift_y, shift_y + s->nz);
    if (rotate) s->rot.assign(rotate, rotate + s->nz);
    s->offset = angle_offset;
  });
}

mf_status mf_coarse_setup(mf_series* s) {
  return sguard(s, [&] { series_coarse_setup(s); });
}

mf_status mf_coarse_iter(mf_series* s, float* maxerr) {
  return sguard(s, [&] {
    float e = series_coarse_iter(s);
    if (maxerr) *maxerr = e;
  });
}

mf_status mf_tiltaxis(mf_series* s, float range, int num, float* axis) {
  return sguard(s, [&] {
    float a = series_tilt_axis(s, range, num);
    if (axis) *axis = a;
  });
}

mf_status mf_find_offset(mf_series* s, float* best) {
  return sguard(s, [&] {
    float b = series_find_offset(s);
    if (best) *best = b;
  });
}

mf_status mf_apply_offset(mf_series* s) {
  return sguard(s, [&] { series_apply_offset(s); });
}

mf_status mf_projalign_setup(mf_series* s, float zshift, int align_z,
                             int nproj, int reproj_mode) {
  return sguard(s, [&] {
    MF_REQUIRE(reproj_mode >= 0 && reproj_mode <= 8, "reproj_mode is 0..8");
    s->reproj_mode = reproj_mode;
    series_projalign_setup(s, zshift, align_z, nproj);
  });
}

mf_status mf_projalign_iter(mf_series* s, float* maxerr) {
  return sguard(s, [&] {
    float e = series_projalign_iter(s);
    if (maxerr) *maxerr = e;
  });
}

mf_status mf_projalign_once(mf_series* s) {
  return sguard(s, [&] { series_projalign_once(s); });
}

mf_status mf_correct(mf_series* s, int out_bin, float* out, int* bx, int* by) {
  return sguard(s, [&] { series_correct(s, out_bin, out, bx, by); });
}

mf_status mf_correct_async(mf_series* s, int out_bin, float* out, int* bx,
                           int* by) {
  return sguard(s, [&] { series_correct(s, out_bin, out, bx, by, false); });
}

mf_status mf_series_wait(mf_series* s) {
  return sguard(s, [&] { series_wait(s); });
}

mf_status mf_run(mf_series* s, const mf_options* o) {
  return sguard(s, [&] {
    MF_REQUIRE(o, "null options");
    MF_REQUIRE(o->nproj > 0, "nProj must be > 0");
    MF_REQUIRE(o->reproj_mode >= 0 && o->reproj_mode <= 8,
               "reproj_mode is 0..8");
    series_run(s, o);
  });
}

mf_status mf_pa_measurements(mf_series* s, float* meas, float* err) {
  return sguard(s, [&] {
    MF_REQUIRE(!s->pa_err.empty(), "no projection-matching sweep has run");
    if (meas) std::copy(s->pa_meas.begin(), s->pa_meas.end(), meas);
    if (err) std::copy(s->pa_err.begin(), s->pa_err.end(), err);
  });
}

mf_status mf_axis_scores(mf_series* s, float* scores, int cap, int* num) {
  return sguard(s, [&] {
    MF_REQUIRE(!s->axis_scores.empty(), "no axis search has run");
    if (num) *num = (int)s->axis_scores.size();
    if (scores)
      std::copy(s->axis_scores.begin(),
                s->axis_scores.begin() +
                    std::min((size_t)std::max(cap, 0), s->axis_scores.size()),
                scores);
  });
}

mf_status mf_offset_scores(mf_series* s, float* ccs) {
  return sguard(s, [&] {
    MF_REQUIRE(!s->offset_scores.empty(), "no offset scan has run");
    if (ccs) std::copy(s->offset_scores.begin(), s->offset_scores.end(), ccs);
  });
}

mf_status mf_series_trace(mf_series* s, int* kinds, float* values, int cap,
                          int* count) {
  return sguard(s, [&] {
    const int n = (int)s->trace.size();
    for (int i = 0; i < n && i < cap; i++) {
      if (kinds) kinds[i] = s->trace[i].first;
      if (values) values[i] = s->trace[i].second;
    }
    if (count) *count = n;
  });
}

mf_status mf_series_trace_clear(mf_series* s) {
  return sguard(s, [&] { s->trace.clear(); });
}

}  // extern "C"
