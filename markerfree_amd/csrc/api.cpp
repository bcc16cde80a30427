// extern "C" entry points: argument checks, exception-to-status translation.
#include <cstring>
#include <new>

#include "bprp.h"
#include "context.h"
#include "fft.h"

using namespace mf;

namespace {

template <typename F>
mf_status guard(mf_ctx* ctx, F&& f) {
  try {
    f();
    return MF_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->last_error = "host allocation failed";
    return MF_ENOMEM;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return MF_EDEVICE;
  }
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
    MF_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->budget = budget;
    ctx->fft.reset(new FftCache(ctx));
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
    MF_REQUIRE(mode == 0 || mode == 1, "mode must be 0 or 1");
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

}  // extern "C"
