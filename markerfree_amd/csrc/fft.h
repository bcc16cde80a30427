// Hand-written FFTs for gfx950: mixed-radix Stockham transforms staged in LDS,
// Bluestein's chirp-z for lengths with a prime factor above kMaxDirectPrime,
// and the real<->complex packing used by the in-place padded layout that the
// reference gets from cuFFT (method/Cufft1&2.cu:17-76, :109-165).
#pragma once
#include <map>
#include <memory>
#include <mutex>

#include "common.h"

namespace mf {

constexpr int kMaxRad = 24;
constexpr int kMaxDirectPrime = 61;

// Device-side description of a complex transform of length n, passed to
// kernels by value.
struct FftDevPlan {
  int n;       // transform length
  int blue;    // 1: Bluestein through a length-m direct transform
  int m;       // Bluestein convolution length (smooth), else n
  int nrad;    // radices of the direct transform (of n, or of m)
  int rad[kMaxRad];
  const int* drad;      // the same radices in device memory (read per stage)
  const float2* tw;     // exp(-2 pi i k / L), L = direct length
  const float2* chirp;  // Bluestein: exp(-i pi k^2 / n), k < n
  const float2* bf;     // Bluestein: FFT_m(b)/m for the forward direction
  const float2* bi;     // Bluestein: same for the inverse direction
  const float2* half;   // exp(-i pi k / n), k <= n (real-transform packing)
};

class FftCache {
 public:
  explicit FftCache(mf_ctx* ctx) : ctx_(ctx) {}
  ~FftCache();
  // Plan for complex length n (n >= 1), built once per n.
  const FftDevPlan& plan(int n);

 private:
  struct Entry {
    FftDevPlan p;
    void* mem = nullptr;
    size_t bytes = 0;
  };
  mf_ctx* ctx_;
  std::mutex mu_;
  std::map<int, std::unique_ptr<Entry>> plans_;
};

// Lines a workgroup holds for a plan under an LDS budget (bytes).
int fft_lines_per_group(const FftDevPlan& p, size_t lds_budget, int max_lines);

// 2-D in-place R2C of an nx x ny image with row pitch pad_x(nx) floats,
// batched over nimg images spaced img_stride floats apart; the result is
// multiplied by 'scale' (1 for cufftExecR2C, 1/(nx*ny) for the reference's
// bNorm option, Cufft1&2.cu:47-62).
void fft_r2c_2d(mf_ctx* ctx, hipStream_t st, float* buf, int nx, int ny,
                int nimg, size_t img_stride, float scale);
// 2-D in-place C2R (unnormalised, cufftExecC2R).
void fft_c2r_2d(mf_ctx* ctx, hipStream_t st, float* buf, int nx, int ny,
                int nimg, size_t img_stride);
// Batched 1-D in-place R2C / C2R of length n on lines 'pitch' floats apart.
void fft_r2c_rows(mf_ctx* ctx, hipStream_t st, float* buf, int n, int nlines,
                  int pitch, float scale);
void fft_c2r_rows(mf_ctx* ctx, hipStream_t st, float* buf, int n, int nlines,
                  int pitch, float scale);

}  // namespace mf
