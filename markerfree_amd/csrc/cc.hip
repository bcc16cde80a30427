// Stretched cross-correlation (CalcCC.cu) and the projection-matching
// measurement (ProjAlign.cpp:243-375) on the GPU.
//
// The reference runs one pair at a time with two cuFFT plan creations, ~6
// allocations and a row-wise D2H of the whole correlation map per pair
// (CalcCC.cu:237-271, 366-379), then scans the map on the host.  Here every
// pair of a coarse-alignment sweep is processed in the same launches, the map
// never leaves HBM and the peak search and parabolic fit run on the device.
#include <cmath>
#include <string>
#include <vector>

#include "cc.h"
#include "context.h"
#include "devutil.h"
#include "fft.h"

namespace mf {

void stretch_matrix(float stretch, float axis_deg, float m[3]) {
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

const float* cc_mask(mf_ctx* ctx, int w, int h, bool pa) {
  std::string key = "mask/" + std::to_string(w) + "/" + std::to_string(h) +
                    (pa ? "/pa" : "/cc");
  return (const float*)ctx_table(ctx, key, [&] {
    // GRoundEdge weights: 0 outside the ellipse, else
    // 1 - (0.5(1-cos pi r))^4.
    const float cx = (float)w * 0.5f, cy = (float)h * 0.5f;
    const float sx = pa ? (float)pad_x(w) : (float)w * 1.0f;
    const float sy = pa ? (float)h : (float)h * 1.0f;
    std::vector<float> m((size_t)w * h);
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        float fx = 2 * std::fabs((float)x - cx) / sx;
        float fy = 2 * std::fabs((float)y - cy) / sy;
        float fr = std::sqrt(fx * fx + fy * fy);
        float v = 0.0f;
        if (fr < 1.0f) {
          fr = 0.5f * (1 - hcos(3.1415926f * fr));
          v = 1.0f - (float)std::pow((double)fr, 4.0);
        }
        m[(size_t)y * w + x] = v;
      }
    return m;
  });
}

size_t stretch_cc_work_floats(int w, int h, int npair) {
  return (size_t)3 * npair * h * pad_x(w);
}

namespace {

// Partial blocks per moment job and per peak map; the combining step uses
// one 64-lane wave, one partial per lane.
constexpr int kParts = 64;

// Pixel of an unpadded w-wide image read through the padded pitch the
// reference uses (pad columns hold zero here; the reference leaves them
// uninitialised, quirk Q9).
__device__ __forceinline__ float pix(const float* img, int w, int x, int y) {
  return x < w ? img[(size_t)y * w + x] : 0.0f;
}

// Grandomfill (CalcCC.cu:23-46): window 31, next = 7 next mod 961.
__device__ float cc_random(float fx, float fy, int pitch, int h, int w,
                           const float* img) {
  int x = (int)fabsf(fx);
  int y = (int)fabsf(fy);
  if (x >= pitch) x = pitch - 1 - (x % (pitch - 1));
  if (y >= h) y = h - 1 - (y % (h - 1));
  unsigned next = (unsigned)(y * pitch + x);
  for (int i = 0; i < 961; i++) {
    next = (next * 7u) % 961u;
    int ix = (int)next - 15 + x;
    if (ix < 0 || ix >= pitch) continue;
    int iy = (int)(next / 31u) - 15 + y;
    if (iy < 0 || iy >= h) continue;
    return pix(img, w, ix, iy);
  }
  return pix(img, w, w / 2, h / 2);
}

// GStretchRandom (CalcCC.cu:149-171) with bilinear interpolation
// (mGBilinear, :8-21); output rows at the padded pitch.
__global__ __launch_bounds__(256) void k_stretch_pairs(
    const float* stack, int w, int h, const PairDesc* pairs, float* out,
    int fill) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= w * h) return;
  const int y = idx / w, x = idx - y * w;
  const PairDesc pd = pairs[blockIdx.y];
  const int pitch = pad_x(w);
  const float* img = stack + (size_t)pd.img * w * h;
  const float hw = 0.5f * (float)w, hh = 0.5f * (float)h;
  const float X = (float)x - hw + 0.5f;
  const float Y = (float)((double)y - 0.5 * (double)h + 0.5);
  const float buf = X * pd.m[0] + Y * pd.m[1];
  const float Y2 = X * pd.m[1] + Y * pd.m[2] + hh - 0.5f;
  const float X2 = buf + hw - 0.5f;
  float v;
  if (X2 >= 0.f && X2 < (float)(w - 1) && Y2 >= 0.f && Y2 < (float)(h - 1)) {
    const int ix = (int)X2, iy = (int)Y2;
    const float fx = X2 - (float)ix, fy = Y2 - (float)iy;
    const float f2 = 1.0f - fx, f3 = 1.0f - fy;
    const float* p = img + (size_t)iy * w + ix;
    v = p[0] * f2 * f3 + p[1] * fx * f3 + p[w] * f2 * fy + p[w + 1] * fx * fy;
  } else if (fill) {
    v = cc_random(X2, Y2, pitch, h, w, img);
  } else {
    v = kInvalid;
  }
  out[(size_t)blockIdx.y * h * pitch + (size_t)y * pitch + x] = v;
}

struct MomJob {
  const float* src;
  int pitch;
};

struct NormJob {
  const float* src;
  float* dst;
  int src_pitch;
  int mom;  // index of the (sum, sum of squares) pair
};

// Job b from a device table, or for two-job launches (jobs == nullptr) from
// the kernel arguments.  The argument structs are selected field by field:
// selecting a whole by-value argument struct makes the compiler copy both
// into per-thread scratch memory.
__device__ __forceinline__ MomJob pick_job(const MomJob* jobs,
                                           MomJob j0, MomJob j1,
                                           unsigned b) {
  if (jobs) return jobs[b];
  return MomJob{b ? j1.src : j0.src, b ? j1.pitch : j0.pitch};
}
__device__ __forceinline__ NormJob pick_job(const NormJob* jobs,
                                            NormJob j0,
                                            NormJob j1, unsigned b) {
  if (jobs) return jobs[b];
  return NormJob{b ? j1.src : j0.src, b ? j1.dst : j0.dst,
                 b ? j1.src_pitch : j0.src_pitch, b ? j1.mom : j0.mom};
}

// Partial (sum, sum of squares) of valid pixels over a band of rows: grid
// (kParts, njobs).  Jobs come from a device table, or for two-job launches
// from the kernel arguments (jobs == nullptr).  Each job's kParts partials
// are combined in a fixed order by k_norm_mask, so results are reproducible.
__global__ __launch_bounds__(256) void k_moments_part(const MomJob* jobs,
                                                      MomJob j0, MomJob j1,
                                                      int w, int h,
                                                      double* part) {
  __shared__ double sh[16];
  const MomJob j = pick_job(jobs, j0, j1, blockIdx.y);
  const int rows = cdiv(h, kParts);
  const int ya = blockIdx.x * rows, yb = min(h, ya + rows);
  double s1 = 0.0, s2 = 0.0;
  for (int y = ya; y < yb; y++) {
    const float* row = j.src + (size_t)y * j.pitch;
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
      const float v = row[x];
      if (v < -1e10f) continue;
      s1 += v;
      s2 += (double)(v * v);
    }
  }
  s1 = block_reduce(s1, sh, [](double a, double b) { return a + b; }, 0.0);
  s2 = block_reduce(s2, sh, [](double a, double b) { return a + b; }, 0.0);
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)blockIdx.y * kParts + blockIdx.x) * 2;
    o[0] = s1;
    o[1] = s2;
  }
}

// mNormalize / PANorm normalisation (GNormalize, CalcCC.cu:137-147) fused
// with GRoundEdge (mask table): invalid pixels become 0.
__global__ __launch_bounds__(256) void k_norm_mask(const NormJob* jobs,
                                                   NormJob j0, NormJob j1,
                                                   int w, int h,
                                                   const double* part,
                                                   const float* mask,
                                                   int unit_std) {
  __shared__ float sh_ms[2];
  const NormJob j = pick_job(jobs, j0, j1, blockIdx.y);
  if (threadIdx.x < 64) {
    // kParts (= 64) partials of this job, summed in a fixed tree order.
    const double* p = part + ((size_t)j.mom * kParts + threadIdx.x) * 2;
    double s1 = p[0], s2 = p[1];
    for (int o = 32; o > 0; o >>= 1) {
      s1 += __shfl_down(s1, o, 64);
      s2 += __shfl_down(s2, o, 64);
    }
    if (threadIdx.x == 0) {
      const double npx = (double)w * (double)h;
      const float mean = (float)(s1 / npx);
      float sd = 1.0f;
      if (!unit_std) {
        const float m2 = (float)(s2 / npx);
        const float var = m2 - mean * mean;
        sd = var <= 0.f ? 0.0f : sqrtf(var);
      }
      sh_ms[0] = mean;
      sh_ms[1] = sd;
    }
  }
  __syncthreads();
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= w * h) return;
  const int y = idx / w, x = idx - y * w;
  const float mean = sh_ms[0], sd = sh_ms[1];
  const float v = j.src[(size_t)y * j.src_pitch + x];
  const float mk = mask[idx];
  float o = 0.0f;
  if (v >= -1e10f) {
    const float nv = (v - mean) / sd;
    if (!(nv < -1e10f) && mk != 0.0f) o = nv * mk;
  }
  j.dst[(size_t)y * pad_x(w) + x] = o;
}

// GConv + GWiener + GCenterOrigin (CalcCC.cu:95-135): b = conj(a) * b,
// times exp(-2B (k^2+y^2)/(W^2+H^2)) / sqrt(|b| + 0.01), times (-1)^(k+y).
__global__ __launch_bounds__(256) void k_spectral(const float2* a, float2* b,
                                                  int w, int h,
                                                  size_t img_stride_c,
                                                  float bfactor) {
  const int nc = w / 2 + 1;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nc * h) return;
  const int y = idx / nc, k = idx - y * nc;
  const size_t i = (size_t)blockIdx.y * img_stride_c + idx;
  const float2 r = a[i], g = b[i];
  float re = r.x * g.x + r.y * g.y;
  float im = r.x * g.y - r.y * g.x;
  const int yy = (y > h / 2) ? y - h : y;
  const float fnx = 2.0f * (float)(nc - 1);
  const float f0 = -2.0f * bfactor / (fnx * fnx + (float)(h * h));
  float f = expf(f0 * (float)(unsigned)(k * k + yy * yy));
  const float amp = sqrtf(re * re + im * im);
  f /= sqrtf(amp + 0.01f);
  re *= f;
  im *= f;
  if ((k + y) % 2 != 0) {
    re = -re;
    im = -im;
  }
  b[i] = make_float2(re, im);
}

struct Best {
  float v;
  int i;
};
__device__ __forceinline__ Best best_of(Best a, Best b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

// Search window of the peak: mode 0 CalcCC::FindPeak (CalcCC.cu:381-435),
// mode 1 PAFindPeak (ProjAlign.cpp:325-375).
struct PeakWin {
  int x0, y0, x1, y1;
  float init;
  __device__ PeakWin(int w, int h, int mode) {
    if (mode == 0) {
      const int sx = w * 8 / 20 * 2, sy = h * 8 / 20 * 2;
      x0 = (w - sx) / 2;
      y0 = (h - sy) / 2;
      x1 = x0 + sx;
      y1 = y0 + sy;
      init = -1e30f;
    } else {
      x0 = max(w / 20, 3);
      y0 = max(h / 20, 3);
      x1 = w - x0;
      y1 = h - y0;
      init = -1e20f;
    }
  }
};

// Wave argmax with the smallest row-major index winning ties.
__device__ __forceinline__ Best wave_best(Best b) {
  for (int o = 32; o > 0; o >>= 1) {
    Best c{__shfl_down(b.v, o, 64), __shfl_down(b.i, o, 64)};
    b = best_of(b, c);
  }
  return b;
}

// Stage 1: argmax over a band of the window's rows, grid (kParts, nmaps).
__global__ __launch_bounds__(256) void k_peak_part(const float* maps, int w,
                                                   int h, size_t img_stride,
                                                   int mode, Best* part) {
  __shared__ Best sh[4];
  const int pitch = pad_x(w);
  const float* cc = maps + (size_t)blockIdx.y * img_stride;
  const PeakWin win(w, h, mode);
  const int rows = cdiv(win.y1 - win.y0, kParts);
  const int ya = win.y0 + blockIdx.x * rows, yb = min(win.y1, ya + rows);
  Best b{win.init, 0x7fffffff};
  for (int y = ya; y < yb; y++)
    for (int x = win.x0 + threadIdx.x; x < win.x1; x += blockDim.x) {
      const float v = cc[(size_t)y * pitch + x];
      if (v > b.v) b = Best{v, y * w + x};
    }
  b = wave_best(b);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = b;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < (int)(blockDim.x >> 6); k++) b = best_of(b, sh[k]);
  part[(size_t)blockIdx.y * kParts + blockIdx.x] = b;
}

// Stage 2 (one wave per map): combine the partials, then the parabolic
// sub-pixel fit of the reference (FP32 terms in mode 0, FP64 in mode 1).
__global__ __launch_bounds__(64) void k_peak_fit(const float* maps, int w,
                                                 int h, size_t img_stride,
                                                 int mode, const Best* part,
                                                 float* out) {
  const int pitch = pad_x(w);
  const float* cc = maps + (size_t)blockIdx.x * img_stride;
  const PeakWin win(w, h, mode);
  Best b = wave_best(part[(size_t)blockIdx.x * kParts + threadIdx.x]);
  if (threadIdx.x != 0) return;
  int px = mode == 0 ? 0 : win.x0, py = mode == 0 ? 0 : win.y0;
  if (b.i != 0x7fffffff) {
    px = b.i % w;
    py = b.i / w;
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
    A = (double)(at(px + 1, py) + at(px - 1, py)) * 0.5 - at(px, py);
    B = (double)(at(px + 1, py) - at(px - 1, py)) * 0.5;
    C = (double)(at(px, py + 1) + at(px, py - 1)) * 0.5 - at(px, py);
    D = (double)(at(px, py + 1) - at(px, py - 1)) * 0.5;
  }
  double dx = -B / (2 * A + 1e-30), dy = -D / (2 * C + 1e-30);
  if (fabs(dx) > 1) dx = 0;
  if (fabs(dy) > 1) dy = 0;
  const float fpx = (float)(px + dx), fpy = (float)(py + dy);
  out[2 * blockIdx.x] = fpx - (float)(w / 2);
  out[2 * blockIdx.x + 1] = fpy - (float)(h / 2);
}

void find_peaks(mf_ctx* ctx, hipStream_t st, const float* maps, int w, int h,
                size_t img_stride, int nmaps, int mode, Best* part,
                float* out) {
  TimedLaunch tl(ctx->timer, "cc_peak", st);
  hipLaunchKernelGGL(k_peak_part, dim3(kParts, nmaps), dim3(256), 0, st, maps,
                     w, h, img_stride, mode, part);
  hipLaunchKernelGGL(k_peak_fit, dim3(nmaps), dim3(64), 0, st, maps, w, h,
                     img_stride, mode, (const Best*)part, out);
  MF_HIP(hipGetLastError());
}

}  // namespace

void stretch_cc_pairs(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                      int w, int h, const std::vector<PairDesc>& pairs,
                      float bfactor, float* d_work, double* d_mom,
                      DevBuf& tab, float* d_shift) {
  const int npair = (int)pairs.size();
  if (npair <= 0) return;
  const int pitch = pad_x(w);
  const size_t img = (size_t)h * pitch;
  float* A = d_work;                    // normalised, masked references
  float* B = d_work + npair * img;      // normalised, masked stretched images
  float* S = d_work + 2 * npair * img;  // stretched images
  std::vector<MomJob> mj(2 * npair);
  std::vector<NormJob> nj(2 * npair);
  for (int p = 0; p < npair; p++) {
    mj[p] = MomJob{d_stack + (size_t)pairs[p].ref * w * h, w};
    mj[npair + p] = MomJob{S + p * img, pitch};
    nj[p] = NormJob{mj[p].src, A + p * img, w, p};
    nj[npair + p] = NormJob{S + p * img, B + p * img, pitch, npair + p};
  }
  const size_t b0 = npair * sizeof(PairDesc), b1 = mj.size() * sizeof(MomJob),
               b2 = nj.size() * sizeof(NormJob);
  char* dt = (char*)tab.ensure(ctx, b0 + b1 + b2);
  MF_HIP(hipMemcpyAsync(dt, pairs.data(), b0, hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(dt + b0, mj.data(), b1, hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(dt + b0 + b1, nj.data(), b2, hipMemcpyHostToDevice, st));
  const PairDesc* d_pairs = (const PairDesc*)dt;
  const dim3 g2(cdiv((long long)w * h, 256), npair);
  hipLaunchKernelGGL(k_stretch_pairs, g2, dim3(256), 0, st, d_stack, w, h,
                     d_pairs, S, 1);
  MF_HIP(hipGetLastError());
  const MomJob m0{};
  const NormJob n0{};
  hipLaunchKernelGGL(k_moments_part, dim3(kParts, 2 * npair), dim3(256), 0, st,
                     (const MomJob*)(dt + b0), m0, m0, w, h, d_mom);
  const float* mask = cc_mask(ctx, w, h, false);
  hipLaunchKernelGGL(k_norm_mask, dim3(cdiv((long long)w * h, 256), 2 * npair),
                     dim3(256), 0, st, (const NormJob*)(dt + b0 + b1), n0, n0,
                     w, h, d_mom, mask, 0);
  MF_HIP(hipGetLastError());
  // A and B are adjacent: one batched forward transform, scaled 1/(w h)
  // (cufft2D::ForwardFFT with bNorm).
  fft_r2c_2d(ctx, st, A, w, h, 2 * npair, img, (float)(1.0 / w / h));
  hipLaunchKernelGGL(k_spectral, dim3(cdiv((long long)(w / 2 + 1) * h, 256), npair),
                     dim3(256), 0, st, (const float2*)A, (float2*)B, w, h,
                     img / 2, bfactor);
  MF_HIP(hipGetLastError());
  fft_c2r_2d(ctx, st, B, w, h, npair, img);
  find_peaks(ctx, st, B, w, h, img, npair, 0,
             (Best*)(d_mom + (size_t)4 * npair * kParts), d_shift);
  MF_HIP(hipStreamSynchronize(st));  // 'tab' is reusable after return
}

size_t cc_scratch_doubles(int npair) {
  // Moment partials: 2 jobs x kParts x (sum, sum of squares) per pair; peak
  // partials: kParts x (value, index) per pair.
  return (size_t)5 * npair * kParts;
}

void pa_measure(mf_ctx* ctx, hipStream_t st, const float* d_reproj,
                const float* d_proj, int w, int h, float* d_work,
                double* d_mom, float* d_ishift) {
  const int pitch = pad_x(w);
  const size_t img = (size_t)h * pitch;
  float* A = d_work;
  float* B = d_work + img;
  const MomJob m0{d_reproj, w}, m1{d_proj, w};
  const NormJob n0{d_reproj, A, w, 0}, n1{d_proj, B, w, 1};
  hipLaunchKernelGGL(k_moments_part, dim3(kParts, 2), dim3(256), 0, st,
                     nullptr, m0, m1, w, h, d_mom);
  const float* mask = cc_mask(ctx, w, h, true);
  hipLaunchKernelGGL(k_norm_mask, dim3(cdiv((long long)w * h, 256), 2),
                     dim3(256), 0, st, nullptr, n0, n1, w, h, d_mom, mask, 1);
  MF_HIP(hipGetLastError());
  fft_r2c_2d(ctx, st, A, w, h, 2, img, 1.0f);
  hipLaunchKernelGGL(k_spectral, dim3(cdiv((long long)(w / 2 + 1) * h, 256), 1),
                     dim3(256), 0, st, (const float2*)A, (float2*)B, w, h,
                     img / 2, 500.0f);
  MF_HIP(hipGetLastError());
  fft_c2r_2d(ctx, st, B, w, h, 1, img);
  find_peaks(ctx, st, B, w, h, img, 1, 1, (Best*)(d_mom + (size_t)4 * kParts),
             d_ishift);
}

}  // namespace mf
