// Stretched cross-correlation (CalcCC.cu) and the projection-matching
// measurement (ProjAlign.cpp:243-375) on the GPU.
//
// The reference runs one pair at a time with two cuFFT plan creations, ~6
// allocations and a row-wise D2H of the whole correlation map per pair
// (CalcCC.cu:237-271, 366-379), then scans the map on the host.  Here every
// pair of a coarse-alignment sweep is processed in the same launches, the map
// never leaves HBM and the peak search and parabolic fit run on the device.
#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <string>
#include <vector>

#include "cc.h"
#include "context.h"
#include "devutil.h"
#include "fft_dev.h"

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

constexpr int kCcThreads = 256;
// Partial (sum, sum of squares) blocks per image; their combination is a
// fixed-order tree, so every consumer computes the same moments.
constexpr int kMomParts = 256;
constexpr int kMomBatch = 8;

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
  const PairDesc pd = pairs[blockIdx.y];
  const int pitch = pad_x(w);
  const float* img = stack + (size_t)pd.img * w * h;
  const float hw = 0.5f * (float)w, hh = 0.5f * (float)h;
  // Rows blockIdx.x, blockIdx.x + gridDim.x, ...; the threads stride along
  // each row (no division per pixel).
  for (int y = blockIdx.x; y < h; y += gridDim.x)
  for (int x = threadIdx.x; x < w; x += blockDim.x) {
    const float X = (float)x - hw + 0.5f;
    const float Y = (float)((double)y - 0.5 * (double)h + 0.5);
    const float buf = X * pd.m[0] + Y * pd.m[1];
    const float Y2 = X * pd.m[1] + Y * pd.m[2] + hh - 0.5f;
    const float X2 = buf + hw - 0.5f;
    float v;
    if ((X2 >= 0.f) & (X2 < (float)(w - 1)) & (Y2 >= 0.f) &
        (Y2 < (float)(h - 1))) {
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
}

// Images of a launch: a device table (TAB), or for two-image launches the
// kernel argument itself, so the per-target measurement needs no upload.  The
// two inline images are flat scalar fields selected one by one, and the
// kernels are instantiated per form: a runtime choice between the table and
// the argument makes the compiler copy the argument into scratch memory.
struct CcImgs {
  const CcImg* tab;
  const float *src0, *src1;
  float *dst0, *dst1;
  int pitch0, pitch1;
};
template <bool TAB>
__device__ __forceinline__ CcImg cc_img(const CcImgs& j, unsigned b) {
  if (TAB) return j.tab[b];
  return CcImg{b ? j.src1 : j.src0, b ? j.pitch1 : j.pitch0,
               b ? j.dst1 : j.dst0};
}

// Partial (sum, sum of squares) of the valid pixels of a band of rows,
// grid (kMomParts, nimg).  FP64 accumulation (the reference's K7/K8 FP32
// tree order is not reproduced).
template <bool TAB>
__global__ __launch_bounds__(kCcThreads) void k_cc_moments(const CcImgs imgs,
                                                           int w, int h,
                                                           double* part) {
  __shared__ double sh[16];
  const CcImg im = cc_img<TAB>(imgs, blockIdx.y);
  const int rows = cdiv(h, kMomParts);
  const int ya = min(h, (int)blockIdx.x * rows), yb = min(h, ya + rows);
  double s1 = 0.0, s2 = 0.0;
  // Batches of kMomBatch pixels whose loads are issued before any is used
  // (a load per loop step left one memory latency per pixel); the sums take
  // the pixels in the same order.
  const int nr = yb - ya;
  for (LineIter it(w); it.l < nr;) {
    float v[kMomBatch];
#pragma unroll
    for (int j = 0; j < kMomBatch; j++) {
      v[j] = it.l < nr ? im.src[(size_t)(ya + it.l) * im.pitch + it.k] : -1e30f;
      it.next();
    }
#pragma unroll
    for (int j = 0; j < kMomBatch; j++) {
      const bool ok = v[j] >= -1e10f;
      s1 += ok ? (double)v[j] : 0.0;
      s2 += ok ? (double)(v[j] * v[j]) : 0.0;
    }
  }
  s1 = block_reduce(s1, sh, [](double a, double b) { return a + b; }, 0.0);
  s2 = block_reduce(s2, sh, [](double a, double b) { return a + b; }, 0.0);
  if (threadIdx.x == 0) {
    double* o = part + ((size_t)blockIdx.y * kMomParts + blockIdx.x) * 2;
    o[0] = s1;
    o[1] = s2;
  }
}

// Mean and standard deviation of image b from its kMomParts partials, in a
// fixed order (lane l sums partials l, l+64, ...; then a shuffle tree).  Call
// from one whole wave; the result is valid in lane 0.
__device__ void cc_mean_sd(const double* part, unsigned b, int w, int h,
                           int unit_std, float* mean, float* sd) {
  const int lane = threadIdx.x & 63;
  const double* p = part + (size_t)b * kMomParts * 2;
  double s1 = 0.0, s2 = 0.0;
  for (int k = lane; k < kMomParts; k += 64) {
    s1 += p[2 * k];
    s2 += p[2 * k + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    s1 += __shfl_down(s1, o, 64);
    s2 += __shfl_down(s2, o, 64);
  }
  const double npx = (double)w * (double)h;
  *mean = (float)(s1 / npx);
  *sd = 1.0f;
  if (!unit_std) {
    const float m2 = (float)(s2 / npx);
    const float var = m2 - *mean * *mean;
    *sd = var <= 0.f ? 0.0f : sqrtf(var);
  }
}

// mNormalize / PANorm (GNormalize, CalcCC.cu:137-147; ProjAlign.cpp:295-312)
// fused with GRoundEdge (mask table) and the row pass of the 2-D R2C: nl rows
// of image blockIdx.y per workgroup are normalised (invalid pixels and masked
// ones become 0), transformed and unpacked into the padded work image.
template <bool TAB>
__global__ __launch_bounds__(kCcThreads) void k_cc_rows_fwd(
    const CcImgs imgs, int w, int h, const double* part, const float* mask,
    int unit_std, FftDevPlan P, int nl) {
  extern __shared__ float2 sm2[];
  __shared__ float sh_ms[2];
  const int N = P.n, ls = fft_lds_len(P);
  float2* a = sm2;
  float2* b = a + nl * ls;
  float2* tw = b + nl * ls;
  const CcImg im = cc_img<TAB>(imgs, blockIdx.y);
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, h - r0);
  // The first kRowPre entries of the thread (pixel and mask pairs) are
  // loaded before the tables and the moments, so their memory round trip
  // overlaps those.
  constexpr int kRowPre = 8;
  LineIter it(N);
  float pv[kRowPre][2], pm[kRowPre][2];
  int pat[kRowPre];
#pragma unroll
  for (int j = 0; j < kRowPre; j++) {
    const int l = it.l, k = it.k;
    pat[j] = l < nl ? l * ls + k : -1;
    pv[j][0] = pv[j][1] = -1e30f;
    pm[j][0] = pm[j][1] = 0.f;
    if (l < nv) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        pv[j][q] = im.src[(size_t)(r0 + l) * im.pitch + 2 * k + q];
        pm[j][q] = mask[(size_t)(r0 + l) * w + 2 * k + q];
      }
    }
    it.next();
  }
  fft_load_tw(P, tw);
  if (threadIdx.x < 64) {
    float mean, sd;
    cc_mean_sd(part, blockIdx.y, w, h, unit_std, &mean, &sd);
    if (threadIdx.x == 0) {
      sh_ms[0] = mean;
      sh_ms[1] = sd;
    }
  }
  __syncthreads();
  const float mean = sh_ms[0], sd = sh_ms[1];
  auto normv = [&](float v, float mk) {
    float o = 0.0f;
    if (v >= -1e10f) {
      const float nv_ = (v - mean) / sd;
      if (!(nv_ < -1e10f) && mk != 0.0f) o = nv_ * mk;
    }
    return o;
  };
  auto norm = [&](int y, int x) {
    return normv(im.src[(size_t)y * im.pitch + x], mask[(size_t)y * w + x]);
  };
#pragma unroll
  for (int j = 0; j < kRowPre; j++)
    if (pat[j] >= 0)
      a[pat[j]] = make_float2(normv(pv[j][0], pm[j][0]), normv(pv[j][1], pm[j][1]));
  // The rest: four of a thread's entries per batch, their loads issued
  // together.
  for (; it.l < nl;) {
    float2 v[4];
    int at[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int l = it.l, k = it.k;
      at[j] = l < nl ? l * ls + k : -1;
      v[j] = make_float2(0.f, 0.f);
      if (l < nv) v[j] = make_float2(norm(r0 + l, 2 * k), norm(r0 + l, 2 * k + 1));
      it.next();
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (at[j] >= 0) a[at[j]] = v[j];
  }
  __syncthreads();
  float2* Z = fft_lines(a, b, P, nl, false, tw);
  const int cp = w / 2 + 1;  // complex row pitch of the work image
  for (LineIter it(N + 1); it.l < nv; it.next()) {
    const int l = it.l, k = it.k;
    float2* out = reinterpret_cast<float2*>(im.dst) + (size_t)(r0 + l) * cp;
    r2c_unpack(Z + l * ls, N, P.half, 1.0f, k,
               [&](int kk, float2 v) { out[kk] = v; });
  }
}

// GConv + GWiener + GCenterOrigin (CalcCC.cu:95-135) for spectral element
// (k, y): conj(r) * g, times exp(-2B (k^2+yy^2)/(W^2+H^2)) / sqrt(|.| + 0.01),
// times (-1)^(k+y).
__device__ __forceinline__ float2 cc_spectral(float2 r, float2 g, int k, int y,
                                              int nc, int h, float bfactor) {
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
  return make_float2(re, im);
}

// Column pass of both forward transforms of a pair (times 'scale'), the
// spectral product, and the column pass of the inverse transform, for nl
// adjacent spectral columns per workgroup; the product's columns return to
// the pair's second work image.
template <bool TAB>
__global__ __launch_bounds__(kCcThreads) void k_cc_cols(
    const CcImgs imgs, int w, int h, int ngroups, FftDevPlan P, int nl,
    float scale, float bfactor) {
  extern __shared__ float2 sm2[];
  const int ls = fft_lds_len(P);
  float2* x = sm2;
  float2* y = x + 2 * nl * ls;
  float2* tw = y + 2 * nl * ls;
  const int nc = w / 2 + 1;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int pair = blk / ngroups;
  const int c0 = (blk - pair * ngroups) * nl;
  const int nv = min(nl, nc - c0);
  const CcImg ia = cc_img<TAB>(imgs, 2 * pair), ib = cc_img<TAB>(imgs, 2 * pair + 1);
  const float2* A = reinterpret_cast<const float2*>(ia.dst);
  float2* B = reinterpret_cast<float2*>(ib.dst);
  fft_load_tw(P, tw);
  for (LineIter it(nl); it.l < 2 * h; it.next()) {
    const int r = it.l, l = it.k;  // r = image * h + row
    const int img = r >= h, row = r - img * h;
    float2 v = make_float2(0.f, 0.f);
    if (l < nv) v = (img ? B : A)[(size_t)row * nc + c0 + l];
    x[(img * nl + l) * ls + row] = v;
  }
  __syncthreads();
  float2* z = fft_lines(x, y, P, 2 * nl, false, tw);
  float2* o = z == x ? y : x;
  for (LineIter it(h); it.l < nl; it.next()) {
    const int l = it.l, row = it.k;
    const float2 r = cscale(z[l * ls + row], scale);
    const float2 g = cscale(z[(nl + l) * ls + row], scale);
    z[(nl + l) * ls + row] = cc_spectral(r, g, c0 + l, row, nc, h, bfactor);
  }
  __syncthreads();
  float2* q = fft_lines(z + nl * ls, o + nl * ls, P, nl, true, tw);
  for (LineIter it(nl); it.l < h; it.next()) {
    const int row = it.l, l = it.k;
    if (l < nv) B[(size_t)row * nc + c0 + l] = q[l * ls + row];
  }
}

// k_cc_cols for one spectral column per workgroup, 256 threads and
// power-of-two columns of N points, with the transforms in place: LDS holds
// the pair's two columns and the twiddles (3N float2, 24 KB at N = 1024,
// against 5N), so twice as many workgroups are resident per CU -- all of a
// two-pair sweep step at 1024^2 (1026) in one round.  Same arithmetic.
template <bool TAB, int N>
__global__ __launch_bounds__(kCcThreads) void k_cc_cols_inpl(
    const CcImgs imgs, int w, int h, int ngroups, FftDevPlan P, float scale,
    float bfactor) {
  extern __shared__ float2 sm2[];
  float2* x = sm2;          // [2][N]: the pair's columns
  float2* tw = x + 2 * N;
  const int nc = w / 2 + 1;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int pair = blk / ngroups;
  const int c0 = blk - pair * ngroups;
  const CcImg ia = cc_img<TAB>(imgs, 2 * pair), ib = cc_img<TAB>(imgs, 2 * pair + 1);
  const float2* A = reinterpret_cast<const float2*>(ia.dst);
  float2* B = reinterpret_cast<float2*>(ib.dst);
  // The column entries of both images (2N / kCcThreads per thread), all
  // loads issued before the first LDS write, then the twiddles.
  {
    constexpr int kPer = 2 * N / kCcThreads;
    float2 v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int r = (int)threadIdx.x + j * kCcThreads;
      const float2* src = j < kPer / 2 ? A : B;  // r < N: image A
      v[j] = src[(size_t)(r - (j < kPer / 2 ? 0 : N)) * nc + c0];
    }
    fft_load_tw(P, tw);
#pragma unroll
    for (int j = 0; j < kPer; j++) x[threadIdx.x + j * kCcThreads] = v[j];
  }
  __syncthreads();
  fft_pow2_inplace<N, false, 2, kCcThreads, N>(x, tw);
  for (int row = threadIdx.x; row < N; row += kCcThreads) {
    const float2 r = cscale(x[row], scale);
    const float2 g = cscale(x[N + row], scale);
    x[N + row] = cc_spectral(r, g, c0, row, nc, h, bfactor);
  }
  __syncthreads();
  fft_pow2_inplace<N, true, 1, kCcThreads, N>(x + N, tw);
  for (int row = threadIdx.x; row < N; row += kCcThreads)
    B[(size_t)row * nc + c0] = x[N + row];
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

// Argmax with the smallest row-major index winning ties (the reference's
// first maximum), over a whole workgroup; valid in thread 0.
__device__ Best block_best(Best b, Best* sh) {
  for (int o = 32; o > 0; o >>= 1) {
    Best c{__shfl_down(b.v, o, 64), __shfl_down(b.i, o, 64)};
    b = best_of(b, c);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = b;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) b = best_of(b, sh[k]);
  return b;
}

// The reference's parabolic sub-pixel fit around the integer peak b of a
// w x h map with row pitch 'pitch' (CalcCC::FindPeak, CalcCC.cu:406-435: the
// peak clamped to [1, n-2], FP32 terms; PAFindPeak, ProjAlign.cpp:349-370:
// FP64 terms); returns the fitted peak's offset from the map centre.  When
// nothing in the window beat the initial value, mode 0 keeps the reference's
// (0, 0) (clamped to (1, 1)) and mode 1 takes the window's corner, where the
// reference would read outside the map.
__device__ float2 peak_fit(const float* cc, int pitch, int w, int h, int mode,
                           const PeakWin& win, Best b) {
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
  return make_float2(fpx - (float)(w / 2), fpy - (float)(h / 2));
}

// Standalone peak search of one map (the entry point for maps made
// elsewhere): argmax partials of nl rows of the window per workgroup, each
// thread visiting its pixels in increasing row-major order.
__global__ __launch_bounds__(256) void k_peak_rows(const float* map, int w,
                                                   int h, int pitch, int mode,
                                                   int nl, Best* part) {
  __shared__ Best sh[16];
  const PeakWin win(w, h, mode);
  const int y0 = win.y0 + blockIdx.x * nl;
  const int y1 = min(y0 + nl, win.y1), wx = win.x1 - win.x0;
  Best best{win.init, 0x7fffffff};
  const long long n = (long long)max(y1 - y0, 0) * wx;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const int y = y0 + (int)(i / wx), x = win.x0 + (int)(i % wx);
    const float v = map[(size_t)y * pitch + x];
    if (v > best.v) best = Best{v, y * w + x};
  }
  best = block_best(best, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = best;
}

__global__ __launch_bounds__(256) void k_peak_fit(const float* map, int w,
                                                  int h, int pitch, int mode,
                                                  const Best* part, int nparts,
                                                  float* out) {
  __shared__ Best sh[16];
  const PeakWin win(w, h, mode);
  Best b{win.init, 0x7fffffff};
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) b = best_of(b, part[k]);
  b = block_best(b, sh);
  if (threadIdx.x != 0) return;
  const float2 r = peak_fit(map, pitch, w, h, mode, win, b);
  out[0] = r.x;
  out[1] = r.y;
}

// Row pass of the inverse 2-D transform (cufftExecC2R, unnormalised) of nl
// rows of the pair's correlation map per workgroup, written back in place,
// with the argmax of these rows inside the peak window as one partial.
template <bool TAB>
__global__ __launch_bounds__(kCcThreads) void k_cc_rows_inv(
    const CcImgs imgs, int w, int h, FftDevPlan P, int nl, int mode,
    Best* part) {
  extern __shared__ float2 sm2[];
  __shared__ Best sh[16];
  const int N = P.n, ls = fft_lds_len(P);
  float2* a = sm2;
  float2* b = a + nl * ls;  // nl * (ls + 1) entries
  float2* tw = b + nl * (ls + 1);
  const CcImg im = cc_img<TAB>(imgs, 2 * blockIdx.y + 1);
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, h - r0);
  const int bs = ls + 1, cp = w / 2 + 1;
  float2* map = reinterpret_cast<float2*>(im.dst);
  // The thread's first kRowPre spectral entries are loaded before the
  // twiddles, so the two round trips overlap.
  constexpr int kRowPre = 8;
  LineIter it(N + 1);
  {
    float2 v[kRowPre];
    int at[kRowPre];
#pragma unroll
    for (int j = 0; j < kRowPre; j++) {
      const int l = it.l, k = it.k;
      at[j] = l < nl ? l * bs + k : -1;
      v[j] = make_float2(0.f, 0.f);
      if (l < nv) v[j] = map[(size_t)(r0 + l) * cp + k];
      it.next();
    }
    fft_load_tw(P, tw);
#pragma unroll
    for (int j = 0; j < kRowPre; j++)
      if (at[j] >= 0) b[at[j]] = v[j];
  }
  // The rest: four of a thread's entries per batch, their loads issued
  // together.
  for (; it.l < nl;) {
    float2 v[4];
    int at[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int l = it.l, k = it.k;
      at[j] = l < nl ? l * bs + k : -1;
      v[j] = make_float2(0.f, 0.f);
      if (l < nv) v[j] = map[(size_t)(r0 + l) * cp + k];
      it.next();
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (at[j] >= 0) b[at[j]] = v[j];
  }
  __syncthreads();
  for (LineIter it(N); it.l < nl; it.next()) {
    const int l = it.l, k = it.k;
    a[l * ls + k] = c2r_pack(b[l * bs + k], b[l * bs + N - k], P.half[k]);
  }
  __syncthreads();
  float2* z = fft_lines(a, b, P, nl, true, tw);
  const PeakWin win(w, h, mode);
  Best best{win.init, 0x7fffffff};
  for (LineIter it(N); it.l < nv; it.next()) {
    const int l = it.l, k = it.k;
    const float2 v = z[l * ls + k];
    map[(size_t)(r0 + l) * cp + k] = v;
    const int y = r0 + l;
    // A thread visits its pixels in increasing row-major order, so a
    // strict '>' keeps its first maximum; block_best breaks ties by index.
    if (y >= win.y0 && y < win.y1) {
      const int xa = 2 * k, xb = 2 * k + 1;
      if (xa >= win.x0 && xa < win.x1 && v.x > best.v)
        best = Best{v.x, y * w + xa};
      if (xb >= win.x0 && xb < win.x1 && v.y > best.v)
        best = Best{v.y, y * w + xb};
    }
  }
  best = block_best(best, sh);
  if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = best;
}

// Combines a map's row-group partials, then the parabolic sub-pixel fit of
// the reference (FP32 terms in mode 0, FP64 in mode 1); one workgroup per
// map.  Writes the shift of the peak from the map centre.
template <bool TAB>
__global__ __launch_bounds__(kCcThreads) void k_cc_fit(const CcImgs imgs,
                                                       int w, int h, int mode,
                                                       const Best* part,
                                                       int nparts, float* out,
                                                       const PaUpdate up) {
  __shared__ Best sh[16];
  const int pitch = pad_x(w);
  const float* cc = cc_img<TAB>(imgs, 2 * blockIdx.x + 1).dst;
  const PeakWin win(w, h, mode);
  Best b{win.init, 0x7fffffff};
  for (int k = threadIdx.x; k < nparts; k += blockDim.x)
    b = best_of(b, part[(size_t)blockIdx.x * nparts + k]);
  b = block_best(b, sh);
  if (threadIdx.x != 0) return;
  const float2 pk = peak_fit(cc, pitch, w, h, mode, win, b);
  const float ishx = pk.x, ishy = pk.y;
  out[2 * blockIdx.x] = ishx;
  out[2 * blockIdx.x + 1] = ishy;
  if (up.sx) {
    const int t = up.t[blockIdx.x];
    const float a0 = ishx * up.binx, a1 = ishy * up.biny;
    up.meas[2 * t] = a0;
    up.meas[2 * t + 1] = a1;
    up.err[t] = (float)sqrt((double)(a0 * a0 + a1 * a1));
    const float c = up.rcs[2 * t], sn = up.rcs[2 * t + 1];
    float r0 = a0 * c - a1 * sn;
    float r1 = a0 * sn + a1 * c;
    r0 += up.ind[2 * t];
    r1 += up.ind[2 * t + 1];
    up.sx[t] += (double)r0;
    up.sy[t] += (double)r1;
    up.cp[t].shift_x = (float)up.sx[t];
    up.cp[t].shift_y = (float)up.sy[t];
  }
}

void set_lds(mf_ctx* ctx, const void* k, size_t bytes) {
  MF_REQUIRE(bytes <= 160 * 1024, "correlation lines exceed LDS");
  ensure_dyn_lds(ctx, k, bytes);
}

// Launch shape of one line-transform kernel: lines per workgroup and
// workgroup size.  Tunables "<family>/<w>x<h>x<npair>/lines" and
// ".../threads" override the heuristic: one line per workgroup (measured
// fastest for these latency-bound transforms, for one pair and for a whole
// sweep of pairs alike), at most what 96 KB of LDS holds (per_line bytes
// each, 'fixed' once); one thread per radix-8 butterfly of those lines, 128
// to kCcThreads.
struct Shape {
  int nl, threads;
};
Shape cc_shape(mf_ctx* ctx, const char* family, int w, int h, int npair,
               size_t per_line, size_t fixed, int n) {
  const std::string key = std::string(family) + "/" + std::to_string(w) + "x" +
                          std::to_string(h) + "x" + std::to_string(npair) + "/";
  const int fit =
      std::max(1, (int)(((size_t)96 * 1024 - std::min(fixed, (size_t)64 * 1024)) /
                        per_line));
  Shape s;
  s.nl = std::max(1, std::min(tune_get(ctx, key + "lines", 1), fit));
  const int t = (s.nl * std::max(n / 8, 1) + 63) / 64 * 64;
  s.threads = std::max(128, std::min(t, kCcThreads));
  s.threads = tune_get(ctx, key + "threads", s.threads);
  MF_REQUIRE(s.threads >= 64 && s.threads % 64 == 0 && s.threads <= kCcThreads,
             "workgroup size must be a multiple of 64 up to 256");
  return s;
}

template <bool TAB>
void cc_launch(mf_ctx* ctx, hipStream_t st, const CcImgs& imgs, int npair,
               int w, int h, int unit_std, float scale, float bfactor,
               int mode, const float* mask, double* part, float* d_shift,
               const PaUpdate& up);

// The correlation of npair pairs (images 2p, 2p+1 of 'imgs'): moments,
// normalisation, forward transforms, spectral product, inverse transform,
// peak search and fit.  'part' holds cc_scratch_doubles(npair, h) doubles.
void cc_run(mf_ctx* ctx, hipStream_t st, const CcImgs& imgs, int npair, int w,
            int h, int unit_std, float scale, float bfactor, int mode,
            const float* mask, double* part, float* d_shift,
            const PaUpdate* up = nullptr) {
  MF_REQUIRE(w % 2 == 0 && w >= 2 && h >= 1, "bad correlation shape");
  MF_REQUIRE(!up || npair <= 2, "a parameter update covers at most two maps");
  const PaUpdate none{};
  if (imgs.tab)
    cc_launch<true>(ctx, st, imgs, npair, w, h, unit_std, scale, bfactor,
                    mode, mask, part, d_shift, up ? *up : none);
  else
    cc_launch<false>(ctx, st, imgs, npair, w, h, unit_std, scale, bfactor,
                     mode, mask, part, d_shift, up ? *up : none);
}

template <bool TAB>
void cc_launch(mf_ctx* ctx, hipStream_t st, const CcImgs& imgs, int npair,
               int w, int h, int unit_std, float scale, float bfactor,
               int mode, const float* mask, double* part, float* d_shift,
               const PaUpdate& up) {
  const int nimg = 2 * npair;
  TimedLaunch tl(ctx->timer, "cc", st);
  hipLaunchKernelGGL(k_cc_moments<TAB>, dim3(kMomParts, nimg), dim3(kCcThreads), 0,
                     st, imgs, w, h, part);
  MF_HIP(hipGetLastError());
  const FftDevPlan& PR = ctx->fft->plan(w / 2);
  const size_t lr = fft_lds_len(PR);
  {
    const Shape sh = cc_shape(ctx, "cc_rowsf", w, h, npair, 2 * lr * 8,
                              lr * 8, (int)lr);
    const size_t lds = (2 * sh.nl * lr + lr) * sizeof(float2);
    set_lds(ctx, (const void*)k_cc_rows_fwd<TAB>, lds);
    hipLaunchKernelGGL(k_cc_rows_fwd<TAB>, dim3(cdiv(h, sh.nl), nimg),
                       dim3(sh.threads), lds, st, imgs, w, h,
                       (const double*)part, mask, unit_std, PR, sh.nl);
    MF_HIP(hipGetLastError());
  }
  {
    const FftDevPlan& PC = ctx->fft->plan(h);
    const size_t lc = fft_lds_len(PC);
    const int nc = w / 2 + 1;
    // Two lines (the pair's two columns) per column of the group.
    const Shape sh = cc_shape(ctx, "cc_cols", w, h, npair, 4 * lc * 8, lc * 8,
                              2 * (int)lc);
    // One column per workgroup of 256 threads over power-of-two columns:
    // the in-place form.
    decltype(&k_cc_cols_inpl<TAB, 1024>) inpl = nullptr;
    if (!PC.blue && sh.nl == 1 && sh.threads == kCcThreads) {
      if (h == 512) inpl = k_cc_cols_inpl<TAB, 512>;
      if (h == 1024) inpl = k_cc_cols_inpl<TAB, 1024>;
      if (h == 2048) inpl = k_cc_cols_inpl<TAB, 2048>;
    }
    if (inpl) {
      const size_t lds = (size_t)3 * h * sizeof(float2);
      set_lds(ctx, (const void*)inpl, lds);
      hipLaunchKernelGGL(inpl, dim3(nc * npair), dim3(kCcThreads), lds, st,
                         imgs, w, h, nc, PC, scale, bfactor);
      MF_HIP(hipGetLastError());
    } else {
      const size_t lds = (4 * sh.nl * lc + lc) * sizeof(float2);
      set_lds(ctx, (const void*)k_cc_cols<TAB>, lds);
      const int ngroups = cdiv(nc, sh.nl);
      hipLaunchKernelGGL(k_cc_cols<TAB>, dim3(ngroups * npair),
                         dim3(sh.threads), lds, st, imgs, w, h, ngroups, PC,
                         sh.nl, scale, bfactor);
      MF_HIP(hipGetLastError());
    }
  }
  Best* peaks = (Best*)(part + (size_t)2 * nimg * kMomParts);
  const Shape sh = cc_shape(ctx, "cc_rowsi", w, h, npair, (2 * lr + 1) * 8,
                            lr * 8, (int)lr);
  const size_t lds = ((2 * lr + 1) * sh.nl + lr) * sizeof(float2);
  set_lds(ctx, (const void*)k_cc_rows_inv<TAB>, lds);
  const int nparts = cdiv(h, sh.nl);
  hipLaunchKernelGGL(k_cc_rows_inv<TAB>, dim3(nparts, npair), dim3(sh.threads),
                     lds, st, imgs, w, h, PR, sh.nl, mode, peaks);
  hipLaunchKernelGGL(k_cc_fit<TAB>, dim3(npair), dim3(kCcThreads), 0, st, imgs, w, h,
                     mode, (const Best*)peaks, nparts, d_shift, up);
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
  float* A = d_work;                    // normalised references, spectra
  float* B = d_work + npair * img;      // normalised stretched images, maps
  float* S = d_work + 2 * npair * img;  // stretched images
  std::vector<CcImg> im(2 * npair);
  for (int p = 0; p < npair; p++) {
    im[2 * p] = CcImg{d_stack + (size_t)pairs[p].ref * w * h, w, A + p * img};
    im[2 * p + 1] = CcImg{S + p * img, pitch, B + p * img};
  }
  const size_t b0 = npair * sizeof(PairDesc), b1 = im.size() * sizeof(CcImg);
  char* dt = (char*)tab.ensure(ctx, b0 + b1, st);
  MF_HIP(hipMemcpyAsync(dt, pairs.data(), b0, hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(dt + b0, im.data(), b1, hipMemcpyHostToDevice, st));
  const PairDesc* d_pairs = (const PairDesc*)dt;
  const dim3 g2(std::max(1, std::min(h, 256)), npair);
  hipLaunchKernelGGL(k_stretch_pairs, g2, dim3(256), 0, st, d_stack, w, h,
                     d_pairs, S, 1);
  MF_HIP(hipGetLastError());
  CcImgs imgs{};
  imgs.tab = (const CcImg*)(dt + b0);
  // Transform's correlation normalises the forward FFT by 1/(w h)
  // (cufft2D::ForwardFFT with bNorm, CalcCC.cu:257,261).
  cc_run(ctx, st, imgs, npair, w, h, 0, (float)(1.0 / w / h), bfactor, 0,
         cc_mask(ctx, w, h, false), d_mom, d_shift);
  MF_HIP(hipStreamSynchronize(st));  // 'tab' is reusable after return
}

size_t cc_scratch_doubles(int npair, int h) {
  // Moment partials: 2 images x kMomParts x (sum, sum of squares) per pair;
  // peak partials: at most one (value, index) per map row.
  return (size_t)npair * (4 * kMomParts + h);
}

void pa_measure(mf_ctx* ctx, hipStream_t st, const float* d_reproj,
                const float* d_proj, int w, int h, float* d_work,
                double* d_mom, float* d_ishift, const PaUpdate* up) {
  const size_t img = (size_t)h * pad_x(w);
  CcImgs imgs{};
  imgs.src0 = d_reproj;
  imgs.pitch0 = w;
  imgs.dst0 = d_work;
  imgs.src1 = d_proj;
  imgs.pitch1 = w;
  imgs.dst1 = d_work + img;
  // PANorm: std := 1 (ProjAlign.cpp:295-312); unnormalised transforms and
  // B = 500 (:258-271); PAFindPeak window (:325-375).
  cc_run(ctx, st, imgs, 1, w, h, 1, 1.0f, 500.0f, 1, cc_mask(ctx, w, h, true),
         d_mom, d_ishift, up);
}

void pa_measure_tab(mf_ctx* ctx, hipStream_t st, const CcImg* d_tab,
                    int npair, int w, int h, double* d_mom, float* d_ishift,
                    const PaUpdate* up) {
  MF_REQUIRE(npair > 0, "no pairs to measure");
  CcImgs imgs{};
  imgs.tab = d_tab;
  // The same normalisation, filter and peak window as pa_measure.
  cc_run(ctx, st, imgs, npair, w, h, 1, 1.0f, 500.0f, 1,
         cc_mask(ctx, w, h, true), d_mom, d_ishift, up);
}

namespace {

void check_peak_args(int w, int h, int pitch, int mode) {
  MF_REQUIRE(w >= 4 && h >= 4 && pitch >= w && (mode == 0 || mode == 1),
             "bad peak-search arguments");
  if (mode == 1)
    MF_REQUIRE(w > 2 * std::max(w / 20, 3) && h > 2 * std::max(h / 20, 3),
               "map too small for the projection-matching peak window");
}

constexpr int kPeakRowsPerPart = 8;

// The two launches of a standalone peak search: cdiv(h, 8) partials in
// 'part', the shift in d_out.
void launch_peak(hipStream_t st, const float* d_map, int w, int h, int pitch,
                 int mode, Best* part, float* d_out) {
  const int nparts = cdiv(h, kPeakRowsPerPart);
  hipLaunchKernelGGL(k_peak_rows, dim3(nparts), dim3(256), 0, st, d_map, w, h,
                     pitch, mode, kPeakRowsPerPart, part);
  hipLaunchKernelGGL(k_peak_fit, dim3(1), dim3(256), 0, st, d_map, w, h, pitch,
                     mode, part, nparts, d_out);
  MF_HIP(hipGetLastError());
}

// Scratch of the allocation-free peak search (mf_peak2d_noalloc): the
// partials of maps of up to kPeakNoAllocMaxRows rows, then the float2
// result.  One copy per GPU, held until the process ends; every call on a
// GPU shares it, so calls that overlap on one GPU race.  Each call writes
// every partial and the result it reads, so nothing needs clearing.
constexpr int kPeakNoAllocMaxRows = 32768;
constexpr int kPeakNoAllocParts = kPeakNoAllocMaxRows / kPeakRowsPerPart;
constexpr size_t kPeakResultOffset = kPeakNoAllocParts * sizeof(Best);
static __device__ __align__(256) unsigned char
    scratch[kPeakResultOffset + 2 * sizeof(float)];

}  // namespace

void peak2d(mf_ctx* ctx, hipStream_t st, const float* d_map, int w, int h,
            int pitch, int mode, DevBuf& ws, float* d_shift) {
  check_peak_args(w, h, pitch, mode);
  Best* part = (Best*)ws.ensure(ctx, cdiv(h, kPeakRowsPerPart) * sizeof(Best),
                                st);
  launch_peak(st, d_map, w, h, pitch, mode, part, d_shift);
}

void* peak_scratch_address(int device) {
  // The array's device address, looked up once per device with that device
  // current (each GPU holds its own copy at its own address).
  static std::mutex mu;
  static std::map<int, void*> addr;
  std::lock_guard<std::mutex> g(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  int cur = 0;
  MF_HIP(hipGetDevice(&cur));
  MF_REQUIRE(cur == device, "peak scratch looked up with another device current");
  void* p = nullptr;
  MF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  MF_REQUIRE(p != nullptr, "no device address for the peak scratch");
  addr[device] = p;
  return p;
}

void peak2d_noalloc(mf_ctx* ctx, hipStream_t st, const float* d_map, int w,
                    int h, int pitch, int mode, float* h_shift) {
  check_peak_args(w, h, pitch, mode);
  MF_REQUIRE(h <= kPeakNoAllocMaxRows,
             "map taller than the allocation-free peak search supports "
             "(32768 rows)");
  MF_REQUIRE(ctx->peak_scratch != nullptr, "context has no peak scratch");
  char* base = (char*)ctx->peak_scratch;
  float* d_out = (float*)(base + kPeakResultOffset);
  launch_peak(st, d_map, w, h, pitch, mode, (Best*)base, d_out);
  MF_HIP(hipMemcpyAsync(h_shift, d_out, 2 * sizeof(float),
                        hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
}

}  // namespace mf
