// Common-line tilt-axis search (CalcTIltAxis, method/Rotate.cu) and the
// tilt-angle offset search (FindOffset, method/FindOffset.cu) on the GPU.
//
// The reference computes each image's Radon lines after a disk read, runs
// 5 API calls per line for its 1-D FFT, copies every spectrum plane to the
// host and then scores candidates with ~10 allocations/launches/copies per
// (candidate, image) -- about 1e5 runtime calls per search (Rotate.cu:292-561).
// Here one launch per step covers all images and candidates.
#include <cmath>
#include <string>
#include <vector>

#include "axis.h"
#include "cc.h"
#include "context.h"
#include "devutil.h"
#include "fft_dev.h"
#include "stages.h"

namespace mf {

AxisSetup AxisSetup::make(int nx, int ny, float axis0, float range, int num) {
  // CLsetup (Rotate.cu:339-374).
  AxisSetup a;
  a.num = num;
  const float step = range / (float)(num - 1);
  a.rot.resize(num);
  a.rot[0] = axis0 - step * (float)num / 2.0f;
  for (int i = 1; i < num; i++) a.rot[i] = a.rot[i - 1] + step;
  float fmin = 1e20f;
  for (int i = 0; i < num; i++) {
    float fs = (float)std::fabs(std::sin(kD2R * a.rot[i]));
    float fc = (float)std::fabs(std::cos(kD2R * a.rot[i]));
    float s1 = (float)nx / (fs + 0.000001f);
    float s2 = (float)ny / (fc + 0.000001f);
    if (fmin > s1) fmin = s1;
    if (fmin > s2) fmin = s2;
  }
  a.L = (int)fmin / 2 * 2 - 100;
  a.cs.resize(2 * num);
  for (int i = 0; i < num; i++) {
    // GRadon / GCalcCommonRegion angle: rot * 3.1415926f / 180.0f.
    float r = a.rot[i] * 3.1415926f / 180.0f;
    a.cs[2 * i] = hcos(r);
    a.cs[2 * i + 1] = hsin(r);
  }
  return a;
}

namespace {

__global__ __launch_bounds__(256) void k_common_region(int num, int L, int nx,
                                                       int ny, float ox,
                                                       float oy,
                                                       const float* cs,
                                                       int* reg) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (y >= L) return;
  const float c = cs[2 * k], s = cs[2 * k + 1];
  const float fy = (float)y - (float)L * 0.5f;
  const int half = L / 2;
  int left = 0, right = 0;  // entries never reached stay 0
  for (int x = 1; x < half; x++) {
    float X = (float)(-x) * c - fy * s + ox;
    float Y = (float)(-x) * s + fy * c + oy;
    if (X < 0.f || X >= (float)nx) break;
    if (Y < 0.f || Y >= (float)ny) break;
    left = -x;
  }
  for (int x = 1; x < half; x++) {
    float X = (float)x * c - fy * s + ox;
    float Y = (float)x * s + fy * c + oy;
    if (X < 0.f || X >= (float)nx) break;
    if (Y < 0.f || Y >= (float)ny) break;
    right = x;
  }
  reg[2 * ((size_t)k * L + y)] = left;
  reg[2 * ((size_t)k * L + y) + 1] = right;
}

constexpr int kRadonLines = 64;  // lines (lanes) per workgroup

struct RadonImg {
  float ct;      // (float)cos(tilt * D2R)
  float ox, oy;  // nx/2 + shift_x, ny/2 + shift_y
};

// GRadon (Rotate.cu:42-95): mean of the nearest non-negative pixels along
// the candidate direction over the cos(tilt)-cropped common region; the sum
// is sequential in x exactly as the reference thread sums it.  One-dimensional
// grid over (candidate, line block, image), image slowest and candidate
// fastest, so that blocks launched together trace nearly the same pixels.
// In the hardware's round-robin over the XCDs every XCD works on the same
// image at once; with 'remap' consecutive blocks (neighbouring candidates)
// share one XCD's L2 instead, which pays for images much larger than the L2
// and candidates a small fraction of a degree apart (5760x4092, 0.2 degree:
// 99 -> 76 ms) and costs elsewhere (1024^2, or 1.8 degree steps).  One wave
// per workgroup and eight samples in flight per lane; measured 1.4x faster
// at 1024^2 and 1.8x at 5760x4092 than 256-lane workgroups with four
// samples in flight.  The lanes of a wave trace
// parallel lines a pixel apart, stepping
// by (-sin, cos); for candidates nearer the image's x axis than its y axis
// that step crosses rows, so those read the transposed stack (tstack, [nz]
// [nx][ny]), where it runs along them.  Samples inside the analytic
// in-image interval of the line (with a one-pixel margin) skip the bounds
// test; the others take it.
// Pixel (ix, iy) of one image: row-major, or from the transposed copy.
// 24-bit products (both factors are below 2^24 and the index below 2^32) and
// a 32-bit byte offset from the image base keep the address arithmetic at
// full rate.
template <bool TRANS>
__device__ __forceinline__ float radon_pix(const float* img, int nx, int ny,
                                           int ix, int iy) {
  const unsigned e = TRANS ? __umul24((unsigned)ix, (unsigned)ny) + (unsigned)iy
                           : __umul24((unsigned)iy, (unsigned)nx) + (unsigned)ix;
  return *reinterpret_cast<const float*>(
      reinterpret_cast<const char*>(img) + (e << 2));
}

// The same pixel with its element index clamped to the image, for lanes
// whose coordinates may lie outside it (their value is discarded).
template <bool TRANS>
__device__ __forceinline__ float radon_pix_clamped(const float* img, int nx,
                                                   int ny, int ix, int iy,
                                                   unsigned last) {
  const unsigned e = TRANS ? __umul24((unsigned)ix, (unsigned)ny) + (unsigned)iy
                           : __umul24((unsigned)iy, (unsigned)nx) + (unsigned)ix;
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(img) +
                                         (min(e, last) << 2));
}

// Steps of misalignment between the lanes above which the interior loop is
// sheared.
constexpr int kRadonShearMin = 2;

template <bool TRANS, bool NEG>
__device__ __forceinline__ void radon_line(const float* img, int nx, int ny,
                                           float c, float s, float fy,
                                           const RadonImg& im, int left,
                                           int right, float* sum_out,
                                           int* cnt_out);

constexpr int kNegParts = 32;

// Flag of part blockIdx.x of image blockIdx.y: whether it holds a value
// below 0 (GRadon's skipped samples).  Parts are row bands.
__global__ __launch_bounds__(256) void k_image_negative(const float* stack,
                                                        int nx, int ny,
                                                        int* negf) {
  const size_t plane = (size_t)nx * ny;
  const int rows = cdiv(ny, kNegParts);
  const int ya = min(ny, (int)blockIdx.x * rows), yb = min(ny, ya + rows);
  const float* p = stack + blockIdx.y * plane + (size_t)ya * nx;
  const size_t n = (size_t)(yb - ya) * nx;
  int any = 0;
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) any |= p[i] < 0.f;
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) negf[blockIdx.y * kNegParts + blockIdx.x] = any;
}

__global__ __launch_bounds__(kRadonLines) void k_radon(const float* stack,
                                               const float* tstack, int nx,
                                               int ny, int num, int L,
                                               const float* cs, const int* reg,
                                               const RadonImg* imgs,
                                               float* lines, int pitch,
                                               const int* negf, int remap) {
  const int yblocks = cdiv(L, kRadonLines);
  const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const int z = blk / (yblocks * num);
  const int rem = blk - z * yblocks * num;
  const int yb = rem / num;
  const int k = rem - yb * num;
  // Lanes past the last line shadow it (the interior loop takes wave-wide
  // reductions) and store nothing.
  const int ylive = yb * blockDim.x + threadIdx.x;
  const int y = min(ylive, L - 1);
  const RadonImg im = imgs[z];
  const float c = cs[2 * k], s = cs[2 * k + 1];
  const bool trans = tstack && fabsf(c) > fabsf(s);
  const float* img = (trans ? tstack : stack) + (size_t)z * nx * ny;
  const float fy = (float)y - (float)L * 0.5f;
  const int* r = reg + 2 * ((size_t)k * L + y);
  const int left = (int)((float)r[0] * im.ct + 0.5f);
  const int right = (int)((float)r[1] * im.ct + 0.5f);
  int neg = 0;  // uniform over the workgroup (one image)
#pragma unroll
  for (int j = 0; j < kNegParts; j++) neg |= negf[z * kNegParts + j];
  float sum;
  int cnt;
  if (trans) {
    if (neg)
      radon_line<true, true>(img, nx, ny, c, s, fy, im, left, right, &sum, &cnt);
    else
      radon_line<true, false>(img, nx, ny, c, s, fy, im, left, right, &sum, &cnt);
  } else {
    if (neg)
      radon_line<false, true>(img, nx, ny, c, s, fy, im, left, right, &sum, &cnt);
    else
      radon_line<false, false>(img, nx, ny, c, s, fy, im, left, right, &sum, &cnt);
  }
  if (ylive < L)
    lines[((size_t)z * num + k) * pitch + y] = cnt == 0 ? kInvalid : sum / (float)cnt;
}

// The sequential sum of one line over x in [left, right) (left <= 0 <=
// right, as the reference's two loops [left, 0) and [0, right) run).  NEG:
// the image may hold negative values, which the sum skips; without them
// the in-image samples need no test at all.
template <bool TRANS, bool NEG>
__device__ __forceinline__ void radon_line(const float* img, int nx, int ny,
                                           float c, float s, float fy,
                                           const RadonImg& im, int left,
                                           int right, float* sum_out,
                                           int* cnt_out) {
  float sum = 0.0f;
  int cnt = 0;
  const float fys = fy * s, fyc = fy * c;
  const float fnx = (float)nx, fny = (float)ny;
  // In-image interval [xa, xb) from the exact linear forms with a one-pixel
  // margin (the FP32 evaluation errs by far less).
  double lo = -1e30, hi = 1e30;
  auto clip = [&](double a0, double a1, double mx) {
    // a0 + a1 x in [1, mx - 1]
    if (fabs(a1) < 1e-12) {
      if (a0 < 1.0 || a0 > mx - 1.0) {
        lo = 1e30;
        hi = -1e30;
      }
      return;
    }
    const double t1 = (1.0 - a0) / a1, t2 = (mx - 1.0 - a0) / a1;
    lo = fmax(lo, fmin(t1, t2));
    hi = fmin(hi, fmax(t1, t2));
  };
  clip((double)im.ox - (double)fys, c, nx);
  clip((double)im.oy + (double)fyc, s, ny);
  int xa = right, xb = right;
  if (lo <= hi) {
    xa = (int)fmax((double)left, fmin((double)right, ceil(lo)));
    xb = (int)fmax((double)xa, fmin((double)right, floor(hi) + 1.0));
  }
  // Branch-free so that the loads of several samples are in flight at once:
  // a skipped sample adds +0.0f to a sum that is never -0, which leaves it
  // unchanged.
  auto tested = [&](int from, int to) {
#pragma unroll 4
    for (int x = from; x < to; x++) {
      const float X = (float)x * c - fys + im.ox;
      const float Y = (float)x * s + fyc + im.oy;
      const bool in = X >= 0.f && X < fnx && Y >= 0.f && Y < fny;
      const float v =
          radon_pix<TRANS>(img, nx, ny, in ? (int)X : 0, in ? (int)Y : 0);
      const bool ok = in && !(v < 0.f);
      sum += ok ? v : 0.0f;
      cnt += ok ? 1 : 0;
    }
  };
  tested(left, xa);
  // The wave's lanes trace parallel lines, but their in-image intervals
  // start at different points along them, so at equal steps they can sit
  // many rows (columns in the transposed copy) apart, one cache line per
  // lane.  If so, each lane enters the loop o steps late, o chosen so that
  // all lanes step through one row together; a lane outside its own
  // interval reads a clamped address and adds nothing.
  const float slope = TRANS ? c : s;  // row (column) per step, |.| >= 0.7
  const bool has = xb > xa;
  const float fxa = (float)xa;
  const float ka = TRANS ? fxa * c - fys + im.ox : fxa * s + fyc + im.oy;
  float kref = has ? ka : (slope > 0.f ? 1e30f : -1e30f);
  for (int o = 32; o > 0; o >>= 1) {
    const float other = __shfl_xor(kref, o, 64);
    kref = slope > 0.f ? fminf(kref, other) : fmaxf(kref, other);
  }
  const int off = has ? max(0, (int)floorf((ka - kref) / slope + 0.5f)) : 0;
  const int n = has ? xb - xa : 0;
  int omax = off, span = off + n, nmax = n;
  for (int o = 32; o > 0; o >>= 1) {
    omax = max(omax, __shfl_xor(omax, o, 64));
    span = max(span, __shfl_xor(span, o, 64));
    nmax = max(nmax, __shfl_xor(nmax, o, 64));
  }
  // Sheared only while the idle steps stay under a fifth of the loop.
  if (omax > kRadonShearMin && 4 * span <= 5 * nmax) {
    const unsigned last = (unsigned)nx * (unsigned)ny - 1u;
    // x as a float counter: consecutive integers below 2^24 are exact.
    float fx = (float)(xa - off);
#pragma unroll 8
    for (int i = 0; i < span; i++, fx += 1.0f) {
      const bool act = (unsigned)(i - off) < (unsigned)n;
      const float X = fx * c - fys + im.ox;
      const float Y = fx * s + fyc + im.oy;
      const float v = radon_pix_clamped<TRANS>(img, nx, ny, (int)X, (int)Y, last);
      if (NEG) {
        const bool ok = act && !(v < 0.f);
        sum += ok ? v : 0.0f;
        cnt += ok ? 1 : 0;
      } else {
        sum += act ? v : 0.0f;
      }
    }
    if (!NEG) cnt += n;
    tested(xb, right);
    *sum_out = sum;
    *cnt_out = cnt;
    return;
  }
  // x as a float counter: consecutive integers below 2^24 are exact.
  float fx = (float)xa;
  if (NEG) {
#pragma unroll 8
    for (int x = xa; x < xb; x++, fx += 1.0f) {
      const float X = fx * c - fys + im.ox;
      const float Y = fx * s + fyc + im.oy;
      const float v = radon_pix<TRANS>(img, nx, ny, (int)X, (int)Y);
      const bool ok = !(v < 0.f);
      sum += ok ? v : 0.0f;
      cnt += ok ? 1 : 0;
    }
  } else {
#pragma unroll 8
    for (int x = xa; x < xb; x++, fx += 1.0f) {
      const float X = fx * c - fys + im.ox;
      const float Y = fx * s + fyc + im.oy;
      sum += radon_pix<TRANS>(img, nx, ny, (int)X, (int)Y);
    }
    cnt += xb - xa;
  }
  tested(xb, right);
  *sum_out = sum;
  *cnt_out = cnt;
}

// [nz][ny][nx] -> [nz][nx][ny] in 32 x 32 tiles staged in LDS.
__global__ __launch_bounds__(256) void k_transpose(const float* in, float* out,
                                                   int nx, int ny) {
  __shared__ float t[32][33];
  const size_t plane = (size_t)nx * ny;
  const float* src = in + blockIdx.z * plane;
  float* dst = out + blockIdx.z * plane;
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int x = x0 + tx, y = y0 + j;
    if (x < nx && y < ny) t[j][tx] = src[(size_t)y * nx + x];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int y = y0 + tx, x = x0 + j;
    if (x < nx && y < ny) dst[(size_t)x * ny + y] = t[tx][j];
  }
}

constexpr int kLineThreads = 256;

// GCalcMean + GRemoveMean + 1-D R2C (Rotate.cu:97-157, 430-451) for nl
// lines per workgroup, in place (pitch = L + 2 floats).
__global__ __launch_bounds__(kLineThreads) void k_line_fft(
    float* lines, int nlines, int L, int pitch, const float* taper,
    FftDevPlan P, int nl) {
  extern __shared__ float2 sm2[];
  __shared__ double ssum[16];
  __shared__ int scnt[16];
  __shared__ float smean[16];
  const int N = P.n, ls = fft_lds_len(P);
  float2* a = sm2;
  float2* b = a + nl * ls;
  float* raw = (float*)b;  // staging: nl * L floats (fits in nl * ls float2)
  float2* tw = b + nl * ls;
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, nlines - r0);
  fft_load_tw(P, tw);
  for (LineIter it(L); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    raw[l * L + k] = l < nv ? lines[(size_t)(r0 + l) * pitch + k] : 0.0f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int l = 0; l < nl; l++) {
    double s = 0.0;
    int c = 0;
    for (int k = threadIdx.x; k < L; k += blockDim.x) {
      float v = raw[l * L + k];
      if (v > -1e10f) {
        s += v;
        c++;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_down(s, o, 64);
      c += __shfl_down(c, o, 64);
    }
    if (lane == 0) {
      ssum[w] = s;
      scnt[w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double ts = 0.0;
      int tc = 0;
      for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
        ts += ssum[q];
        tc += scnt[q];
      }
      smean[l] = tc == 0 ? 0.0f : (float)(ts / tc);
    }
    __syncthreads();
  }
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    const float m = smean[l];
    float v0 = raw[l * L + 2 * k], v1 = raw[l * L + 2 * k + 1];
    v0 = (v0 < 0.f ? 0.0f : v0 - m) * taper[2 * k];
    v1 = (v1 < 0.f ? 0.0f : v1 - m) * taper[2 * k + 1];
    a[l * ls + k] = make_float2(v0, v1);
  }
  __syncthreads();
  float2* Z = fft_lines(a, b, P, nl, false, tw);
  for (LineIter it(N + 1); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    if (l >= nv) continue;
    float2* out = reinterpret_cast<float2*>(lines + (size_t)(r0 + l) * pitch);
    r2c_unpack(Z + l * ls, N, P.half, 1.0f, k,
               [&](int kk, float2 v) { out[kk] = v; });
  }
}

// Sum over images of each candidate's spectrum, sequential in z as the
// reference's GSumCmp accumulation (Rotate.cu:498-503).
__global__ void k_line_sum(const float2* spec, int nz, int num, int nc,
                           int cpitch, float2* S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num * nc) return;
  const int k = i / nc, j = i - k * nc;
  float2 acc = make_float2(0.f, 0.f);
  for (int z = 0; z < nz; z++) {
    const float2 v = spec[((size_t)z * num + k) * cpitch + j];
    acc = make_float2(acc.x + v.x, acc.y + v.y);
  }
  S[i] = acc;
}

// GConv/GSum (Rotate.cu:179-274) for one (candidate, image): filtered
// normalised correlation of the line with the sum of the others.
__global__ __launch_bounds__(256) void k_line_cc(const float2* spec,
                                                 const float2* S, int nz,
                                                 int num, int nc, int cpitch,
                                                 const float* filt,
                                                 double fsum, float* ccv) {
  __shared__ double sh[4];
  const int k = blockIdx.x, z = blockIdx.y;
  const float2* l = spec + ((size_t)z * num + k) * cpitch;
  const float2* s = S + (size_t)k * nc;
  double acc = 0.0;
  for (int j = 1 + threadIdx.x; j < nc; j += blockDim.x) {
    const float2 lv = l[j], sv = s[j];
    const float rx = sv.x - lv.x, ry = sv.y - lv.y;
    const float a1 = sqrtf(rx * rx + ry * ry);
    const float a2 = sqrtf(lv.x * lv.x + lv.y * lv.y);
    acc += (double)((lv.x * rx + lv.y * ry) * filt[j] / (a1 * a2 + 1e-20f));
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sh[0] + sh[1] + sh[2] + sh[3];
    ccv[(size_t)k * nz + z] = (float)(t / fsum);
  }
}

__global__ void k_line_score(const float* ccv, int nz, int num, float* score) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= num) return;
  float s = 0.0f;
  for (int z = 0; z < nz; z++) s += ccv[(size_t)k * nz + z];
  score[k] = s / (float)nz;
}

struct OffJob {
  const float* ref;
  const float* img;
  float m[3];
};

// FindOffset::Correlate (FindOffset.cu:176-238): stretch (bilinear, no
// fill) and the six masked moments over pixels valid in both images.
__global__ __launch_bounds__(256) void k_offset_moments(int w, int h,
                                                        const OffJob* jobs,
                                                        double* out) {
  __shared__ double sh[6][4];
  const OffJob jb = jobs[blockIdx.x];
  const float* ref = jb.ref;
  const float* img = jb.img;
  const float hw = 0.5f * (float)w, hh = 0.5f * (float)h;
  double m[6] = {0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < w * h; i += blockDim.x) {
    const int y = i / w, x = i - y * w;
    const float v1 = ref[i];
    if (v1 < -1e10f) continue;
    const float X = (float)x - hw + 0.5f;
    const float Y = (float)((double)y - 0.5 * (double)h + 0.5);
    const float buf = X * jb.m[0] + Y * jb.m[1];
    const float Y2 = X * jb.m[1] + Y * jb.m[2] + hh - 0.5f;
    const float X2 = buf + hw - 0.5f;
    if (!(X2 >= 0.f && X2 < (float)(w - 1) && Y2 >= 0.f && Y2 < (float)(h - 1)))
      continue;
    const int ix = (int)X2, iy = (int)Y2;
    const float fx = X2 - (float)ix, fy = Y2 - (float)iy;
    const float f2 = 1.0f - fx, f3 = 1.0f - fy;
    const float* p = img + (size_t)iy * w + ix;
    const float v2 = p[0] * f2 * f3 + p[1] * fx * f3 + p[w] * f2 * fy +
                     p[w + 1] * fx * fy;
    if (v2 < -1e10f) continue;
    m[0] += v1;
    m[1] += v2;
    m[2] += (double)(v1 * v1);
    m[3] += (double)(v2 * v2);
    m[4] += (double)(v1 * v2);
    m[5] += 1.0;
  }
  for (int q = 0; q < 6; q++) {
    double v = m[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6)
    out[(size_t)blockIdx.x * 6 + threadIdx.x] =
        sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] +
        sh[threadIdx.x][3];
}

}  // namespace

const float* stack_transposed(mf_ctx* ctx, hipStream_t st,
                              const float* d_stack, int nx, int ny, int nz,
                              AxisWork& ws) {
  float* t = (float*)ws.tstack.ensure(ctx, (size_t)nx * ny * nz * sizeof(float),
                                      st);
  if (ws.t_of != d_stack) {
    hipLaunchKernelGGL(k_transpose, dim3(cdiv(nx, 32), cdiv(ny, 32), nz),
                       dim3(256), 0, st, d_stack, t, nx, ny);
    MF_HIP(hipGetLastError());
    ws.t_of = d_stack;
  }
  return t;
}

float* common_lines(mf_ctx* ctx, hipStream_t st, const float* d_stack, int nx,
                    int ny, int nz, const float* tilts, const double* sx,
                    const double* sy, const AxisSetup& a, AxisWork& ws) {
  // GetFrameIdxFromTilt(0) (Util.cpp:8-20).
  int z0 = 0;
  {
    float m = std::fabs(tilts[0]);
    for (int i = 1; i < nz; i++) {
      float d = std::fabs(tilts[i]);
      if (d < m) {
        m = d;
        z0 = i;
      }
    }
  }
  const int num = a.num, L = a.L, nc = L / 2 + 1, pitch = L + 2;
  MF_REQUIRE(L >= 4, "common-line length too small for this image size");
  std::vector<RadonImg> ri(nz);
  for (int z = 0; z < nz; z++) {
    ri[z].ct = (float)std::cos(tilts[z] * kD2R);
    ri[z].ox = (float)nx * 0.5f + (float)sx[z];
    ri[z].oy = (float)ny * 0.5f + (float)sy[z];
  }
  // Workspace (the 'small' block also holds axis_search's score buffers).
  const size_t lines_f = (size_t)nz * num * pitch;
  float* lines = (float*)ws.lines.ensure(ctx, lines_f * sizeof(float), st);
  int* reg = (int*)ws.reg.ensure(ctx, (size_t)num * L * 2 * sizeof(int), st);
  char* small = (char*)ws.small.ensure(
      ctx, a.cs.size() * 4 + nz * sizeof(RadonImg) + (size_t)num * nc * 8 +
               (size_t)num * nz * 4 + num * 4 + 256, st);
  float* d_cs = (float*)small;
  RadonImg* d_ri = (RadonImg*)(small + a.cs.size() * 4);
  MF_HIP(hipMemcpyAsync(d_cs, a.cs.data(), a.cs.size() * 4,
                        hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(d_ri, ri.data(), nz * sizeof(RadonImg),
                        hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_common_region, dim3(cdiv(L, 256), num), dim3(256), 0, st,
                     num, L, nx, ny, (float)nx * 0.5f + (float)sx[z0],
                     (float)ny * 0.5f + (float)sy[z0], d_cs, reg);
  // Column-like candidates read a transposed copy of the stack, made once
  // per stack.
  bool colike = false;
  for (int k = 0; k < num; k++)
    colike = colike || std::fabs(a.cs[2 * k]) > std::fabs(a.cs[2 * k + 1]);
  const float* tstack =
      colike ? stack_transposed(ctx, st, d_stack, nx, ny, nz, ws) : nullptr;
  int* negf = (int*)ws.negf.ensure(ctx, (size_t)nz * kNegParts * sizeof(int),
                                   st);
  if (ws.neg_of != d_stack) {
    hipLaunchKernelGGL(k_image_negative, dim3(kNegParts, nz), dim3(256), 0, st,
                       d_stack, nx, ny, negf);
    MF_HIP(hipGetLastError());
    ws.neg_of = d_stack;
  }
  {
    TimedLaunch tl(ctx->timer, "radon", st);
    // Neighbouring candidates on one XCD for images well beyond its 4 MB L2
    // and candidate steps under a degree.
    const float step = num > 1 ? std::fabs(a.rot[1] - a.rot[0]) : 180.0f;
    const int remap = (size_t)nx * ny * 4 > ((size_t)32 << 20) && step < 1.0f;
    hipLaunchKernelGGL(k_radon, dim3(cdiv(L, kRadonLines) * num * nz),
                       dim3(kRadonLines), 0, st,
                       d_stack, tstack, nx, ny, num, L, d_cs, reg, d_ri, lines,
                       pitch, negf, remap);
  }
  MF_HIP(hipGetLastError());
  return lines;
}

AxisResult axis_search(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                       int nx, int ny, int nz, const float* tilts,
                       const double* sx, const double* sy, float axis0,
                       float range, int num, AxisWork& ws) {
  AxisSetup a = AxisSetup::make(nx, ny, axis0, range, num);
  const int L = a.L, nc = L / 2 + 1, pitch = L + 2;
  float* lines =
      common_lines(ctx, st, d_stack, nx, ny, nz, tilts, sx, sy, a, ws);
  // Host tables: taper (GRemoveMean) and filter (GConv) depend on L only.
  const float* d_taper =
      (const float*)ctx_table(ctx, "taper/" + std::to_string(L), [&] {
        std::vector<float> taper(L);
        const float half = 0.5f * (float)L;
        for (int i = 0; i < L; i++) {
          float r = std::fabs((float)i - half) / half;
          r = 0.5f * (1 - hcos((float)(3.14159 * (double)r)));
          r = 1.0f - (float)std::pow((double)r, 100.0);
          taper[i] = r * r;
        }
        return taper;
      });
  std::vector<float> filt(nc, 0.0f);
  double fsum = 0.0;
  for (int j = 1; j < nc; j++) {
    float f = (float)j / (2.0f * (float)(nc - 1));
    filt[j] = (float)std::exp((double)(-2.0f * 10.0f * f * f));
    fsum += filt[j];
  }
  const float* d_filt = (const float*)ctx_table(
      ctx, "linefilt/" + std::to_string(nc), [&] { return filt; });
  char* small = (char*)ws.small.get();
  RadonImg* d_ri = (RadonImg*)(small + a.cs.size() * 4);
  float2* d_S = (float2*)((char*)d_ri + nz * sizeof(RadonImg));
  float* d_ccv = (float*)(d_S + (size_t)num * nc);
  float* d_score = d_ccv + (size_t)num * nz;
  MF_HIP(hipGetLastError());
  const FftDevPlan& P = ctx->fft->plan(L / 2);
  const int ls = fft_lds_len(P);
  const size_t per = (size_t)2 * ls * sizeof(float2);
  const int nl = std::max(1, std::min(16, (int)((64 * 1024) / per)));
  const size_t lds = nl * per + (size_t)ls * sizeof(float2);
  MF_REQUIRE(lds + 512 <= 160 * 1024, "common line exceeds LDS");
  ensure_dyn_lds(ctx, (const void*)k_line_fft, lds);
  const int nlines = nz * num;
  hipLaunchKernelGGL(k_line_fft, dim3(cdiv(nlines, nl)), dim3(kLineThreads),
                     lds, st, lines, nlines, L, pitch, d_taper, P, nl);
  MF_HIP(hipGetLastError());
  const float2* spec = (const float2*)lines;
  hipLaunchKernelGGL(k_line_sum, dim3(cdiv((long long)num * nc, 256)),
                     dim3(256), 0, st, spec, nz, num, nc, pitch / 2, d_S);
  hipLaunchKernelGGL(k_line_cc, dim3(num, nz), dim3(256), 0, st, spec, d_S, nz,
                     num, nc, pitch / 2, d_filt, fsum, d_ccv);
  hipLaunchKernelGGL(k_line_score, dim3(cdiv(num, 64)), dim3(64), 0, st, d_ccv,
                     nz, num, d_score);
  MF_HIP(hipGetLastError());
  AxisResult res;
  res.scores.resize(num);
  res.L = L;
  MF_HIP(hipMemcpyAsync(res.scores.data(), d_score, num * sizeof(float),
                        hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
  // findTiltAxis (Rotate.cu:468-487): strict maximum from an initial 0.
  int best = 0;
  float bs = 0.0f;
  for (int k = 0; k < num; k++)
    if (bs < res.scores[k]) {
      bs = res.scores[k];
      best = k;
    }
  res.axis = a.rot[best];
  return res;
}

float offset_cc(const double m[6]) {
  // FindOffset::Correlate's finish (FindOffset.cu:218-235): moments
  // rounded to float as the reference reads them back, quirk Q7.
  float a[6];
  for (int t = 0; t < 6; t++) a[t] = (float)m[t];
  for (int t = 0; t < 5; t++) a[t] /= a[5];
  double s1 = a[2] - a[0], s2 = a[3] - a[1];
  double cc = a[4] - a[0] * a[1];
  if (s1 < 0) s1 = 0;
  if (s2 < 0) s2 = 0;
  s1 = std::sqrt(s1);
  s2 = std::sqrt(s2);
  cc = cc / (s1 * s2 + 1e-30);
  if (cc < -1) cc = -1;
  else if (cc > 1) cc = 1;
  return (float)cc;
}

void ncc_moments(mf_ctx* ctx, hipStream_t st, const float* d_ref,
                 const float* d_img, int w, int h, float ref_tilt, float tilt,
                 OffsetWork& ws, double m6[6]) {
  OffJob j;
  j.ref = d_ref;
  j.img = d_img;
  const double ds = std::cos(kD2R * ref_tilt) / std::cos(kD2R * tilt);
  stretch_matrix((float)ds, 0.0f, j.m);
  char* buf = (char*)ws.buf.ensure(ctx, sizeof(OffJob) + 6 * 8, st);
  OffJob* d_job = (OffJob*)buf;
  double* d_mom = (double*)(buf + sizeof(OffJob));
  MF_HIP(hipMemcpyAsync(d_job, &j, sizeof(OffJob), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_offset_moments, dim3(1), dim3(256), 0, st, w, h, d_job,
                     d_mom);
  MF_HIP(hipGetLastError());
  MF_HIP(hipMemcpyAsync(m6, d_mom, 6 * 8, hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
}

float offset_search(mf_ctx* ctx, hipStream_t st, const float* d_corr, int w,
                    int h, int nz, const float* tilts, OffsetWork& ws,
                    std::vector<float>* ccs) {
  // SearchOffset / CalcAverangedCC (FindOffset.cu:129-174) with the tilt copy
  // drifting exactly as the reference's member copy does.
  const int num = 31;
  std::vector<float> ang(tilts, tilts + nz);
  std::vector<OffJob> jobs;
  std::vector<int> z0s(num);
  std::vector<float> offs(num);
  for (int i = 0; i < num; i++) {
    const float off = 0.0f + (float)(i - num / 2) * 1.0f;
    offs[i] = off;
    for (int z = 0; z < nz; z++) ang[z] += off;
    int z0 = 0;
    float m = std::fabs(0.0f - ang[0]);
    for (int q = 1; q < nz; q++) {
      float d = std::fabs(0.0f - ang[q]);
      if (d >= m) continue;
      m = d;
      z0 = q;
    }
    z0s[i] = z0;
    for (int k = 0; k < nz; k++) {
      if (k == z0) continue;
      const int r = (k < z0) ? k + 1 : k - 1;
      OffJob j;
      j.ref = d_corr + (size_t)r * w * h;
      j.img = d_corr + (size_t)k * w * h;
      double ds = std::cos(kD2R * ang[r]) / std::cos(kD2R * ang[k]);
      stretch_matrix((float)ds, 0.0f, j.m);
      jobs.push_back(j);
    }
    for (int z = 0; z < nz; z++) ang[z] -= off;
  }
  const size_t nj = jobs.size();
  char* buf = (char*)ws.buf.ensure(ctx, nj * sizeof(OffJob) + nj * 6 * 8, st);
  OffJob* d_jobs = (OffJob*)buf;
  double* d_mom = (double*)(buf + nj * sizeof(OffJob));
  MF_HIP(hipMemcpyAsync(d_jobs, jobs.data(), nj * sizeof(OffJob),
                        hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_offset_moments, dim3((unsigned)nj), dim3(256), 0, st,
                     w, h, d_jobs, d_mom);
  MF_HIP(hipGetLastError());
  std::vector<double> mom(nj * 6);
  MF_HIP(hipMemcpyAsync(mom.data(), d_mom, nj * 6 * 8, hipMemcpyDeviceToHost,
                        st));
  MF_HIP(hipStreamSynchronize(st));
  float maxcc = 0.0f, best = 0.0f;
  size_t q = 0;
  if (ccs) ccs->assign(num, 0.0f);
  for (int i = 0; i < num; i++) {
    float sum = 0.0f;
    int count = 0;
    for (int k = 0; k < nz; k++) {
      if (k == z0s[i]) continue;
      const float cc = offset_cc(&mom[6 * q++]);
      sum += cc;
      count++;
    }
    const float mean = sum / (float)count;
    if (ccs) (*ccs)[i] = mean;
    if (mean > maxcc) {
      maxcc = mean;
      best = offs[i];
    }
  }
  return best;
}

}  // namespace mf
