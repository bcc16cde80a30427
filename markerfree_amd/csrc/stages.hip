// Preprocessing, stack correction / binning and R-weighting kernels.
//
// Reference: PreProcess (method/Preprocess.cu), CorrTomoStack, BinImage,
// GCorrPatchShift and RWeight (method/CorrStack.cu).  The reference corrects
// one projection per call after a disk read and H2D copy of it; here the
// preprocessed stack stays in HBM and one launch corrects any number of
// projections.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "context.h"
#include "devutil.h"
#include "fft_dev.h"
#include "stages.h"

namespace mf {

CorrGeom CorrGeom::make(int nx, int ny, int bin) {
  CorrGeom g;
  g.nx = nx;
  g.ny = ny;
  g.cx = (pad_x(nx) / 2 - 1) * 2;
  g.bx = nx / bin / 2 * 2;
  g.by = ny / bin / 2 * 2;
  MF_REQUIRE(g.bx >= 2 && g.by >= 2, "binned image too small");
  g.binx = g.cx / g.bx;
  g.biny = ny / g.by;
  return g;
}

namespace {

constexpr int kRedThreads = 1024;

// Min and max of part blockIdx.x (of kMinMaxParts) of slice blockIdx.y;
// out [nz][kMinMaxParts][2] (the host folds the parts: min and max do not
// depend on the order).
constexpr int kMinMaxParts = 16;
__global__ __launch_bounds__(kRedThreads) void k_minmax(const float* d,
                                                        size_t plane,
                                                        float* out) {
  __shared__ float sh[16];
  const size_t part = (plane + kMinMaxParts - 1) / kMinMaxParts;
  const size_t i0 = blockIdx.x * part, i1 = min(plane, i0 + part);
  const float* p = d + (size_t)blockIdx.y * plane;
  float mn = 1e20f, mx = -1e20f;
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    float v = p[i];
    if (v < -1e10f) continue;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  mn = block_reduce(mn, sh, [](float a, float b) { return fminf(a, b); }, 1e20f);
  mx = block_reduce(mx, sh, [](float a, float b) { return fmaxf(a, b); }, -1e20f);
  if (threadIdx.x == 0) {
    const size_t o = ((size_t)blockIdx.y * kMinMaxParts + blockIdx.x) * 2;
    out[o] = mn;
    out[o + 1] = mx;
  }
}

// Image blockIdx.y of the raw stack to position blockIdx.y (nz - 1 -
// blockIdx.y when 'rev') of the stack, minus 'sub' where valid (0: a plain
// copy): the reordering and the positivity shift in one pass.
__global__ void k_reorder_subtract(const float* raw, float* stack,
                                   size_t plane, int nz, int rev, float sub) {
  const int a = blockIdx.y;
  const float* src = raw + (size_t)a * plane;
  float* dst = stack + (size_t)(rev ? nz - 1 - a : a) * plane;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
       i += (size_t)gridDim.x * blockDim.x) {
    const float v = src[i];
    dst[i] = (sub != 0.0f && !(v < -1e10f)) ? v - sub : v;
  }
}

__global__ void k_subtract(float* d, size_t n, float sub) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    float v = d[i];
    if (v < -1e10f) continue;
    d[i] = v - sub;
  }
}

// Sum and count of the pixels above -1e20 in the box [x0, x0+w) x [y0, y0+h)
// of image blockIdx.y, in kBoxParts row bands (part blockIdx.x); FP64 sums,
// combined by the host in part order.
constexpr int kBoxParts = 16;
__global__ __launch_bounds__(256) void k_box_mean(
    const float* d, int nx, int ny, int x0, int y0, int w, int h,
    double* out) {
  __shared__ double sh[16];
  __shared__ int shc[16];
  const float* p = d + (size_t)blockIdx.y * nx * ny;
  const int rows = (h + kBoxParts - 1) / kBoxParts;
  const int ya = min(h, (int)blockIdx.x * rows), yb = min(h, ya + rows);
  double s = 0.0;
  int c = 0;
  for (int y = ya; y < yb; y++) {
    const float* row = p + (size_t)(y + y0) * nx + x0;
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
      const float v = row[x];
      if (v <= -1e20f) continue;
      s += v;
      c++;
    }
  }
  s = block_reduce(s, sh, [](double a, double b) { return a + b; }, 0.0);
  c = block_reduce(c, shc, [](int a, int b) { return a + b; }, 0);
  if (threadIdx.x == 0) {
    double* o = out + ((size_t)blockIdx.y * kBoxParts + blockIdx.x) * 2;
    o[0] = s;
    o[1] = (double)c;
  }
}

__global__ void k_scale(float* d, size_t plane, const float* scale) {
  float* p = d + (size_t)blockIdx.y * plane;
  const float f = scale[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
       i += (size_t)gridDim.x * blockDim.x) {
    float v = p[i];
    if (v <= -1e20f) continue;
    p[i] = v * f;
  }
}

// Raw pixel (x, y) of an nx-wide image: 24-bit product (both factors below
// 2^24, the index below 2^32) and a 32-bit offset, at full rate.
__device__ __forceinline__ float raw_pix(const float* raw, int nx, int x,
                                         int y) {
  return raw[__umul24((unsigned)y, (unsigned)nx) + (unsigned)x];
}

// mGRandom (CorrStack.cu:116-140): window 51, next = (19 next + 57) mod 2601.
__device__ float corr_random(int x, int y, int nx, int ny, const float* raw) {
  if (x < 0) x = -x;
  if (y < 0) y = -y;
  if (x >= nx) x = nx - 1 - (x % nx);
  if (y >= ny) y = ny - 1 - (y % ny);
  unsigned next = (unsigned)(y * nx + x);
  for (int i = 0; i < 20; i++) {
    next = (next * 19u + 57u) % 2601u;
    int ix = (int)next - 25 + x;
    if (ix < 0 || ix >= nx) continue;
    int iy = (int)(next / 51u) - 25 + y;
    if (iy < 0 || iy >= ny) continue;
    return raw_pix(raw, nx, ix, iy);
  }
  return raw_pix(raw, nx, x, y);
}

struct CorrArgs {
  const float* raw;   // [nz][ny][nx]
  const float* rawT;  // the same stack transposed, [nz][nx][ny], or null
  const CorrProj* proj;
  float* out;
  int nx, ny, cx, bx, by, binx, biny, z0, fill, out_pitch;
  size_t out_stride;
  int zstep;  // k_correct_rweight: image z0 + blockIdx.y * zstep
};

// Whether projection pr reads the transposed copy: its rotation is nearer a
// quarter turn than none, so an output row runs down the raw image's
// columns, which are the transposed copy's rows (the same pixels: the
// values are unchanged).
__device__ __forceinline__ bool corr_transposed(const CorrArgs& a,
                                                const CorrProj& pr) {
  return a.rawT && fabsf(pr.sin_a) > fabsf(pr.cos_a);
}

// Binned output pixel (x, y) of projection z: each of its bin x bin corrected
// pixels is the nearest raw pixel at R(axis) p + shift + centre (mGCorrect),
// random filled or invalid outside; the bin sum is invalid if any is
// (mGBinImage).
template <bool TR>
__device__ __forceinline__ float corr_pixel(const CorrArgs& a,
                                            const CorrProj& pr,
                                            const float* raw, const float* rt,
                                            int x, int y) {
  const float hx = (float)a.cx * 0.5f, hy = (float)a.ny * 0.5f;
  const float ox = pr.shift_x + (float)a.nx * 0.5f;
  const float oy = pr.shift_y + (float)a.ny * 0.5f;
  // The bin's binx*biny corrected pixels are summed in the reference's
  // (row-major) order.  They are visited in batches of kB whose in-image
  // reads are all issued before any is used, so they are in flight together;
  // slots past the bin are skipped by a uniform branch and read nothing.
  // (The reference stops at the first invalid pixel, after which its sum is
  // discarded: summing on changes nothing.)
  constexpr int kB = 8;
  const int nb = a.binx * a.biny;
  float sum = 0.f;
  bool bad = false;
  int ix = 0, iy = 0;  // bin position of the batch's first pixel
  for (int b0 = 0; b0 < nb; b0 += kB) {
    int rx[kB], ry[kB];
    bool in[kB];
    float v[kB];
    int jx = ix, jy = iy;
#pragma unroll
    for (int j = 0; j < kB; j++) {
      if (b0 + j < nb) {
        const float px = (float)(x * a.binx + jx) + 0.5f - hx;
        const float py = (float)(y * a.biny + jy) + 0.5f - hy;
        float qx = px * pr.cos_a - py * pr.sin_a;
        float qy = px * pr.sin_a + py * pr.cos_a;
        qx += ox;
        qy += oy;
        rx[j] = (int)(qx + 0.5f);
        ry[j] = (int)(qy + 0.5f);
        in[j] = rx[j] >= 0 && rx[j] < a.nx && ry[j] >= 0 && ry[j] < a.ny;
        v[j] = in[j] ? (TR ? raw_pix(rt, a.ny, ry[j], rx[j])
                           : raw_pix(raw, a.nx, rx[j], ry[j]))
                     : 0.f;
        if (++jx == a.binx) {
          jx = 0;
          jy++;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kB; j++) {
      if (b0 + j < nb) {
        if (!in[j])
          v[j] = a.fill ? corr_random(rx[j], ry[j], a.nx, a.ny, raw) : kInvalid;
        bad = bad || v[j] < -1e10f;
        sum += v[j];
      }
    }
    ix = jx;
    iy = jy;
  }
  return bad ? kInvalid : sum;
}

// Binned output pixels, grid-strided over the image (a few per thread: one
// per thread made the launch dispatch-bound).
constexpr int kBinChunk = 1024;  // pixels of a row per workgroup
__global__ __launch_bounds__(256) void k_correct_bin(const CorrArgs a) {
  const int z = a.z0 + blockIdx.y;
  const CorrProj pr = a.proj[z];
  const size_t plane = (size_t)a.nx * a.ny;
  const float* raw = a.raw + z * plane;
  const bool tr = corr_transposed(a, pr);
  // Workgroup b: row b / nch, columns [c, c + kBinChunk) with c the chunk
  // b % nch times kBinChunk (one division per workgroup, none per pixel).
  const int nch = (a.bx + kBinChunk - 1) / kBinChunk;
  const int y = (int)blockIdx.x / nch;
  const int c = ((int)blockIdx.x - y * nch) * kBinChunk;
  const int ce = min(c + kBinChunk, a.bx);
  for (int x = c + (int)threadIdx.x; x < ce; x += blockDim.x) {
    float* o = a.out + (size_t)blockIdx.y * a.out_stride + (size_t)y * a.out_pitch + x;
    if (tr)
      *o = corr_pixel<true>(a, pr, raw, a.rawT + z * plane, x, y);
    else
      *o = corr_pixel<false>(a, pr, raw, nullptr, x, y);
  }
}

constexpr int kRowThreads = 256;

// R-weighting of nl padded rows per workgroup: R2C (length bx), scale by
// 1/N and the ramp filter, C2R, unpadded write (RWeight::DoIt + Unpad).
__global__ __launch_bounds__(kRowThreads) void k_rweight(
    const float* in, int pitch, float* out, int nrows, FftDevPlan P, int nl,
    const float* filt) {
  extern __shared__ float2 sm2[];
  const int N = P.n, ls = fft_lds_len(P), xs = N + 1;
  const float fn = 2.0f * ((float)(N + 1) - 1.0f);
  float2* a = sm2;
  float2* b = a + nl * ls;
  float2* X = b + nl * ls;
  float2* tw = X + nl * xs;
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, nrows - r0);
  fft_load_tw(P, tw);
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    float2 v = make_float2(0.f, 0.f);
    if (l < nv) v = reinterpret_cast<const float2*>(in + (size_t)(r0 + l) * pitch)[k];
    a[l * ls + k] = v;
  }
  __syncthreads();
  float2* Z = fft_lines(a, b, P, nl, false, tw);
  for (LineIter it(xs); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    r2c_unpack(Z + l * ls, N, P.half, 1.0f, k, [&](int kk, float2 v) {
      // GRWeight (CorrStack.cu:431-446): x /= N, then x *= ramp.
      const float f = filt[kk];
      X[l * xs + kk] = make_float2((v.x / fn) * f, (v.y / fn) * f);
    });
  }
  __syncthreads();
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    a[l * ls + k] = c2r_pack(X[l * xs + k], X[l * xs + N - k], P.half[k]);
  }
  __syncthreads();
  float2* z = fft_lines(a, b, P, nl, true, tw);
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    if (l >= nv) continue;
    reinterpret_cast<float2*>(out + (size_t)(r0 + l) * 2 * N)[k] = z[l * ls + k];
  }
}

// Unbinned corrected pixel (x, y) of projection pr from the raw value v it
// gathered (or did not: 'in' false), as corr_pixel computes it for a 1 x 1
// bin: the random fill or the invalid value outside the image, and the bin
// sum's 0 + v.
__device__ __forceinline__ void corr_coords(const CorrArgs& a,
                                            const CorrProj& pr, int x, int y,
                                            int* rx, int* ry, bool* in) {
  const float hx = (float)a.cx * 0.5f, hy = (float)a.ny * 0.5f;
  const float ox = pr.shift_x + (float)a.nx * 0.5f;
  const float oy = pr.shift_y + (float)a.ny * 0.5f;
  const float px = (float)x + 0.5f - hx;
  const float py = (float)y + 0.5f - hy;
  float qx = px * pr.cos_a - py * pr.sin_a;
  float qy = px * pr.sin_a + py * pr.cos_a;
  qx += ox;
  qy += oy;
  *rx = (int)(qx + 0.5f);
  *ry = (int)(qy + 0.5f);
  *in = *rx >= 0 && *rx < a.nx && *ry >= 0 && *ry < a.ny;
}
__device__ __forceinline__ float corr_finish(const CorrArgs& a, const float* raw,
                                             float v, int rx, int ry, bool in) {
  if (!in) v = a.fill ? corr_random(rx, ry, a.nx, a.ny, raw) : kInvalid;
  float sum = 0.f;
  sum += v;
  return v < -1e10f ? kInvalid : sum;
}

// Correction + binning (k_correct_bin) fused with the R-weighting of the
// rows it produces (k_rweight): nl binned rows of projection z0 + blockIdx.y
// per workgroup go from the raw image to the unpadded R-weighted output
// without passing through HBM.  Same operations, same results.  The
// workgroup's latency is its chain of memory round trips, so they are
// shortened: unbinned images gather all of a thread's pixels (kCorrBatch
// complex entries) before any is used, and the twiddles and the ramp filter
// go to LDS in the same round trip as the first gathers.
constexpr int kCorrBatch = 4;
__global__ __launch_bounds__(kRowThreads) void k_correct_rweight(
    const CorrArgs ca, FftDevPlan P, int nl, const float* filt) {
  extern __shared__ float2 sm2[];
  const int N = P.n, ls = fft_lds_len(P), xs = N + 1;
  const float fn = 2.0f * ((float)(N + 1) - 1.0f);
  float2* a = sm2;
  float2* b = a + nl * ls;
  float2* X = b + nl * ls;
  float2* tw = X + nl * xs;
  float* fl = reinterpret_cast<float*>(tw + ls);  // ramp filter, N + 1 values
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, ca.by - r0);
  const int z = ca.z0 + (int)blockIdx.y * ca.zstep;
  const CorrProj pr = ca.proj[z];
  const size_t plane = (size_t)ca.nx * ca.ny;
  const float* raw = ca.raw + z * plane;
  const bool tr = corr_transposed(ca, pr);
  const float* rt = tr ? ca.rawT + z * plane : nullptr;
  if (ca.binx == 1 && ca.biny == 1) {
    // Entries e = l * N + k (row l, complex column k: pixels 2k, 2k + 1),
    // kCorrBatch per thread per round, every gather of a round in flight.
    const int ne = nl * N;
    const int rounds = cdiv(ne, kCorrBatch * (int)blockDim.x);
    // The first kCorrBatch twiddles and filter values of each thread, loaded
    // into registers ahead of the first gathers and written to LDS after
    // them (the rest, for lines above 1024 points, in a loop).
    const int bd = blockDim.x, ln = N + 1;
    float2 twr[kCorrBatch];
    float flr[kCorrBatch];
#pragma unroll
    for (int j = 0; j < kCorrBatch; j++) {
      const int i = (int)threadIdx.x + j * bd;
      twr[j] = i < ls ? P.tw[i] : make_float2(0.f, 0.f);
      flr[j] = i < ln ? filt[i] : 0.f;
    }
    for (int rd = 0; rd < rounds; rd++) {
      const int e0 = rd * kCorrBatch * (int)blockDim.x + (int)threadIdx.x;
      int rx[kCorrBatch][2], ry[kCorrBatch][2];
      bool in[kCorrBatch][2];
      float v[kCorrBatch][2];
#pragma unroll
      for (int j = 0; j < kCorrBatch; j++) {
        const int e = e0 + j * (int)blockDim.x;
        const int l = e / N, k = e - l * N;
        const bool use = e < ne && l < nv;
#pragma unroll
        for (int h = 0; h < 2; h++) {
          corr_coords(ca, pr, 2 * k + h, r0 + l, &rx[j][h], &ry[j][h], &in[j][h]);
          in[j][h] = in[j][h] && use;
          v[j][h] = in[j][h] ? (tr ? raw_pix(rt, ca.ny, ry[j][h], rx[j][h])
                                   : raw_pix(raw, ca.nx, rx[j][h], ry[j][h]))
                             : 0.f;
        }
      }
      if (rd == 0) {
#pragma unroll
        for (int j = 0; j < kCorrBatch; j++) {
          const int i = (int)threadIdx.x + j * bd;
          if (i < ls) tw[i] = twr[j];
          if (i < ln) fl[i] = flr[j];
        }
        for (int i = (int)threadIdx.x + kCorrBatch * bd; i < ls; i += bd)
          tw[i] = P.tw[i];
        for (int i = (int)threadIdx.x + kCorrBatch * bd; i < ln; i += bd)
          fl[i] = filt[i];
      }
#pragma unroll
      for (int j = 0; j < kCorrBatch; j++) {
        const int e = e0 + j * (int)blockDim.x;
        if (e >= ne) continue;
        const int l = e / N, k = e - l * N;
        float2 o = make_float2(0.f, 0.f);
        if (l < nv)
          o = make_float2(
              corr_finish(ca, raw, v[j][0], rx[j][0], ry[j][0], in[j][0]),
              corr_finish(ca, raw, v[j][1], rx[j][1], ry[j][1], in[j][1]));
        a[l * ls + k] = o;
      }
    }
  } else {
    fft_load_tw(P, tw);
    for (int k = threadIdx.x; k <= N; k += blockDim.x) fl[k] = filt[k];
    if (tr) {
      for (LineIter it(N); it.l < nl; it.next()) {
        int l = it.l, k = it.k;
        float2 v = make_float2(0.f, 0.f);
        if (l < nv)
          v = make_float2(corr_pixel<true>(ca, pr, raw, rt, 2 * k, r0 + l),
                          corr_pixel<true>(ca, pr, raw, rt, 2 * k + 1, r0 + l));
        a[l * ls + k] = v;
      }
    } else {
      for (LineIter it(N); it.l < nl; it.next()) {
        int l = it.l, k = it.k;
        float2 v = make_float2(0.f, 0.f);
        if (l < nv)
          v = make_float2(corr_pixel<false>(ca, pr, raw, nullptr, 2 * k, r0 + l),
                          corr_pixel<false>(ca, pr, raw, nullptr, 2 * k + 1, r0 + l));
        a[l * ls + k] = v;
      }
    }
  }
  __syncthreads();
  float2* Z = fft_lines(a, b, P, nl, false, tw);
  for (LineIter it(xs); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    r2c_unpack(Z + l * ls, N, P.half, 1.0f, k, [&](int kk, float2 v) {
      const float f = fl[kk];
      X[l * xs + kk] = make_float2((v.x / fn) * f, (v.y / fn) * f);
    });
  }
  __syncthreads();
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    a[l * ls + k] = c2r_pack(X[l * xs + k], X[l * xs + N - k], P.half[k]);
  }
  __syncthreads();
  float2* zz = fft_lines(a, b, P, nl, true, tw);
  float* out = ca.out + (ptrdiff_t)blockIdx.y * ca.zstep * (ptrdiff_t)ca.out_stride;
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    if (l >= nv) continue;
    reinterpret_cast<float2*>(out + (size_t)(r0 + l) * 2 * N)[k] =
        zz[l * ls + k];
  }
}

__global__ __launch_bounds__(kRedThreads) void k_moments(const float* d, int w,
                                                         int h, int pitch,
                                                         size_t stride,
                                                         double* out) {
  __shared__ double sh[16];
  const float* p = d + (size_t)blockIdx.x * stride;
  double s1 = 0.0, s2 = 0.0;
  const int n = w * h;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int y = i / w, x = i - y * w;
    float v = p[(size_t)y * pitch + x];
    if (v < -1e10f) continue;
    s1 += v;
    s2 += (double)(v * v);
  }
  s1 = block_reduce(s1, sh, [](double a, double b) { return a + b; }, 0.0);
  s2 = block_reduce(s2, sh, [](double a, double b) { return a + b; }, 0.0);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = s1;
    out[2 * blockIdx.x + 1] = s2;
  }
}

// Ramp-filter table of GRWeight for row length n (host-evaluated cosine).
const float* rweight_filter(mf_ctx* ctx, int n) {
  return (const float*)ctx_table(ctx, "rweight/" + std::to_string(n), [&] {
    const int nc = n / 2 + 1;
    const float fn = 2.0f * ((float)nc - 1.0f);
    std::vector<float> h(nc);
    for (int k = 0; k < nc; k++) {
      float r = (float)k / fn;
      h[k] = 2.0f * r * (0.55f + 0.45f * hcos(6.2831852f * r));
    }
    return h;
  });
}

}  // namespace

void slice_minmax(mf_ctx* ctx, hipStream_t st, const float* d, size_t plane,
                  int nz, DevBuf& tmp, float* h_out) {
  const size_t nparts = (size_t)nz * kMinMaxParts;
  float* buf = (float*)tmp.ensure(ctx, sizeof(float) * 2 * nparts, st);
  hipLaunchKernelGGL(k_minmax, dim3(kMinMaxParts, nz), dim3(kRedThreads), 0, st,
                     d, plane, buf);
  MF_HIP(hipGetLastError());
  std::vector<float> parts(2 * nparts);
  MF_HIP(hipMemcpyAsync(parts.data(), buf, sizeof(float) * 2 * nparts,
                        hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
  for (int z = 0; z < nz; z++) {
    float mn = 1e20f, mx = -1e20f;
    for (int k = 0; k < kMinMaxParts; k++) {
      mn = std::fmin(mn, parts[2 * ((size_t)z * kMinMaxParts + k)]);
      mx = std::fmax(mx, parts[2 * ((size_t)z * kMinMaxParts + k) + 1]);
    }
    h_out[2 * z] = mn;
    h_out[2 * z + 1] = mx;
  }
}

void reorder_subtract(mf_ctx* ctx, hipStream_t st, const float* d_raw,
                      float* d_stack, size_t plane, int nz, bool rev,
                      float sub) {
  (void)ctx;
  const int blocks = (int)std::min<size_t>(cdiv((long long)plane, 256 * 4), 1024);
  hipLaunchKernelGGL(k_reorder_subtract, dim3(blocks, nz), dim3(256), 0, st,
                     d_raw, d_stack, plane, nz, rev ? 1 : 0, sub);
  MF_HIP(hipGetLastError());
}

void box_means(mf_ctx* ctx, hipStream_t st, const float* d, int nx, int ny,
               int nz, int x0, int y0, int w, int h, DevBuf& tmp,
               float* h_out) {
  const size_t nb = (size_t)2 * nz * kBoxParts;
  double* buf = (double*)tmp.ensure(ctx, sizeof(double) * nb, st);
  hipLaunchKernelGGL(k_box_mean, dim3(kBoxParts, nz), dim3(256), 0, st, d, nx,
                     ny, x0, y0, w, h, buf);
  MF_HIP(hipGetLastError());
  std::vector<double> hb(nb);
  MF_HIP(hipMemcpyAsync(hb.data(), buf, sizeof(double) * nb,
                        hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
  for (int z = 0; z < nz; z++) {
    double sum = 0.0, cnt = 0.0;
    for (int k = 0; k < kBoxParts; k++) {
      sum += hb[(z * kBoxParts + k) * 2];
      cnt += hb[(z * kBoxParts + k) * 2 + 1];
    }
    h_out[z] = cnt == 0 ? 0.0f : (float)(sum / cnt);
  }
}

void scale_slices(mf_ctx* ctx, hipStream_t st, float* d, size_t plane, int nz,
                  DevBuf& tmp, const float* h_scale) {
  float* s = (float*)tmp.ensure(ctx, sizeof(float) * nz, st);
  MF_HIP(hipMemcpyAsync(s, h_scale, sizeof(float) * nz, hipMemcpyHostToDevice,
                        st));
  hipLaunchKernelGGL(k_scale, dim3(256, nz), dim3(256), 0, st, d, plane, s);
  MF_HIP(hipGetLastError());
  MF_HIP(hipStreamSynchronize(st));
}

void correct_bin(mf_ctx* ctx, hipStream_t st, const float* d_raw,
                 const CorrGeom& g, const CorrProj* d_proj, int z0, int count,
                 bool random_fill, float* d_out, int out_pitch,
                 size_t out_stride, const float* d_raw_t) {
  if (count <= 0) return;
  CorrArgs a;
  a.raw = d_raw;
  a.rawT = d_raw_t;
  a.proj = d_proj;
  a.out = d_out;
  a.nx = g.nx;
  a.ny = g.ny;
  a.cx = g.cx;
  a.bx = g.bx;
  a.by = g.by;
  a.binx = g.binx;
  a.biny = g.biny;
  a.z0 = z0;
  a.fill = random_fill ? 1 : 0;
  a.out_pitch = out_pitch;
  a.out_stride = out_stride;
  a.zstep = 1;
  TimedLaunch tl(ctx->timer, "correct", st);
  // One row chunk of up to kBinChunk pixels per workgroup (four per thread).
  hipLaunchKernelGGL(k_correct_bin,
                     dim3(g.by * cdiv(g.bx, kBinChunk), count),
                     dim3(256), 0, st, a);
  MF_HIP(hipGetLastError());
}

void rweight_rows(mf_ctx* ctx, hipStream_t st, const float* d_pad, int bx,
                  int nrows, float* d_out) {
  MF_REQUIRE(bx % 2 == 0 && bx >= 2, "R-weighting needs an even row length");
  const FftDevPlan& P = ctx->fft->plan(bx / 2);
  const int ls = fft_lds_len(P);
  const size_t per = (size_t)(2 * ls + P.n + 1) * sizeof(float2);
  const int nl = std::max(1, std::min(16, (int)((64 * 1024) / per)));
  const size_t lds = nl * per + (size_t)ls * sizeof(float2);
  MF_REQUIRE(lds <= 160 * 1024, "R-weighting row exceeds LDS");
  ensure_dyn_lds(ctx, (const void*)k_rweight, lds);
  const float* filt = rweight_filter(ctx, bx);
  TimedLaunch tl(ctx->timer, "rweight", st);
  hipLaunchKernelGGL(k_rweight, dim3(cdiv(nrows, nl)), dim3(kRowThreads), lds,
                     st, d_pad, pad_x(bx), d_out, nrows, P, nl, filt);
  MF_HIP(hipGetLastError());
}

void correct_rweight(mf_ctx* ctx, hipStream_t st, const float* d_raw,
                     const CorrGeom& g, const CorrProj* d_proj, int z0,
                     int count, bool random_fill, float* d_out, int zstep,
                     const float* d_raw_t) {
  if (count <= 0) return;
  MF_REQUIRE(zstep != 0 || count == 1, "image step must be nonzero");
  MF_REQUIRE(g.bx % 2 == 0 && g.bx >= 2, "R-weighting needs an even row length");
  CorrArgs a;
  a.raw = d_raw;
  a.rawT = d_raw_t;
  a.proj = d_proj;
  a.out = d_out;
  a.nx = g.nx;
  a.ny = g.ny;
  a.cx = g.cx;
  a.bx = g.bx;
  a.by = g.by;
  a.binx = g.binx;
  a.biny = g.biny;
  a.z0 = z0;
  a.fill = random_fill ? 1 : 0;
  a.out_pitch = g.bx;
  a.out_stride = (size_t)g.bx * g.by;
  a.zstep = zstep;
  const FftDevPlan& P = ctx->fft->plan(g.bx / 2);
  const int ls = fft_lds_len(P);
  const size_t per = (size_t)(2 * ls + P.n + 1) * sizeof(float2);
  // Rows per workgroup: within 48 KB of LDS, and no more than keeps about
  // eight workgroups per CU busy; 256 threads (the per-pixel correction
  // gathers dominate).  Tunables "corr_rw/<bx>x<by>x<count>/lines" and
  // ".../threads" override the heuristic.
  const std::string key = "corr_rw/" + std::to_string(g.bx) + "x" +
                          std::to_string(g.by) + "x" + std::to_string(count) + "/";
  const long long want = (long long)g.by * count / (8LL * std::max(ctx->num_cu, 1));
  const int fit = std::max(
      1, (int)((48 * 1024 - (size_t)ls * sizeof(float2) -
                (size_t)(P.n + 1) * sizeof(float)) / per));
  int nl = (int)std::max(1LL, std::min<long long>({want, (long long)fit, 16LL}));
  nl = std::max(1, std::min(tune_get(ctx, key + "lines", nl), fit));
  int threads = tune_get(ctx, key + "threads", kRowThreads);
  MF_REQUIRE(threads >= 64 && threads % 64 == 0 && threads <= kRowThreads,
             "workgroup size must be a multiple of 64 up to 256");
  const size_t lds = nl * per + (size_t)ls * sizeof(float2) +
                     (size_t)(P.n + 1) * sizeof(float);
  MF_REQUIRE(lds <= 160 * 1024, "R-weighting row exceeds LDS");
  ensure_dyn_lds(ctx, (const void*)k_correct_rweight, lds);
  const float* filt = rweight_filter(ctx, g.bx);
  TimedLaunch tl(ctx->timer, "correct_rweight", st);
  hipLaunchKernelGGL(k_correct_rweight, dim3(cdiv(g.by, nl), count),
                     dim3(threads), lds, st, a, P, nl, filt);
  MF_HIP(hipGetLastError());
}

void image_moments(mf_ctx* ctx, hipStream_t st, const float* d, int w, int h,
                   int pitch, size_t stride, int nimg, double* d_out) {
  (void)ctx;
  hipLaunchKernelGGL(k_moments, dim3(nimg), dim3(kRedThreads), 0, st, d, w, h,
                     pitch, stride, d_out);
  MF_HIP(hipGetLastError());
}

}  // namespace mf
