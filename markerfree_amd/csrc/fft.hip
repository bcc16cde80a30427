// FFT plans (host) and the generic batched row / column kernels.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "context.h"
#include "devutil.h"
#include "fft_dev.h"

namespace mf {

namespace {

typedef std::complex<double> cd;

// Radix list for a length whose prime factors are all <= kMaxDirectPrime:
// radix-8 first, then 4/2, then odd primes.  Returns false otherwise.
bool factor_direct(int n, std::vector<int>& rad) {
  rad.clear();
  int m = n;
  int e2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    e2++;
  }
  while (e2 >= 3) {
    rad.push_back(8);
    e2 -= 3;
  }
  if (e2 == 2) rad.push_back(4);
  if (e2 == 1) rad.push_back(2);
  for (int p = 3; p <= kMaxDirectPrime && m > 1; p += 2) {
    while (m % p == 0) {
      rad.push_back(p);
      m /= p;
    }
  }
  return m == 1 && (int)rad.size() <= kMaxRad;
}

bool smooth235(int m) {
  for (int p : {2, 3, 5})
    while (m % p == 0) m /= p;
  return m == 1;
}

// Host double-precision DFT via recursive mixed radix (lengths 2^a 3^b 5^c).
void host_fft(std::vector<cd>& a, int sign) {
  const int n = (int)a.size();
  if (n == 1) return;
  int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 5;
  const int m = n / p;
  std::vector<std::vector<cd>> sub(p, std::vector<cd>(m));
  for (int j = 0; j < m; j++)
    for (int r = 0; r < p; r++) sub[r][j] = a[j * p + r];
  for (int r = 0; r < p; r++) host_fft(sub[r], sign);
  const double ang = sign * 2.0 * M_PI / n;
  for (int k = 0; k < n; k++) {
    cd acc = 0;
    for (int r = 0; r < p; r++)
      acc += sub[r][k % m] * std::polar(1.0, ang * (double)((long long)r * k % n));
    a[k] = acc;
  }
}

float2 f2(cd v) { return make_float2((float)v.real(), (float)v.imag()); }

}  // namespace

FftCache::~FftCache() {
  for (auto& kv : plans_) dev_free(ctx_, kv.second->mem, kv.second->bytes);
}

const FftDevPlan& FftCache::plan(int n) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = plans_.find(n);
  if (it != plans_.end()) return it->second->p;
  MF_REQUIRE(n >= 1 && n <= (1 << 22), "FFT length out of range");
  std::unique_ptr<Entry> e(new Entry);
  FftDevPlan& P = e->p;
  memset(&P, 0, sizeof(P));
  P.n = n;
  std::vector<int> rad;
  int L = n;
  if (!factor_direct(n, rad)) {
    P.blue = 1;
    int m = 2 * n - 1;
    while (!smooth235(m)) m++;
    P.m = m;
    L = m;
    MF_REQUIRE(factor_direct(m, rad), "Bluestein length not factorable");
  } else {
    P.m = n;
  }
  P.nrad = (int)rad.size();
  for (int i = 0; i < P.nrad; i++) P.rad[i] = rad[i];
  // Tables: tw (L), half (n+1), and for Bluestein chirp (n), bf, bi (m).
  std::vector<float2> host;
  auto push = [&](const std::vector<float2>& v) {
    size_t off = host.size();
    host.insert(host.end(), v.begin(), v.end());
    return off;
  };
  std::vector<float2> tw(L);
  for (int k = 0; k < L; k++)
    tw[k] = f2(std::polar(1.0, -2.0 * M_PI * (double)k / L));
  std::vector<float2> half(n + 1);
  for (int k = 0; k <= n; k++)
    half[k] = f2(std::polar(1.0, -M_PI * (double)k / n));
  size_t o_tw = push(tw), o_half = push(half), o_chirp = 0, o_bf = 0, o_bi = 0;
  // Radices, two ints per table entry.
  std::vector<float2> rv((P.nrad + 1) / 2 + 1);
  std::memcpy(rv.data(), P.rad, P.nrad * sizeof(int));
  const size_t o_rad = push(rv);
  if (P.blue) {
    const int m = P.m;
    std::vector<cd> chirp(n);
    for (int k = 0; k < n; k++) {
      long long q = ((long long)k * k) % (2LL * n);
      chirp[k] = std::polar(1.0, -M_PI * (double)q / n);
    }
    std::vector<float2> ch(n);
    for (int k = 0; k < n; k++) ch[k] = f2(chirp[k]);
    o_chirp = push(ch);
    for (int dir = 0; dir < 2; dir++) {
      // b_j = conj(chirp_s(j)) at j mod m, |j| < n; chirp_s = chirp (fwd)
      // or conj(chirp) (inverse).
      std::vector<cd> b(m, cd(0, 0));
      for (int k = 0; k < n; k++) {
        cd v = dir == 0 ? std::conj(chirp[k]) : chirp[k];
        b[k] = v;
        if (k) b[m - k] = v;
      }
      host_fft(b, -1);
      std::vector<float2> bv(m);
      for (int k = 0; k < m; k++) bv[k] = f2(b[k] / (double)m);
      (dir == 0 ? o_bf : o_bi) = push(bv);
    }
  }
  e->bytes = host.size() * sizeof(float2);
  e->mem = dev_alloc(ctx_, e->bytes);
  MF_HIP(hipMemcpy(e->mem, host.data(), e->bytes, hipMemcpyHostToDevice));
  float2* base = (float2*)e->mem;
  P.tw = base + o_tw;
  P.half = base + o_half;
  P.drad = reinterpret_cast<const int*>(base + o_rad);
  if (P.blue) {
    P.chirp = base + o_chirp;
    P.bf = base + o_bf;
    P.bi = base + o_bi;
  }
  auto& ref = e->p;
  plans_[n] = std::move(e);
  return ref;
}

int fft_lines_per_group(const FftDevPlan& p, size_t lds_budget, int max_lines) {
  size_t per = (size_t)2 * fft_lds_len(p) * sizeof(float2);
  int nl = (int)(lds_budget / per);
  return std::max(1, std::min(nl, max_lines));
}

namespace {

constexpr int kFftThreads = 256;
constexpr size_t kRowLds = 64 * 1024;
constexpr size_t kColLds = 128 * 1024;

// Batched real-to-complex transform of rows: length 2N reals per row with
// row pitch 'pitch' floats, in place.  Rows are indexed across images by
// row = img * rows_per_img + r, image base img * img_stride.
__global__ __launch_bounds__(kFftThreads) void k_r2c_rows(
    float* buf, int nrows, int rows_per_img, size_t img_stride, int pitch,
    FftDevPlan P, int nl, float scale) {
  extern __shared__ float2 sm2[];
  const int N = P.n, ls = fft_lds_len(P);
  float2* a = sm2;
  float2* b = sm2 + nl * ls;
  float2* tw = b + nl * ls;
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, nrows - r0);
  auto row_ptr = [&](int r) {
    int img = r / rows_per_img, rr = r - img * rows_per_img;
    return buf + img * img_stride + (size_t)rr * pitch;
  };
  fft_load_tw(P, tw);
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    float2 v = make_float2(0.f, 0.f);
    if (l < nv) v = reinterpret_cast<const float2*>(row_ptr(r0 + l))[k];
    a[l * ls + k] = v;
  }
  __syncthreads();
  float2* Z = fft_lines(a, b, P, nl, false, tw);
  for (LineIter it(N + 1); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    if (l >= nv) continue;
    float2* out = reinterpret_cast<float2*>(row_ptr(r0 + l));
    r2c_unpack(Z + l * ls, N, P.half, scale, k,
               [&](int kk, float2 v) { out[kk] = v; });
  }
}

__global__ __launch_bounds__(kFftThreads) void k_c2r_rows(
    float* buf, int nrows, int rows_per_img, size_t img_stride, int pitch,
    FftDevPlan P, int nl, float scale) {
  extern __shared__ float2 sm2[];
  const int N = P.n, ls = fft_lds_len(P);
  float2* a = sm2;
  float2* b = sm2 + nl * ls;  // holds nl * (ls + 1) entries
  float2* tw = b + nl * (ls + 1);
  const int r0 = blockIdx.x * nl;
  const int nv = min(nl, nrows - r0);
  auto row_ptr = [&](int r) {
    int img = r / rows_per_img, rr = r - img * rows_per_img;
    return buf + img * img_stride + (size_t)rr * pitch;
  };
  // Stage X_0..X_N in b (line stride ls + 1), then pack Z' into a.
  const int bs = ls + 1;
  fft_load_tw(P, tw);
  for (LineIter it(N + 1); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    float2 v = make_float2(0.f, 0.f);
    if (l < nv) v = reinterpret_cast<const float2*>(row_ptr(r0 + l))[k];
    b[l * bs + k] = v;
  }
  __syncthreads();
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    a[l * ls + k] = c2r_pack(b[l * bs + k], b[l * bs + N - k], P.half[k]);
  }
  __syncthreads();
  float2* z = fft_lines(a, b, P, nl, true, tw);
  for (LineIter it(N); it.l < nl; it.next()) {
    int l = it.l, k = it.k;
    if (l >= nv) continue;
    reinterpret_cast<float2*>(row_ptr(r0 + l))[k] = cscale(z[l * ls + k], scale);
  }
}

// Complex transform along columns of images of ncols x nrows complex values
// (row pitch 'cpitch' complex), nl adjacent columns per workgroup.  Adjacent
// column groups run on the same XCD (xcd_block), so the rows' cache lines
// they share are fetched into one L2.
__global__ __launch_bounds__(kFftThreads) void k_c2c_cols(
    float2* buf, int ncols, int cpitch, size_t img_stride_c, FftDevPlan P,
    int nl, int inv, float scale) {
  extern __shared__ float2 sm2[];
  const int N = P.n, ls = fft_lds_len(P);
  float2* a = sm2;
  float2* b = sm2 + nl * ls;
  float2* tw = b + nl * ls;
  const int groups = (ncols + nl - 1) / nl;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int img = blk / groups;
  const int c0 = (blk - img * groups) * nl;
  float2* base = buf + img * img_stride_c;
  fft_load_tw(P, tw);
  for (LineIter it(nl); it.l < N; it.next()) {
    int y = it.l, l = it.k;
    float2 v = make_float2(0.f, 0.f);
    if (c0 + l < ncols) v = base[(size_t)y * cpitch + c0 + l];
    a[l * ls + y] = v;
  }
  __syncthreads();
  float2* z = fft_lines(a, b, P, nl, inv != 0, tw);
  for (LineIter it(nl); it.l < N; it.next()) {
    int y = it.l, l = it.k;
    if (c0 + l < ncols)
      base[(size_t)y * cpitch + c0 + l] = cscale(z[l * ls + y], scale);
  }
}

void set_lds(mf_ctx* ctx, const void* k, size_t bytes) {
  ensure_dyn_lds(ctx, k, bytes);
}

void launch_rows(mf_ctx* ctx, hipStream_t st, float* buf, int n, int nrows,
                 int rows_per_img, size_t img_stride, int pitch, float scale,
                 bool r2c) {
  MF_REQUIRE(n % 2 == 0 && n >= 2, "real transform length must be even");
  const FftDevPlan& P = ctx->fft->plan(n / 2);
  const int nl = fft_lines_per_group(P, kRowLds, 16);
  const size_t lds =
      ((size_t)nl * (2 * fft_lds_len(P) + 1) + fft_lds_len(P)) * sizeof(float2);
  MF_REQUIRE(lds <= 160 * 1024, "FFT row exceeds LDS");
  dim3 grid(cdiv(nrows, nl));
  TimedLaunch tl(ctx->timer, "fft_rows", st);
  if (r2c) {
    set_lds(ctx, (const void*)k_r2c_rows, lds);
    hipLaunchKernelGGL(k_r2c_rows, grid, dim3(kFftThreads), lds, st, buf, nrows,
                       rows_per_img, img_stride, pitch, P, nl, scale);
  } else {
    set_lds(ctx, (const void*)k_c2r_rows, lds);
    hipLaunchKernelGGL(k_c2r_rows, grid, dim3(kFftThreads), lds, st, buf, nrows,
                       rows_per_img, img_stride, pitch, P, nl, scale);
  }
  MF_HIP(hipGetLastError());
}

void launch_cols(mf_ctx* ctx, hipStream_t st, float* buf, int nx, int ny,
                 int nimg, size_t img_stride, bool inv, float scale) {
  const FftDevPlan& P = ctx->fft->plan(ny);
  const int ncols = nx / 2 + 1;
  // Columns per workgroup: as many as LDS allows, but no more than keeps
  // about two workgroups per CU busy.
  const long long want = (long long)ncols * nimg / (2LL * std::max(ctx->num_cu, 1));
  const int nl = std::max(1, std::min<int>(fft_lines_per_group(P, kColLds, 16),
                                           (int)std::max(1LL, want)));
  const size_t lds =
      ((size_t)2 * nl * fft_lds_len(P) + fft_lds_len(P)) * sizeof(float2);
  MF_REQUIRE(lds <= 160 * 1024, "FFT column exceeds LDS");
  dim3 grid(cdiv(ncols, nl) * nimg);
  set_lds(ctx, (const void*)k_c2c_cols, lds);
  TimedLaunch tl(ctx->timer, "fft_cols", st);
  hipLaunchKernelGGL(k_c2c_cols, grid, dim3(kFftThreads), lds, st,
                     (float2*)buf, ncols, ncols, img_stride / 2, P, nl,
                     inv ? 1 : 0, scale);
  MF_HIP(hipGetLastError());
}

}  // namespace

void fft_r2c_2d(mf_ctx* ctx, hipStream_t st, float* buf, int nx, int ny,
                int nimg, size_t img_stride, float scale) {
  MF_REQUIRE(nx >= 2 && ny >= 1 && nimg >= 1, "bad 2-D FFT shape");
  const int pitch = pad_x(nx);
  launch_rows(ctx, st, buf, nx, ny * nimg, ny, img_stride, pitch, 1.0f, true);
  launch_cols(ctx, st, buf, nx, ny, nimg, img_stride, false, scale);
}

void fft_c2r_2d(mf_ctx* ctx, hipStream_t st, float* buf, int nx, int ny,
                int nimg, size_t img_stride) {
  MF_REQUIRE(nx >= 2 && ny >= 1 && nimg >= 1, "bad 2-D FFT shape");
  const int pitch = pad_x(nx);
  launch_cols(ctx, st, buf, nx, ny, nimg, img_stride, true, 1.0f);
  launch_rows(ctx, st, buf, nx, ny * nimg, ny, img_stride, pitch, 1.0f, false);
}

void fft_r2c_rows(mf_ctx* ctx, hipStream_t st, float* buf, int n, int nlines,
                  int pitch, float scale) {
  launch_rows(ctx, st, buf, n, nlines, nlines, 0, pitch, scale, true);
}

void fft_c2r_rows(mf_ctx* ctx, hipStream_t st, float* buf, int n, int nlines,
                  int pitch, float scale) {
  launch_rows(ctx, st, buf, n, nlines, nlines, 0, pitch, scale, false);
}

}  // namespace mf
