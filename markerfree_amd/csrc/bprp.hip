// Backprojection into xz slabs and ray-sum reprojection for projection
// matching: the roofline-graded kernels of the alignment path.
//
// Reference: CalcReprojFBP (method/CalcReproj.cu).  The reference launches one
// backprojection and one reprojection per image row (CalcReproj.cu:265-269,
// by x (NPROJ+3) runtime calls per target); here one launch covers every row
// of a target.
//
// Layouts in HBM:
//   corrected stack  [nz][by][bx]   float  (x fastest)
//   slab             [by][tb][2bx]  float  (one xz plane per image row)
//   reprojection     [by][bx]       float
//
// FP32 evaluation order shared with the CPU oracle (index selection is
// bit-identical; the reference's own FMA contraction is compiler-chosen):
//   BP: rho = fma(x, cos_i, fma(z, sin_i, bx/2)); index (int)rho
//   RP: X = fma(-t', sin, tx), Z = fma(t', cos, tz); index (int)Z, (int)X
//
// Kernels are instantiated for 1..kMaxFixedN reference projections so the
// per-projection cos/sin/weight tables sit in scalar registers; larger counts
// use the runtime-n instantiation (N = 0).
#include <algorithm>
#include <cmath>
#include <vector>

#include "bprp.h"
#include "context.h"

namespace mf {

namespace {

constexpr int kBpThreads = 512;
constexpr int kBpBlock = 256;  // voxels along x per wave task (4 per lane)
constexpr int kMaxFixedN = 16;

struct BpParams {
  const float* stack;  // [nz][by][bx]
  float* slab;         // [by][tb][2bx]
  int bx, by, tb, n, i0;
  float cnt_all;       // sum of the n weights in reference order
  float c[kMaxRefProj], s[kMaxRefProj], w[kMaxRefProj];
};

struct RpParams {
  const float* slab;
  float* out;
  int bx, by, tb, L;
  float c, s;
};

// Loads row y of the n reference projections into LDS as value*cos_i (the
// first product of the reference's v*cos*w), with a zero guard entry at
// index bx and a validity byte per entry.  Returns nonzero if any sample is
// invalid (< -1e10, the reference's mDIntProj test).
__device__ int load_sino_rows(const BpParams& p, int n, int y, float* sv,
                              unsigned char* sm, float* tc, float* ts) {
  const int bx = p.bx, rowp = bx + 1;
  const size_t plane = (size_t)bx * p.by;
  int bad = 0;
  for (int k = threadIdx.x; k < n * rowp; k += blockDim.x) {
    int i = k / rowp, x = k - i * rowp;
    float vc = 0.f;
    unsigned char ok = 0;
    if (x < bx) {
      float v = p.stack[(size_t)(p.i0 + i) * plane + (size_t)y * bx + x];
      if (v < -1e10f) {
        bad = 1;
      } else {
        vc = v * p.c[i];
        ok = 1;
      }
    }
    sv[k] = vc;
    sm[k] = ok;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    tc[i] = p.c[i];
    ts[i] = p.s[i];
  }
  return __syncthreads_or(bad);
}

// V voxels per lane with per-projection range and validity tests (the
// reference's general per-voxel loop, CalcReproj.cu:37-63).
template <int N, int V>
__device__ __forceinline__ void bp_general(const BpParams& p, const float* fx,
                                           float fz, const float* sv,
                                           const unsigned char* sm,
                                           int any_bad, float* out) {
  const int n = N ? N : p.n;
  const int bx = p.bx, rowp = bx + 1;
  const float cx = (float)bx * 0.5f, iend = (float)(bx - 1);
  float acc[V], cnt[V];
#pragma unroll
  for (int v = 0; v < V; v++) acc[v] = cnt[v] = 0.f;
#pragma unroll
  for (int i = 0; i < n; i++) {
    const float c = p.c[i], w = p.w[i];
    const float zsc = fmaf(fz, p.s[i], cx);
    const float* row = sv + i * rowp;
    const unsigned char* mrow = sm + i * rowp;
#pragma unroll
    for (int v = 0; v < V; v++) {
      const float r = fmaf(fx[v], c, zsc);
      const bool inr = (r >= 0.f) && (r <= iend);
      const int k = inr ? (int)r : bx;
      bool ok = inr;
      if (any_bad) ok = ok && mrow[k];
      acc[v] = fmaf(row[k], w, acc[v]);
      cnt[v] += ok ? w : 0.f;
    }
  }
#pragma unroll
  for (int v = 0; v < V; v++)
    out[v] = (cnt[v] < 0.001f) ? kInvalid : acc[v] / cnt[v];
}

template <int N>
__global__ __launch_bounds__(kBpThreads) void k_bp_slab(const BpParams p) {
  extern __shared__ float smem[];
  const int n = N ? N : p.n;
  const int y = blockIdx.x;
  const int bx = p.bx, rowp = bx + 1;
  float* sv = smem;
  float* tc = sv + n * rowp;
  float* ts = tc + kMaxRefProj;
  unsigned char* sm = (unsigned char*)(ts + kMaxRefProj);
  const int any_bad = load_sino_rows(p, n, y, sv, sm, tc, ts);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int vx = 2 * bx;
  const int nblk = (vx + kBpBlock - 1) / kBpBlock;
  const float cx = (float)bx * 0.5f, iend = (float)(bx - 1);
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  float* out_row = p.slab + (size_t)y * p.tb * vx;

  for (int task = wave; task < p.tb * nblk; task += nwaves) {
    const int iz = task / nblk, b = task - iz * nblk;
    const int ix0 = b * kBpBlock;
    const int ixl = min(ix0 + kBpBlock - 1, vx - 1);
    const float fz = (float)iz + 0.5f - htb;
    // Classify the block: every projection in range over it (fast), none in
    // range anywhere in it (empty), or mixed.  rho is monotone in x.
    bool inall = true, outall = true;
    if (lane < n) {
      const float zsc = fmaf(fz, ts[lane], cx);
      const float rf = fmaf((float)ix0 + 0.5f - hvx, tc[lane], zsc);
      const float rl = fmaf((float)ixl + 0.5f - hvx, tc[lane], zsc);
      inall = (rf >= 0.f) && (rl <= iend);
      outall = (rl < 0.f) || (rf > iend);
    }
    const bool fast = !any_bad && (__ballot(inall) == ~0ull);
    const bool empty = (__ballot(outall) == ~0ull);
    float* orow = out_row + (size_t)iz * vx;
    float fx[4], res[4];
    int ixs[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      ixs[v] = ix0 + lane + 64 * v;
      fx[v] = (float)min(ixs[v], ixl) + 0.5f - hvx;
    }
    if (empty) {
#pragma unroll
      for (int v = 0; v < 4; v++) res[v] = kInvalid;
    } else if (fast) {
      float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int i = 0; i < n; i++) {
        const float c = p.c[i], w = p.w[i];
        const float zsc = fmaf(fz, p.s[i], cx);
        const float* row = sv + i * rowp;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const float r = fmaf(fx[v], c, zsc);
          acc[v] = fmaf(row[(int)r], w, acc[v]);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; v++) res[v] = acc[v] / p.cnt_all;
    } else {
      bp_general<N, 4>(p, fx, fz, sv, sm, any_bad, res);
    }
#pragma unroll
    for (int v = 0; v < 4; v++)
      if (ixs[v] < vx) orow[ixs[v]] = res[v];
  }
}

// Sample range [t0, t1) of one ray that can fall inside the slab, from the
// monotone coordinate formulas with a 2-sample margin; the exact test is
// still applied to every sample in the range.
__device__ __forceinline__ void rp_sample_range(const RpParams& q, float tx,
                                                float tz, float z0, int* t0,
                                                int* t1) {
  const int vx = 2 * q.bx;
  float lo = -z0 - tz / q.c;
  float hi = -z0 + ((float)(q.tb - 1) - tz) / q.c;
  if (q.s > 1e-6f) {
    lo = fmaxf(lo, -z0 + (tx - (float)(vx - 1)) / q.s);
    hi = fminf(hi, -z0 + tx / q.s);
  } else if (q.s < -1e-6f) {
    lo = fmaxf(lo, -z0 + tx / q.s);
    hi = fminf(hi, -z0 + (tx - (float)(vx - 1)) / q.s);
  }
  int a = (int)floorf(lo) - 2, b = (int)ceilf(hi) + 3;
  *t0 = max(a, 0);
  *t1 = min(b, q.L);
}

// Ray-sum of the materialised slab; one lane per detector position u.  At a
// fixed sample index every ray reads (nearly) the same slab row, so the lanes
// of a wave gather along one row (stride 1/cos voxels).
__global__ __launch_bounds__(256) void k_rp_slab(const RpParams q) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (u >= q.bx) return;
  const int vx = 2 * q.bx;
  const float* vol = q.slab + (size_t)y * q.tb * vx;
  const float xp = (float)u + 0.5f - (float)q.bx * 0.5f;
  const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
  const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
  const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  int t0, t1;
  rp_sample_range(q, tx, tz, z0, &t0, &t1);
  double sum = 0.0;
  int cnt = 0;
#pragma unroll 4
  for (int t = t0; t < t1; t++) {
    const float fz = (float)t + z0;
    const float X = fmaf(-fz, q.s, tx);
    const float Z = fmaf(fz, q.c, tz);
    const bool inb = X >= 0.f && X < iex && Z >= 0.f && Z < iez;
    const int ix = inb ? (int)X : 0, iz = inb ? (int)Z : 0;
    const float v = vol[(size_t)vx * iz + ix];
    const bool ok = inb && (v >= -1e10f);
    sum += ok ? (double)v : 0.0;
    cnt += ok ? 1 : 0;
  }
  q.out[(size_t)y * q.bx + u] =
      ((float)cnt < 0.8f * (float)q.L) ? kInvalid : (float)(sum / (double)cnt);
}

// Fused form: each ray sample evaluates the backprojection of the voxel it
// falls in directly from the LDS-cached sinogram rows, so no slab reaches
// HBM.  The voxel value uses the same operations as k_bp_slab's general
// path, hence the output equals the materialised path bit for bit.
template <int N>
__global__ __launch_bounds__(256) void k_bprp_fused(const BpParams p,
                                                    const RpParams q) {
  extern __shared__ float smem[];
  const int n = N ? N : p.n;
  const int y = blockIdx.y;
  const int bx = p.bx, rowp = bx + 1;
  float* sv = smem;
  float* tc = sv + n * rowp;
  float* ts = tc + kMaxRefProj;
  unsigned char* sm = (unsigned char*)(ts + kMaxRefProj);
  const int any_bad = load_sino_rows(p, n, y, sv, sm, tc, ts);
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= bx) return;
  const int vx = 2 * bx;
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const float xp = (float)u + 0.5f - (float)bx * 0.5f;
  const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
  const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
  const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  int t0, t1;
  rp_sample_range(q, tx, tz, z0, &t0, &t1);
  double sum = 0.0;
  int cnt = 0;
  for (int t = t0; t < t1; t++) {
    const float fz = (float)t + z0;
    const float X = fmaf(-fz, q.s, tx);
    const float Z = fmaf(fz, q.c, tz);
    const bool inb = X >= 0.f && X < iex && Z >= 0.f && Z < iez;
    const float vfx = (float)(inb ? (int)X : 0) + 0.5f - hvx;
    const float vfz = (float)(inb ? (int)Z : 0) + 0.5f - htb;
    float v;
    bp_general<N, 1>(p, &vfx, vfz, sv, sm, any_bad, &v);
    const bool ok = inb && (v >= -1e10f);
    sum += ok ? (double)v : 0.0;
    cnt += ok ? 1 : 0;
  }
  q.out[(size_t)y * bx + u] =
      ((float)cnt < 0.8f * (float)q.L) ? kInvalid : (float)(sum / (double)cnt);
}

__global__ void k_fill(float* p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

size_t bp_lds_bytes(int n, int bx) {
  size_t rowp = (size_t)bx + 1;
  return (size_t)n * rowp * 4 + 2 * kMaxRefProj * 4 + (size_t)n * rowp;
}

typedef void (*BpKernel)(BpParams);
typedef void (*FusedKernel)(BpParams, RpParams);

template <int N>
struct KernelTable {
  static void fill(BpKernel* bp, FusedKernel* fu) {
    bp[N] = k_bp_slab<N>;
    fu[N] = k_bprp_fused<N>;
    KernelTable<N - 1>::fill(bp, fu);
  }
};
template <>
struct KernelTable<0> {
  static void fill(BpKernel* bp, FusedKernel* fu) {
    bp[0] = k_bp_slab<0>;
    fu[0] = k_bprp_fused<0>;
  }
};

struct Kernels {
  BpKernel bp[kMaxFixedN + 1];
  FusedKernel fu[kMaxFixedN + 1];
  Kernels() { KernelTable<kMaxFixedN>::fill(bp, fu); }
};

const Kernels& kernels() {
  static Kernels k;
  return k;
}

}  // namespace

void find_proj_range(const float* tilts, int nz, const unsigned char* skip,
                     int target, int nproj, int range[2]) {
  // CalcReproj.cu:281-334, float/double semantics kept.
  const float kRange = 20.5f, kStretch = 1.2f;
  const float tilt = tilts[target];
  const float fc = (float)std::cos(tilt * kD2R);
  auto accept = [&](int i) {
    if (skip[i]) return false;
    float d = tilt - tilts[i];
    if (std::fabs(d) > kRange) return false;
    float st = (float)(std::cos(tilts[i] * kD2R) / fc);
    return !(st > kStretch);
  };
  int lo = -1, hi = -1;
  for (int i = 0; i < nz; i++)
    if (accept(i)) {
      lo = i;
      break;
    }
  if (lo >= 0)
    for (int i = lo; i < nz; i++)
      if (accept(i)) hi = i;
  if (lo >= 0 && hi - lo > nproj - 1) {
    // GetFrameIdxFromTilt(0) (Util.cpp:8-20): first minimum of |a_i|.
    int z0 = 0;
    float m = std::fabs(tilts[0]);
    for (int i = 1; i < nz; i++) {
      float d = std::fabs(tilts[i]);
      if (d < m) {
        m = d;
        z0 = i;
      }
    }
    if (target < z0)
      hi = lo + nproj - 1;
    else
      lo = hi - nproj + 1;
  }
  range[0] = lo;
  range[1] = hi;
}

void reproject_target(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                      int bx, int by, int nz, int tb, const float* tilts,
                      const unsigned char* skip, int target, int nproj,
                      int mode, float* d_reproj, float* d_slab) {
  MF_REQUIRE(bx > 1 && by > 0 && nz > 0 && tb > 1, "bad reprojection shape");
  MF_REQUIRE(target >= 0 && target < nz, "target out of range");
  MF_REQUIRE(nproj > 0, "nproj must be > 0");
  int range[2];
  find_proj_range(tilts, nz, skip, target, nproj, range);
  const int n = range[0] >= 0 ? range[1] - range[0] + 1 : 0;
  MF_REQUIRE(n <= kMaxRefProj, "more reference projections than supported");

  BpParams bp{};
  bp.stack = d_stack;
  bp.bx = bx;
  bp.by = by;
  bp.tb = tb;
  bp.n = n;
  bp.i0 = std::max(range[0], 0);
  const float pa = tilts[target];
  float cnt = 0.f;
  for (int k = 0; k < n; k++) {
    const float a = tilts[range[0] + k];
    bp.w[k] = hcos((pa - a) * kBPD2R);
    bp.c[k] = hcos(a * kBPD2R);
    bp.s[k] = hsin(a * kBPD2R);
    cnt += bp.w[k];
  }
  bp.cnt_all = cnt;

  RpParams rp{};
  rp.out = d_reproj;
  rp.bx = bx;
  rp.by = by;
  rp.tb = tb;
  rp.c = (float)std::cos(kD2R * pa);
  rp.s = (float)std::sin(kD2R * pa);
  rp.L = (int)((float)tb / rp.c + 0.5f);

  const size_t lds = bp_lds_bytes(n, bx);
  MF_REQUIRE(lds <= 160 * 1024, "sinogram rows exceed LDS");
  const int kid = n <= kMaxFixedN ? n : 0;
  if (mode == 1) {
    if (n == 0) {
      hipLaunchKernelGGL(k_fill, dim3(cdiv((long long)bx * by, 256)),
                         dim3(256), 0, st, d_reproj, (size_t)bx * by,
                         kInvalid);
      MF_HIP(hipGetLastError());
      return;
    }
    FusedKernel k = kernels().fu[kid];
    dim3 grid(cdiv(bx, 256), by);
    if (lds > 64 * 1024)
      MF_HIP(hipFuncSetAttribute((const void*)k,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    TimedLaunch tl(ctx->timer, "bprp_fused", st);
    hipLaunchKernelGGL(k, grid, dim3(256), lds, st, bp, rp);
    MF_HIP(hipGetLastError());
    return;
  }
  const size_t slab_elems = (size_t)by * tb * 2 * bx;
  float* slab = d_slab;
  if (!slab) slab = (float*)ctx->slab_ws.ensure(ctx, slab_elems * 4);
  bp.slab = slab;
  rp.slab = slab;
  if (n == 0) {
    hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, st, slab, slab_elems,
                       kInvalid);
  } else {
    BpKernel k = kernels().bp[kid];
    if (lds > 64 * 1024)
      MF_HIP(hipFuncSetAttribute((const void*)k,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
    TimedLaunch tl(ctx->timer, "bp_slab", st);
    hipLaunchKernelGGL(k, dim3(by), dim3(kBpThreads), lds, st, bp);
  }
  MF_HIP(hipGetLastError());
  {
    TimedLaunch tl(ctx->timer, "rp_slab", st);
    hipLaunchKernelGGL(k_rp_slab, dim3(cdiv(bx, 256), by), dim3(256), 0, st,
                       rp);
  }
  MF_HIP(hipGetLastError());
}

}  // namespace mf
