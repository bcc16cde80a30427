// Backprojection into xz slabs and ray-sum reprojection for projection
// matching: the roofline-graded kernels of the alignment path.
//
// Reference: CalcReprojFBP (method/CalcReproj.cu).  The reference launches one
// backprojection and one reprojection per image row (CalcReproj.cu:190-194,
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
//   RP: X = fma(-t', sin, tx), Z = fma(t', cos, tz); index (int)Z, (int)X;
//       the ray sum is an FP32 sum of the valid samples in increasing t and
//       the output sum / count (the reference's FP32 tree order over 512
//       threads per ray, CalcReproj.cu:68-124, is one admissible order of
//       the same sum; a fixed sequential order makes every kernel form and
//       the oracle agree bit for bit).
//
// The fused kernel is instantiated for 1..kMaxFixedN reference projections so
// the per-projection cos/sin/weight tables sit in scalar registers; larger
// counts use the runtime-n instantiation (N = 0).
#include <algorithm>
#include <map>
#include <cmath>
#include <type_traits>
#include <vector>

#include "bprp.h"
#include "context.h"

namespace mf {

namespace {

constexpr int kBpThreads = 512;
constexpr int kBpBlock = 256;  // voxels along x per wave block (4 per lane)
// The R-row slab's block (2 voxels per lane): a block is classified as a
// whole (every projection in range, none, or mixed), and about a fifth of
// the 128-voxel blocks of a 1024-column slab are mixed against a third of
// the 256-voxel ones.
constexpr int kBpBlockR = 128;
constexpr int kFuThreads = 512;  // rays per fused work unit
constexpr int kMaxFixedN = 16;

struct BpParams {
  const float* stack;  // [nz][by][bx]
  float* slab;         // [by][tb][2bx]
  int bx, by, tb, n, i0;
  float cnt_all;       // sum of the n weights in reference order
  float rcp_all;       // 1 / cnt_all, correctly rounded
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
// index bx.  A sample the reference skips (< -1e10, its mDIntProj test) is
// stored as -inf, which no valid finite sample times cos can equal.  Returns
// nonzero (in every thread) if any sample of the rows is skipped.
__device__ int load_sino_rows(const BpParams& p, int n, int y, float* sv) {
  const int bx = p.bx, rowp = bx + 1;
  const size_t plane = (size_t)bx * p.by;
  const float* src = p.stack + (size_t)p.i0 * plane + (size_t)y * bx;
  int bad = 0;
  for (int i = 0; i < n; i++) {
    const float c = p.c[i];
    for (int x = threadIdx.x; x < rowp; x += blockDim.x) {
      float vc = 0.f;
      if (x < bx) {
        const float v = src[(size_t)i * plane + x];
        if (v < -1e10f) {
          bad = 1;
          vc = -INFINITY;
        } else {
          vc = v * c;
        }
      }
      sv[i * rowp + x] = vc;
    }
  }
  return __syncthreads_or(bad);
}

// Quotient a/d from d's correctly rounded reciprocal: q = a*r, then one
// residual step.  With r = RN(1/d) this returns RN(a/d) for operands away
// from overflow and the subnormal range (Markstein's division theorem).
__device__ __forceinline__ float div_by_rcp(float a, float d, float r) {
  const float q = a * r;
  return fmaf(fmaf(-q, d, a), r, q);
}

// Slab voxels are written once and read back by the next kernel only after
// more data than the last-level cache holds, so they are stored as streaming
// (non-temporal) writes: no dirty slab lines remain in the caches to be
// written back while the reprojection reads.
__device__ __forceinline__ void stream_store(float* at, float v) {
  __builtin_nontemporal_store(v, at);
}
__device__ __forceinline__ void stream_store(float4* at, const float4& v) {
  typedef float Nat __attribute__((ext_vector_type(4)));
  const Nat n = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(n, reinterpret_cast<Nat*>(at));
}
__device__ __forceinline__ void stream_store(float2* at, const float2& v) {
  typedef float Nat __attribute__((ext_vector_type(2)));
  const Nat n = {v.x, v.y};
  __builtin_nontemporal_store(n, reinterpret_cast<Nat*>(at));
}

// A uniform value copied into a vector register, so that a select between it
// and zero takes no scalar operand besides its condition mask.
__device__ __forceinline__ float in_vgpr(float x) {
  float y;
  asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "s"(x));
  return y;
}

// RN(1/d) for d in [2^-10, 2^4] (weight sums): the division sequence of
// 1.0f / d without its range scaling and special-case fix-up, which are the
// identity on that range -- a hardware reciprocal estimate refined by one
// Newton step, then two residual corrections of the quotient.
__device__ __forceinline__ float rcp_rn(float d) {
  const float r0 = __builtin_amdgcn_rcpf(d);
  const float r = fmaf(fmaf(-d, r0, 1.0f), r0, r0);
  const float q = fmaf(fmaf(-d, r, 1.0f), r, r);
  return fmaf(fmaf(-d, q, 1.0f), r, q);
}

// V voxels per lane with per-projection range and validity tests (the
// reference's general per-voxel loop, CalcReproj.cu:37-63).
template <int N, int V>
__device__ __forceinline__ void bp_general(const BpParams& p, const float* fx,
                                           float fz, const float* sv,
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
#pragma unroll
    for (int v = 0; v < V; v++) {
      const float r = fmaf(fx[v], c, zsc);
      const bool inr = (r >= 0.f) && (r <= iend);
      float val = row[inr ? (int)r : bx];
      bool ok = inr;
      if (any_bad) {
        ok = ok && (val != -INFINITY);
        val = ok ? val : 0.f;
      }
      acc[v] = fmaf(val, w, acc[v]);
      cnt[v] += ok ? w : 0.f;
    }
  }
#pragma unroll
  for (int v = 0; v < V; v++)
    out[v] = (cnt[v] < 0.001f) ? kInvalid : acc[v] / cnt[v];
}

// Per-projection constants of one target in lane registers: lane i < n
// holds cos_i, sin_i and w_i (lanes >= n repeat projection 0; their bits are
// masked out of every ballot).
struct LaneProj {
  float c, s, w;
  unsigned long long mask;  // bits 0..n-1
  __device__ LaneProj(const BpParams& p, int n, int lane) {
    const int i = lane < n ? lane : 0;
    c = p.c[i];
    s = p.s[i];
    w = p.w[i];
    mask = n >= 64 ? ~0ull : ((1ull << n) - 1);
  }
};

__device__ __forceinline__ float lane_f(float v, int i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
}

// Persistent backprojection.  Workgroup g owns the contiguous range
// [g*per, (g+1)*per) of slab rows (y, iz) in y-major order and loads the
// sinogram rows of each y it meets into LDS once; its waves take whole slab
// rows (iz) and walk them in blocks of kBpBlock voxels along x.  For each
// block and projection, rho's monotonicity in x classifies the projection as
// in range over the whole block, out of range over the whole block, or
// mixed, from the block's two end voxels:
//   - no projection in range: every voxel has weight 0 -> invalid;
//   - all in range, no skipped samples: no per-voxel tests, and the constant
//     weight sum divides through its correctly rounded reciprocal;
//   - otherwise: out-of-range projections are skipped (the reference's
//     'continue'), in-range ones need no test, mixed ones test each voxel.
// Every voxel accumulates the same terms in the same order as the reference
// loop (CalcReproj.cu:37-63), so the slab is bit-identical to it.
__global__ __launch_bounds__(kBpThreads) void k_bp_slab(const BpParams p,
                                                        int per) {
  extern __shared__ float smem[];
  constexpr int V = kBpBlock / 64;
  const int n = p.n;
  const int bx = p.bx, rowp = bx + 1;
  float* sv = smem;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int vx = 2 * bx;
  const float cx = (float)bx * 0.5f, iend = (float)(bx - 1);
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const LaneProj lp(p, n, lane);
  const bool all_fast_ok = p.cnt_all >= 0.001f;
  const int total = p.by * p.tb;
  const int qa = blockIdx.x * per, qb = min(total, qa + per);

  for (int q0 = qa; q0 < qb;) {
    const int y = q0 / p.tb;
    const int za = q0 - y * p.tb;
    const int zb = min(p.tb, za + (qb - q0));
    q0 += zb - za;
    __syncthreads();  // the previous row's waves are done with the LDS rows
    const int any_bad = load_sino_rows(p, n, y, sv);
    for (int iz = za + wave; iz < zb; iz += nwaves) {
      const float fz = (float)iz + 0.5f - htb;
      const float lz = fmaf(fz, lp.s, cx);  // lane i: zsc of projection i
      float* orow = p.slab + ((size_t)y * p.tb + iz) * vx + lane;
      for (int ix0 = 0; ix0 < vx; ix0 += kBpBlock) {
        const int ixl = min(ix0 + kBpBlock - 1, vx - 1);
        const float rf = fmaf((float)ix0 + 0.5f - hvx, lp.c, lz);
        const float rl = fmaf((float)ixl + 0.5f - hvx, lp.c, lz);
        const unsigned long long m_in =
            __ballot(rf >= 0.f && rl <= iend) & lp.mask;
        const unsigned long long m_out =
            __ballot(rl < 0.f || rf > iend) & lp.mask;
        float fx[V], res[V];
#pragma unroll
        for (int v = 0; v < V; v++)
          fx[v] = (float)min(ix0 + lane + 64 * v, ixl) + 0.5f - hvx;
        if (m_out == lp.mask) {
#pragma unroll
          for (int v = 0; v < V; v++) res[v] = kInvalid;
        } else if (m_in == lp.mask && !any_bad && all_fast_ok) {
          float acc[V];
#pragma unroll
          for (int v = 0; v < V; v++) acc[v] = 0.f;
#pragma unroll 1
          for (int i = 0; i < n; i++) {
            const float c = lane_f(lp.c, i), w = lane_f(lp.w, i);
            const float zsc = lane_f(lz, i);
            const float* row = sv + i * rowp;
#pragma unroll
            for (int v = 0; v < V; v++)
              acc[v] = fmaf(row[(int)fmaf(fx[v], c, zsc)], w, acc[v]);
          }
#pragma unroll
          for (int v = 0; v < V; v++)
            res[v] = div_by_rcp(acc[v], p.cnt_all, p.rcp_all);
        } else {
          float acc[V], cnt[V];
#pragma unroll
          for (int v = 0; v < V; v++) acc[v] = cnt[v] = 0.f;
          const unsigned long long m_fast = any_bad ? 0ull : m_in;
#pragma unroll 1
          for (int i = 0; i < n; i++) {
            if ((m_out >> i) & 1ull) continue;
            const float c = lane_f(lp.c, i), w = lane_f(lp.w, i);
            const float zsc = lane_f(lz, i);
            const float* row = sv + i * rowp;
            if ((m_fast >> i) & 1ull) {
#pragma unroll
              for (int v = 0; v < V; v++) {
                acc[v] = fmaf(row[(int)fmaf(fx[v], c, zsc)], w, acc[v]);
                cnt[v] += w;
              }
            } else {
#pragma unroll
              for (int v = 0; v < V; v++) {
                const float r = fmaf(fx[v], c, zsc);
                const bool inr = (r >= 0.f) && (r <= iend);
                float val = row[inr ? (int)r : bx];
                bool ok = inr;
                if (any_bad) {
                  ok = ok && (val != -INFINITY);
                  val = ok ? val : 0.f;
                }
                acc[v] = fmaf(val, w, acc[v]);
                cnt[v] += ok ? w : 0.f;
              }
            }
          }
#pragma unroll
          for (int v = 0; v < V; v++)
            res[v] = (cnt[v] < 0.001f) ? kInvalid : acc[v] / cnt[v];
        }
        if (ix0 + kBpBlock <= vx) {
#pragma unroll
          for (int v = 0; v < V; v++) stream_store(orow + ix0 + 64 * v, res[v]);
        } else {
#pragma unroll
          for (int v = 0; v < V; v++)
            if (ix0 + lane + 64 * v < vx)
              stream_store(orow + ix0 + 64 * v, res[v]);
        }
      }
    }
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
  float sum = 0.0f;
  int cnt = 0;
#pragma unroll 4
  for (int t = t0; t < t1; t++) {
    const float fz = (float)t + z0;
    const float X = fmaf(-fz, q.s, tx);
    const float Z = fmaf(fz, q.c, tz);
    const bool inb = (X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez);
    const int ix = inb ? (int)X : 0, iz = inb ? (int)Z : 0;
    const float v = vol[(size_t)vx * iz + ix];
    const bool ok = inb && (v >= -1e10f);
    sum += ok ? v : 0.0f;
    cnt += ok ? 1 : 0;
  }
  q.out[(size_t)y * q.bx + u] =
      ((float)cnt < 0.8f * (float)q.L) ? kInvalid : sum / (float)cnt;
}

__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// Sample interval [ta, tb) of one ray on which every sample lies inside the
// slab and every reference projection is in range for the sample's voxel,
// with margins that cover the voxel snapping (|rho - g| <= (|c|+|s|)/2) and
// the FP32 rounding of the sample coordinates (~1e-3 voxel).  Along the ray
// tau = t + z0 the exact coordinates are X = tx - tau s, Z = tz + tau c, and
// rho_i's continuous counterpart g_i is linear in tau, so each condition is
// a tau interval.  Evaluated in FP64; the caller intersects it with the
// ray's own range.
__device__ void fast_interval(const BpParams& p, const RpParams& q, int n,
                              float tx, float tz, float z0, int* ta, int* tb,
                              int* ia = nullptr, int* ib = nullptr) {
  const double vx = 2.0 * p.bx, hvx = p.bx, htb = 0.5 * p.tb;
  const double s = q.s, c = q.c;
  double lo = -1e30, hi = 1e30;
  auto clip = [&](double a, double b, double mn, double mx) {
    // a + b*tau in [mn, mx]
    if (fabs(b) < 1e-12) {
      if (a < mn || a > mx) {
        lo = 1e30;
        hi = -1e30;
      }
      return;
    }
    double t1 = (mn - a) / b, t2 = (mx - a) / b;
    lo = fmax(lo, fmin(t1, t2));
    hi = fmin(hi, fmax(t1, t2));
  };
  clip(tx, -s, 2.0, vx - 3.0);             // X inside the slab
  clip(tz, c, 2.0, (double)p.tb - 3.0);    // Z inside the slab
  if (ia) {
    // The samples inside the slab (same margins), whatever the projections.
    if (!(lo <= hi)) {
      *ia = *ib = 0;
    } else {
      *ia = (int)fmax(-1e9, ceil(lo - (double)z0)) + 1;
      *ib = (int)fmin(1e9, floor(hi - (double)z0));
    }
  }
#pragma unroll 1
  for (int i = 0; i < n; i++) {
    const double ci = p.c[i], si = p.s[i];
    const double a = (tx - hvx) * ci + (tz - htb) * si + 0.5 * p.bx;
    clip(a, -s * ci + c * si, 2.0, (double)p.bx - 3.0);
  }
  if (!(lo <= hi)) {
    *ta = *tb = 0;
    return;
  }
  const double zf = z0;
  *ta = (int)fmax(-1e9, ceil(lo - zf)) + 1;
  *tb = (int)fmin(1e9, floor(hi - zf));
}

// Fused form: each ray sample evaluates the backprojection of the voxel it
// falls in directly from the LDS-cached sinogram rows, so no slab reaches
// HBM.  Persistent: workgroup g owns the units [g*per, (g+1)*per) of
// (image row, chunk of kFuThreads rays), reloading the sinogram rows when
// the row changes.  Each ray is split at its wave's common fast interval:
// outside it the sample takes the reference's tests (slab bounds, range,
// skipped samples, empty voxel); inside it none apply.  Either way the voxel
// value is computed with the same operations and order as k_bp_slab and the
// ray sum visits the samples in increasing t, so the output equals the
// materialised path bit for bit.
template <int N>
__global__ __launch_bounds__(kFuThreads) void k_bprp_fused(const BpParams p,
                                                           const RpParams q,
                                                           int per) {
  extern __shared__ float smem[];
  const int n = N ? N : p.n;
  const int bx = p.bx, rowp = bx + 1;
  float* sv = smem;
  const int vx = 2 * bx;
  const float cx = (float)bx * 0.5f;
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  const float iend = (float)(bx - 1);
  const LaneProj lp(p, n, threadIdx.x & 63);
  // (float)k + 0.5f - h == (float)k + (0.5f - h) exactly for the integer
  // voxel indices and half-sizes here (all terms are small half-integers).
  const float ox = 0.5f - hvx, oz = 0.5f - htb;
  const bool fast_ok = p.cnt_all >= 0.001f;
  const int nch = cdiv(bx, kFuThreads);
  const int units = p.by * nch;
  const int ua = blockIdx.x * per, ub = min(units, ua + per);
  int ycur = -1, any_bad = 0;
  for (int un = ua; un < ub; un++) {
    const int y = un / nch;
    if (y != ycur) {
      __syncthreads();  // all waves are done with the previous row
      any_bad = load_sino_rows(p, n, y, sv);
      ycur = y;
    }
    const int u = (un - y * nch) * kFuThreads + threadIdx.x;
    const bool live = u < bx;
    const int uu = live ? u : bx - 1;  // idle lanes shadow a real ray
    const float xp = (float)uu + 0.5f - (float)bx * 0.5f;
    const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
    const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
    const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
    int t0, t1;
    rp_sample_range(q, tx, tz, z0, &t0, &t1);
    int fa, fb;
    fast_interval(p, q, p.n, tx, tz, z0, &fa, &fb);
    fa = max(fa, t0);
    fb = min(fb, t1);
    // The wave's common fast interval.
    const int ta = wave_max(fa), tb = wave_min(fb);
    const bool has_fast = fast_ok && !any_bad && ta < tb;
    const int a = has_fast ? ta : t1, b = has_fast ? tb : t1;
    float sum = 0.0f;
    int cnt = 0;
    // General samples: [t0, a) and [b, t1); per-projection constants from
    // lane registers, not unrolled (a minority of the samples).
    auto general = [&](int from, int to) {
      for (int t = from; t < to; t++) {
        const float fz = (float)t + z0;
        const float X = fmaf(-fz, q.s, tx);
        const float Z = fmaf(fz, q.c, tz);
        const bool inb = (X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez);
        const float vfx = (float)(inb ? (int)X : 0) + ox;
        const float vfz = (float)(inb ? (int)Z : 0) + oz;
        float acc = 0.f, wsum = 0.f;
#pragma unroll 1
        for (int i = 0; i < n; i++) {
          const float ci = lane_f(lp.c, i), si = lane_f(lp.s, i);
          const float wi = lane_f(lp.w, i);
          const float r = fmaf(vfx, ci, fmaf(vfz, si, cx));
          const bool inr = (r >= 0.f) && (r <= iend);
          float val = sv[i * rowp + (inr ? (int)r : bx)];
          bool okr = inr;
          if (any_bad) {
            okr = okr && (val != -INFINITY);
            val = okr ? val : 0.f;
          }
          acc = fmaf(val, wi, acc);
          wsum += okr ? wi : 0.f;
        }
        const float v = (wsum < 0.001f) ? kInvalid : acc / wsum;
        const bool ok = inb && (v >= -1e10f);
        sum += ok ? v : 0.0f;
        cnt += ok ? 1 : 0;
      }
    };
    general(t0, a);
    if (has_fast) {
      for (int t = a; t < b; t++) {
        const float fz = (float)t + z0;
        const float vfx = (float)(int)fmaf(-fz, q.s, tx) + ox;
        const float vfz = (float)(int)fmaf(fz, q.c, tz) + oz;
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < n; i++) {
          const float zsc = fmaf(vfz, p.s[i], cx);
          acc = fmaf(sv[i * rowp + (int)fmaf(vfx, p.c[i], zsc)], p.w[i], acc);
        }
        sum += div_by_rcp(acc, p.cnt_all, p.rcp_all);
      }
      cnt += b - a;
    }
    general(b, t1);
    if (live)
      q.out[(size_t)y * bx + u] = ((float)cnt < 0.8f * (float)q.L)
                                      ? kInvalid
                                      : sum / (float)cnt;
  }
}

// Table-driven fused form for n <= kTabMaxN reference projections.
//
// Two per-target tables are built in LDS once per workgroup:
//   ztab[iz][i] = fma(iz + 0.5 - tb/2, sin_i, bx/2), the z half of every
//     backprojection coordinate (shared by all voxels of slab row iz), so a
//     lookup costs fma, cvt, shift, read, fma;
//   wtab[m] = {W(m), RN(1/W(m))}, W(m) the reference-order FP32 sum of the
//     weights of the projections whose bits are set in m, so a sample that
//     needs the range tests accumulates a validity mask with one add-carry
//     per projection and reads its weight sum and reciprocal at the end.
// Samples inside the wave's common fast interval take no tests; the others
// test every projection without branches.  Every voxel value, and the ray
// sum, is computed with the same operations in the same order as the
// reference-order loop (CalcReproj.cu:37-63, 97-124), so the output equals
// k_bprp_fused and the materialised path bit for bit.
constexpr int kTabMaxN = 10;

__host__ __device__ constexpr int ztab_stride(int n) { return (n + 3) & ~3; }

size_t tab_lds_bytes(int n, int bx, int tb) {
  return (size_t)tb * ztab_stride(n) * 4 + ((size_t)8 << n) +
         (size_t)n * ((size_t)bx + 1) * 4;
}

template <int N>
__device__ __forceinline__ void load_z(const float* ztab, int iz, float* z) {
  constexpr int ZS = ztab_stride(N);
  const float4* r =
      reinterpret_cast<const float4*>(ztab) + __umul24((unsigned)iz, ZS / 4);
#pragma unroll
  for (int k = 0; k < ZS / 4; k++) {
    const float4 v = r[k];
    z[4 * k] = v.x;
    z[4 * k + 1] = v.y;
    z[4 * k + 2] = v.z;
    z[4 * k + 3] = v.w;
  }
}

template <int N>
__global__ __launch_bounds__(kFuThreads) void k_bprp_tab(const BpParams p,
                                                         const RpParams q,
                                                         int per) {
  extern __shared__ float4 smem4[];
  constexpr int ZS = ztab_stride(N);
  float* ztab = reinterpret_cast<float*>(smem4);
  float2* wtab = reinterpret_cast<float2*>(ztab + p.tb * ZS);
  float* sv = reinterpret_cast<float*>(wtab + (1 << N));
  const int bx = p.bx, rowp = bx + 1;
  const int vx = 2 * bx;
  const float cx = (float)bx * 0.5f;
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  const unsigned iend_bits = __float_as_uint((float)(bx - 1));
  const float ox = 0.5f - hvx, oz = 0.5f - htb;
  const bool fast_ok = p.cnt_all >= 0.001f;

  for (int k = threadIdx.x; k < p.tb * ZS; k += blockDim.x) {
    const int iz = k / ZS, i = k - iz * ZS;
    float v = 0.f;
#pragma unroll
    for (int j = 0; j < N; j++)
      if (i == j) v = fmaf((float)iz + oz, p.s[j], cx);
    ztab[k] = v;
  }
  for (int m = threadIdx.x; m < (1 << N); m += blockDim.x) {
    float w = 0.f;
#pragma unroll
    for (int i = 0; i < N; i++)
      if ((m >> (N - 1 - i)) & 1) w += p.w[i];
    wtab[m] = make_float2(w, 1.0f / w);
  }

  const int nch = cdiv(bx, kFuThreads);
  const int units = p.by * nch;
  const int ua = blockIdx.x * per, ub = min(units, ua + per);
  int ycur = -1, any_bad = 0;
  for (int un = ua; un < ub; un++) {
    const int y = un / nch;
    if (y != ycur) {
      __syncthreads();  // tables built; all waves done with the previous row
      any_bad = load_sino_rows(p, N, y, sv);
      ycur = y;
    }
    const int u = (un - y * nch) * kFuThreads + threadIdx.x;
    const bool live = u < bx;
    const int uu = live ? u : bx - 1;  // idle lanes shadow a real ray
    const float xp = (float)uu + 0.5f - (float)bx * 0.5f;
    const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
    const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
    const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
    int t0, t1;
    rp_sample_range(q, tx, tz, z0, &t0, &t1);
    int fa, fb;
    fast_interval(p, q, N, tx, tz, z0, &fa, &fb);
    fa = max(fa, t0);
    fb = min(fb, t1);
    const int ta = wave_max(fa), tb = wave_min(fb);
    const bool has_fast = fast_ok && !any_bad && ta < tb;
    const int a = has_fast ? ta : t1, b = has_fast ? tb : t1;
    float sum = 0.0f;
    int cnt = 0;
    // Samples that take the reference's tests.  A skipped sinogram sample
    // (stored as -inf) clears its projection's validity bit when the row
    // holds one.
    auto general = [&](int from, int to, auto bad) {
      for (int t = from; t < to; t++) {
        const float fz = (float)t + z0;
        const float X = fmaf(-fz, q.s, tx);
        const float Z = fmaf(fz, q.c, tz);
        const bool inb = (X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez);
        const float vfx = (inb ? truncf(X) : 0.f) + ox;
        float zs[ZS];
        load_z<N>(ztab, inb ? (int)Z : 0, zs);
        float acc = 0.f;
        unsigned m = 0;
#pragma unroll
        for (int i = 0; i < N; i++) {
          const float r = fmaf(vfx, p.c[i], zs[i]);
          // 0 <= r <= bx-1 as one unsigned compare of the bit patterns.
          // r is finite and never -0: a round-to-nearest fma returns -0 only
          // when its product and addend are both -0, and zs[i] is not -0
          // because its own addend, bx/2, is positive.
          bool ok = __float_as_uint(r) <= iend_bits;
          float val = sv[i * rowp + (ok ? (int)r : bx)];
          if (decltype(bad)::value) {
            ok = ok && (val != -INFINITY);
            val = ok ? val : 0.f;
          }
          acc = fmaf(val, p.w[i], acc);
          m = 2 * m + (ok ? 1u : 0u);
        }
        const float2 wr = wtab[m];
        const float v =
            (wr.x < 0.001f) ? kInvalid : div_by_rcp(acc, wr.x, wr.y);
        const bool ok = inb && (v >= -1e10f);
        sum += ok ? v : 0.0f;
        cnt += ok ? 1 : 0;
      }
    };
    if (any_bad) {
      general(t0, t1, std::true_type{});
    } else {
      general(t0, a, std::false_type{});
      if (has_fast) {
        for (int t = a; t < b; t++) {
          const float fz = (float)t + z0;
          const float vfx = truncf(fmaf(-fz, q.s, tx)) + ox;
          float zs[ZS];
          load_z<N>(ztab, (int)fmaf(fz, q.c, tz), zs);
          float acc = 0.f;
#pragma unroll
          for (int i = 0; i < N; i++)
            acc = fmaf(sv[i * rowp + (int)fmaf(vfx, p.c[i], zs[i])], p.w[i],
                       acc);
          sum += div_by_rcp(acc, p.cnt_all, p.rcp_all);
        }
        cnt += b - a;
      }
      general(b, t1, std::false_type{});
    }
    if (live)
      q.out[(size_t)y * bx + u] = ((float)cnt < 0.8f * (float)q.L)
                                      ? kInvalid
                                      : sum / (float)cnt;
  }
}

// Multi-row table-driven form.  The rays of image rows y0 .. y0+R-1 at one
// detector position follow the same path through the same voxels, and every
// voxel looks up the same sinogram positions; only the sinogram values
// differ.  A work unit therefore covers one ray chunk of a group of R rows:
// the sample coordinates, the z table row, each projection's index and range
// test are evaluated once, and the R rows' values come from a row-interleaved
// LDS copy of the sinogram rows (RowSlots) with one address per lookup.
// Each row's voxel values and ray sum still take exactly the single-row
// operations in the same order, so the output equals k_bprp_tab bit for bit.
// Workgroups of up to kRowsThreads rays keep enough waves resident under the
// larger LDS footprint.
constexpr int kRowsThreads = 1024;

// Sinogram slots of R rows: R = 2 keeps one float2 {row 0, row 1} per
// (projection, position).
template <int R>
struct RowSlots;

template <>
struct RowSlots<2> {
  __host__ __device__ static size_t floats(int n, int rowp) {
    return (size_t)n * rowp * 2;
  }
  __device__ static float* at(float* sv, int rowp, int i, int x, int r) {
    return sv + ((size_t)i * rowp + x) * 2 + r;
  }
  __device__ static void read(const float* sv, int rowp, int i, int idx,
                              float* v) {
    const float2 a =
        reinterpret_cast<const float2*>(sv)[(size_t)i * rowp + idx];
    v[0] = a.x;
    v[1] = a.y;
  }
};

size_t rows_lds_bytes(int n, int bx, int tb) {
  return (size_t)tb * ztab_stride(n) * 4 + ((size_t)8 << n) +
         RowSlots<2>::floats(n, bx + 1) * 4;
}

// Rows y[0..R) of the n reference projections into their slots, encoded
// like load_sino_rows.  Returns nonzero if any of the rows holds a skipped
// sample.
template <int R>
__device__ int load_sino_group(const BpParams& p, int n, const int* y,
                               float* sv) {
  const int bx = p.bx, rowp = bx + 1;
  const size_t plane = (size_t)bx * p.by;
  int bad = 0;
  for (int i = 0; i < n; i++) {
    const float c = p.c[i];
    const float* src = p.stack + (size_t)(p.i0 + i) * plane;
    for (int x = threadIdx.x; x < rowp; x += blockDim.x) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        float vc = 0.f;
        if (x < bx) {
          const float v = src[(size_t)y[r] * bx + x];
          const bool b = v < -1e10f;
          bad |= b ? 1 : 0;
          vc = b ? -INFINITY : v * c;
        }
        *RowSlots<R>::at(sv, rowp, i, x, r) = vc;
      }
    }
  }
  return __syncthreads_or(bad);
}

template <int N, int R>
__global__ __launch_bounds__(kRowsThreads) void k_bprp_rows(const BpParams p,
                                                            const RpParams q,
                                                            int per) {
  extern __shared__ float4 smem4[];
  constexpr int ZS = ztab_stride(N);
  float* ztab = reinterpret_cast<float*>(smem4);
  float2* wtab = reinterpret_cast<float2*>(ztab + p.tb * ZS);
  float* sv = reinterpret_cast<float*>(wtab + (1 << N));
  const int bx = p.bx, rowp = bx + 1;
  const int vx = 2 * bx;
  const float cx = (float)bx * 0.5f;
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  const unsigned iend_bits = __float_as_uint((float)(bx - 1));
  const float ox = 0.5f - hvx, oz = 0.5f - htb;
  const bool fast_ok = p.cnt_all >= 0.001f;

  for (int k = threadIdx.x; k < p.tb * ZS; k += blockDim.x) {
    const int iz = k / ZS, i = k - iz * ZS;
    float v = 0.f;
#pragma unroll
    for (int j = 0; j < N; j++)
      if (i == j) v = fmaf((float)iz + oz, p.s[j], cx);
    ztab[k] = v;
  }
  for (int m = threadIdx.x; m < (1 << N); m += blockDim.x) {
    float w = 0.f;
#pragma unroll
    for (int i = 0; i < N; i++)
      if ((m >> (N - 1 - i)) & 1) w += p.w[i];
    wtab[m] = make_float2(w, 1.0f / w);
  }

  const int nt = blockDim.x;
  const int nch = cdiv(bx, nt);
  const int units = cdiv(p.by, R) * nch;
  const int ua = blockIdx.x * per, ub = min(units, ua + per);
  int gcur = -1, any_bad = 0;
  int y[R];
  for (int un = ua; un < ub; un++) {
    const int gr = un / nch;
    if (gr != gcur) {
      __syncthreads();  // tables built; all waves done with the previous rows
      // Rows past the last repeat it (computed, not stored).
#pragma unroll
      for (int r = 0; r < R; r++) y[r] = min(R * gr + r, p.by - 1);
      any_bad = load_sino_group<R>(p, N, y, sv);
      gcur = gr;
    }
    const int u = (un - gr * nch) * nt + threadIdx.x;
    const bool live = u < bx;
    const int uu = live ? u : bx - 1;  // idle lanes shadow a real ray
    const float xp = (float)uu + 0.5f - (float)bx * 0.5f;
    const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
    const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
    const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
    int t0, t1;
    rp_sample_range(q, tx, tz, z0, &t0, &t1);
    int fa, fb;
    fast_interval(p, q, N, tx, tz, z0, &fa, &fb);
    fa = max(fa, t0);
    fb = min(fb, t1);
    const int ta = wave_max(fa), tb = wave_min(fb);
    const bool has_fast = fast_ok && !any_bad && ta < tb;
    const int a = has_fast ? ta : t1, b = has_fast ? tb : t1;
    float sum[R];
    int cnt[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      sum[r] = 0.0f;
      cnt[r] = 0;
    }
    // Samples that take the reference's tests.  With skipped samples in any
    // row each row keeps its own validity mask (a row without them never
    // matches the -inf test, so its mask is the shared one).
    auto general = [&](int from, int to, auto bad) {
      for (int t = from; t < to; t++) {
        const float fz = (float)t + z0;
        const float X = fmaf(-fz, q.s, tx);
        const float Z = fmaf(fz, q.c, tz);
        const bool inb = (X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez);
        const float vfx = (inb ? truncf(X) : 0.f) + ox;
        float zs[ZS];
        load_z<N>(ztab, inb ? (int)Z : 0, zs);
        float vals[N][R];
        unsigned okm = 0;  // bit i: projection i in range
#pragma unroll
        for (int i = 0; i < N; i++) {
          const float rr = fmaf(vfx, p.c[i], zs[i]);
          // The same unsigned range test as k_bprp_tab.
          const bool ok = __float_as_uint(rr) <= iend_bits;
          okm |= ok ? (1u << i) : 0u;
          RowSlots<R>::read(sv, rowp, i, ok ? (int)rr : bx, vals[i]);
        }
        float acc[R];
        unsigned m[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
          acc[r] = 0.f;
          m[r] = 0;
        }
#pragma unroll
        for (int i = 0; i < N; i++) {
          const bool ok = (okm >> i) & 1u;
#pragma unroll
          for (int r = 0; r < R; r++) {
            float val = vals[i][r];
            if (decltype(bad)::value) {
              const bool okr = ok && (val != -INFINITY);
              val = okr ? val : 0.f;
              m[r] = 2 * m[r] + (okr ? 1u : 0u);
            } else if (r == 0) {
              m[0] = 2 * m[0] + (ok ? 1u : 0u);
            }
            acc[r] = fmaf(val, p.w[i], acc[r]);
          }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float2 wr = wtab[decltype(bad)::value ? m[r] : m[0]];
          const float v =
              (wr.x < 0.001f) ? kInvalid : div_by_rcp(acc[r], wr.x, wr.y);
          const bool ok = inb && (v >= -1e10f);
          sum[r] += ok ? v : 0.0f;
          cnt[r] += ok ? 1 : 0;
        }
      }
    };
    if (any_bad) {
      general(t0, t1, std::true_type{});
    } else {
      general(t0, a, std::false_type{});
      if (has_fast) {
        for (int t = a; t < b; t++) {
          const float fz = (float)t + z0;
          const float vfx = truncf(fmaf(-fz, q.s, tx)) + ox;
          float zs[ZS];
          load_z<N>(ztab, (int)fmaf(fz, q.c, tz), zs);
          // All lookups of the sample are issued before the first use, so
          // their LDS latencies overlap.
          float vals[N][R];
#pragma unroll
          for (int i = 0; i < N; i++)
            RowSlots<R>::read(sv, rowp, i, (int)fmaf(vfx, p.c[i], zs[i]),
                              vals[i]);
#pragma unroll
          for (int r = 0; r < R; r++) {
            float acc = 0.f;
#pragma unroll
            for (int i = 0; i < N; i++) acc = fmaf(vals[i][r], p.w[i], acc);
            sum[r] += div_by_rcp(acc, p.cnt_all, p.rcp_all);
          }
        }
#pragma unroll
        for (int r = 0; r < R; r++) cnt[r] += b - a;
      }
      general(b, t1, std::false_type{});
    }
    if (live) {
      const float lim = 0.8f * (float)q.L;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (R * gr + r < p.by)
          q.out[(size_t)y[r] * bx + u] =
              ((float)cnt[r] < lim) ? kInvalid
                                    : sum[r] / (float)cnt[r];
    }
  }
}

// Table-free multi-row form (R = 4 or 2 image rows per work unit).  With four
// rows per unit at 1024 columns, or two rows at 1920-2048 columns, the
// sinogram slots alone fill LDS (one float4 or float2 {row 0..R-1} per
// projection and position, no guard entry), so nothing else is tabulated:
// the z half of each lookup coordinate is evaluated per sample with the z
// table's own fma, and an out-of-range projection reads the clamped slot and
// contributes with weight 0 (adding a zero product leaves the sum
// unchanged).  A tested sample shares its weight sum (in reference order),
// the sum's correctly rounded reciprocal and its count across the rows.  The
// result equals k_bprp_tab bit for bit.  At 1024 image rows, four rows per
// unit is also one unit per CU.
//
// The kernel holds no LDS beyond the slots, so there is no workgroup-wide
// flag for skipped samples (stored as -inf); each ray detects them itself.
// A lookup of -inf makes its sample's quotient NaN, whatever the weights,
// and every counted sample's quotient reaches its row's ray sum.  A ray with
// a NaN sum is evaluated again with per-row validity and weight sums (exact
// in every case), so the output stays bit-identical.
size_t free_lds_bytes(int r, int n, int bx) { return (size_t)n * bx * 4 * r; }
template <int R>
constexpr int kSlotShift = R == 4 ? 4 : 3;  // log2 of a slot's bytes

template <int R>
struct FreeSlot;
template <>
struct FreeSlot<4> {
  typedef float4 T;
  __device__ static T make(const float* v) {
    return make_float4(v[0], v[1], v[2], v[3]);
  }
  __device__ static void split(const T& s, float* v) {
    v[0] = s.x;
    v[1] = s.y;
    v[2] = s.z;
    v[3] = s.w;
  }
};
template <>
struct FreeSlot<2> {
  typedef float2 T;
  __device__ static T make(const float* v) { return make_float2(v[0], v[1]); }
  __device__ static void split(const T& s, float* v) {
    v[0] = s.x;
    v[1] = s.y;
  }
};

// The slot at LDS byte address 'at'.
template <int R>
__device__ __forceinline__ typename FreeSlot<R>::T lds_slot(unsigned at) {
  typedef float Nat __attribute__((ext_vector_type(R)));
  const Nat n = *(const __attribute__((address_space(3))) Nat*)(size_t)at;
  float v[R];
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = n[r];
  return FreeSlot<R>::make(v);
}

// Rows y[0..R) of the n reference projections as slots, encoded like
// load_sino_rows.
template <int R>
__device__ void load_sino_slots(const BpParams& p, int n, const int* y,
                                typename FreeSlot<R>::T* sq) {
  const int bx = p.bx;
  const size_t plane = (size_t)bx * p.by;
  for (int i = 0; i < n; i++) {
    const float c = p.c[i];
    const float* src = p.stack + (size_t)(p.i0 + i) * plane;
    for (int x = threadIdx.x; x < bx; x += blockDim.x) {
      float v[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const float s = src[(size_t)y[r] * bx + x];
        v[r] = s < -1e10f ? -INFINITY : s * c;
      }
      sq[i * bx + x] = FreeSlot<R>::make(v);
    }
  }
  __syncthreads();
}

// DIV1 (mode 8): in the test-free interval, where every sample's voxel has
// the same weight sum W, the ray adds the samples' numerators and divides
// their sum by W once, instead of dividing each sample (three instructions
// per row and sample fewer); tested samples divide by the hardware
// reciprocal of their weight sum (one ulp) with one multiply per row
// instead of the correctly rounded reciprocal and quotient.  The result
// differs from the reference's per-voxel division only in rounding, well
// inside 1e-5 relative of the FP64 evaluation.  A skipped sample leaves a
// non-finite ray sum (-inf, or NaN where a weight of 0 met it), which sends
// the ray to the exact per-row form as before.
template <int N, int R, bool DIV1>
__global__ __launch_bounds__(kRowsThreads) void k_bprp_free(const BpParams p,
                                                            const RpParams q,
                                                            int per) {
  typedef typename FreeSlot<R>::T Slot;
  extern __shared__ float4 smem_free[];
  Slot* sq = reinterpret_cast<Slot*>(smem_free);  // [N][bx]
  const int bx = p.bx;
  const int vx = 2 * bx;
  const float cx = (float)bx * 0.5f;
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  const float iend = (float)(bx - 1);
  const unsigned iend_bits = __float_as_uint(iend);
  const float ox = 0.5f - hvx, oz = 0.5f - htb;
  const bool fast_ok = p.cnt_all >= 0.001f;

  const int nt = blockDim.x;
  const int nch = cdiv(bx, nt);
  const int units = cdiv(p.by, R) * nch;
  const int ua = blockIdx.x * per, ub = min(units, ua + per);
  int gcur = -1;
  int y[R];
  for (int un = ua; un < ub; un++) {
    const int gr = un / nch;
    if (gr != gcur) {
      __syncthreads();  // all waves are done with the previous rows
      // Rows past the last repeat it (computed, not stored).
#pragma unroll
      for (int r = 0; r < R; r++) y[r] = min(R * gr + r, p.by - 1);
      load_sino_slots<R>(p, N, y, sq);
      gcur = gr;
    }
    const int u = (un - gr * nch) * nt + threadIdx.x;
    const bool live = u < bx;
    const int uu = live ? u : bx - 1;  // idle lanes shadow a real ray
    const float xp = (float)uu + 0.5f - (float)bx * 0.5f;
    const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
    const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
    const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
    int t0, t1;
    rp_sample_range(q, tx, tz, z0, &t0, &t1);
    int fa, fb, ia, ib;
    fast_interval(p, q, N, tx, tz, z0, &fa, &fb, &ia, &ib);
    fa = max(fa, t0);
    fb = min(fb, t1);
    // The wave's common fast interval, in scalar registers (the fast loop's
    // counter is then scalar too).
    const int ta = __builtin_amdgcn_readfirstlane(wave_max(fa));
    const int tb = __builtin_amdgcn_readfirstlane(wave_min(fb));
    const bool has_fast = fast_ok && ta < tb;
    const int a = has_fast ? ta : t1, b = has_fast ? tb : t1;
    float sum[R];
    int cnt[R];
    int nuse = 0;  // samples counted by every row (no skipped samples)
    // Samples that take the reference's tests, on the premise that no sample
    // is skipped: then the rows share the weight sum, the validity and the
    // count.  A skipped sample (-inf) makes a counted sample's quotient NaN,
    // which reaches the row's sum and sends the ray to 'exact'.  'inslab'
    // (a std::bool_constant) true: the samples are known to lie inside the
    // slab, so the slab bounds are not tested.
    auto tested = [&](int from, int to, auto inslab) {
      // Weights and slot-row addresses in vector registers: the scalar file
      // cannot hold them next to the range masks of this loop.
      float wv[N];
      unsigned rowv[N];
#pragma unroll
      for (int i = 0; i < N; i++) {
        wv[i] = in_vgpr(p.w[i]);
        rowv[i] = __float_as_uint(in_vgpr(__uint_as_float(
            (unsigned)(size_t)(sq + i * bx))));
      }
      for (int t = from; t < to; t++) {
        const float fz = (float)t + z0;
        const float X = fmaf(-fz, q.s, tx);
        const float Z = fmaf(fz, q.c, tz);
        const bool inb = decltype(inslab)::value ||
                         ((X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez));
        const float vfx = (inb ? truncf(X) : 0.f) + ox;
        const float vfz = (inb ? truncf(Z) : 0.f) + oz;
        Slot vals[N];
        float wsel[N];  // projection i's weight, 0 where out of range
#pragma unroll
        for (int i = 0; i < N; i++) {
          const float rr = fmaf(vfx, p.c[i], fmaf(vfz, p.s[i], cx));
          // The same unsigned range test as k_bprp_tab.
          const bool ok = __float_as_uint(rr) <= iend_bits;
          wsel[i] = ok ? wv[i] : 0.f;
          // Out of range: any slot of the row (the clamped one), read with
          // weight 0.
          const int ix = (int)__builtin_amdgcn_fmed3f(rr, 0.f, iend);
          const unsigned at = rowv[i] + ((unsigned)ix << kSlotShift<R>);
          vals[i] = lds_slot<R>(at);
        }
        // The weight sum in reference order from 0 (0 + w0 = w0: w0 >= +0).
        float acc[R], ws = wsel[0];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] = 0.f;
#pragma unroll
        for (int i = 0; i < N; i++) {
          const float wi = wsel[i];
          if (i > 0) ws += wi;
          float v4[R];
          FreeSlot<R>::split(vals[i], v4);
#pragma unroll
          for (int r = 0; r < R; r++) acc[r] = fmaf(v4[r], wi, acc[r]);
        }
        const float rc = DIV1 ? __builtin_amdgcn_rcpf(ws) : rcp_rn(ws);
        const bool use = inb && ws >= 0.001f;
        // An unused sample adds +0 (its quotient masked to zero bits; the
        // sums never hold -0), so no branch guards the quotients.
        const unsigned um = use ? ~0u : 0u;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float qv = DIV1 ? acc[r] * rc : div_by_rcp(acc[r], ws, rc);
          sum[r] += __uint_as_float(__float_as_uint(qv) & um);
        }
        nuse += use ? 1 : 0;
      }
    };
    // The reference's tests with per-row validity and weight sums, exact
    // whatever the skipped samples.
    auto exact = [&](int from, int to) {
      for (int t = from; t < to; t++) {
        const float fz = (float)t + z0;
        const float X = fmaf(-fz, q.s, tx);
        const float Z = fmaf(fz, q.c, tz);
        const bool inb = (X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez);
        const float vfx = (inb ? truncf(X) : 0.f) + ox;
        const float vfz = (inb ? truncf(Z) : 0.f) + oz;
        Slot vals[N];
        unsigned okm = 0;
#pragma unroll
        for (int i = 0; i < N; i++) {
          const float rr = fmaf(vfx, p.c[i], fmaf(vfz, p.s[i], cx));
          const bool ok = __float_as_uint(rr) <= iend_bits;
          okm |= ok ? (1u << i) : 0u;
          vals[i] = sq[i * bx + (ok ? (int)rr : 0)];
        }
        float acc[R], ws[R];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] = ws[r] = 0.f;
#pragma unroll
        for (int i = 0; i < N; i++) {
          const bool ok = (okm >> i) & 1u;
          float v4[R];
          FreeSlot<R>::split(vals[i], v4);
#pragma unroll
          for (int r = 0; r < R; r++) {
            const bool okr = ok && (v4[r] != -INFINITY);
            acc[r] = fmaf(okr ? v4[r] : 0.f, p.w[i], acc[r]);
            ws[r] += okr ? p.w[i] : 0.f;
          }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float v = (ws[r] < 0.001f) ? kInvalid : acc[r] / ws[r];
          const bool ok = inb && (v >= -1e10f);
          sum[r] += ok ? v : 0.0f;
          cnt[r] += ok ? 1 : 0;
        }
      }
    };
#pragma unroll
    for (int r = 0; r < R; r++) sum[r] = 0.0f;
    // [t0, a): margin samples [t0, m0) with the slab bounds tested, then
    // in-slab ones [m0, m1).  With a fast interval [a, b) lies inside the
    // lane's in-slab interval [ia, ib), so m1 = a; without, a = b = t1 and
    // the rest [m1, t1) is left to the second part.
    const int m0 = min(max(ia, t0), a);
    const int m1 = max(min(ib, a), m0);
    tested(t0, m0, std::false_type{});
    tested(m0, m1, std::true_type{});
    if (has_fast) {
      float sf[R];  // DIV1: the interval's numerators
#pragma unroll
      for (int r = 0; r < R; r++) sf[r] = 0.f;
      for (int t = ta; t < tb; t++) {
        const float fz = (float)t + z0;
        const float vfx = truncf(fmaf(-fz, q.s, tx)) + ox;
        const float vfz = truncf(fmaf(fz, q.c, tz)) + oz;
        // All lookups of the sample are issued before the first use, so
        // their LDS latencies overlap.
        Slot vals[N];
#pragma unroll
        for (int i = 0; i < N; i++)
          vals[i] = sq[i * bx + (int)fmaf(vfx, p.c[i], fmaf(vfz, p.s[i], cx))];
        float acc[R];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] = 0.f;
#pragma unroll
        for (int i = 0; i < N; i++) {
          float v4[R];
          FreeSlot<R>::split(vals[i], v4);
#pragma unroll
          for (int r = 0; r < R; r++) acc[r] = fmaf(v4[r], p.w[i], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
          if (DIV1)
            sf[r] += acc[r];
          else
            sum[r] += div_by_rcp(acc[r], p.cnt_all, p.rcp_all);
        }
      }
      if (DIV1) {
#pragma unroll
        for (int r = 0; r < R; r++) sum[r] += sf[r] / p.cnt_all;
      }
      nuse += tb - ta;
    }
    // [c0, t1): in-slab samples [c0, n1), then margin samples [n1, t1).
    const int c0 = has_fast ? b : m1;
    const int n1 = max(min(ib, t1), c0);
    tested(c0, n1, std::true_type{});
    tested(n1, t1, std::false_type{});
    bool seen = false;
#pragma unroll
    for (int r = 0; r < R; r++)
      seen |= DIV1 ? !isfinite(sum[r]) : sum[r] != sum[r];
    if (seen) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        sum[r] = 0.0f;
        cnt[r] = 0;
      }
      exact(t0, t1);
    } else {
#pragma unroll
      for (int r = 0; r < R; r++) cnt[r] = nuse;
    }
    if (live) {
      const float lim = 0.8f * (float)q.L;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (R * gr + r < p.by)
          q.out[(size_t)y[r] * bx + u] =
              ((float)cnt[r] < lim) ? kInvalid
                                    : sum[r] / (float)cnt[r];
    }
  }
}

// ------------------------------------------- materialised, R rows per voxel
// The slab written for R image rows at once: a voxel's sinogram index depends
// on (x, z) only, so one index computation and one LDS slot read serve the R
// rows, and each voxel is stored as R consecutive floats,
//   slab [cdiv(by, R)][tb][2bx][R]  (row y = R*g + r of group g at lane r).
// Each voxel takes the same terms in the same order as k_bp_slab, so every
// value is bit-identical to it; rows past by repeat the last row and are
// never read back.  Workgroup g owns a contiguous range of (group, iz) slab
// rows; no LDS is left beside the sinogram slots, so a skipped sample (-inf)
// is found from the voxel it makes non-finite, and that block is computed
// again with per-row validity.
template <int N, int R>
__global__ __launch_bounds__(kRowsThreads) void k_bp_slabr(const BpParams p,
                                                           int per) {
  typedef typename FreeSlot<R>::T Slot;
  extern __shared__ float4 smem_slab[];
  Slot* sq = reinterpret_cast<Slot*>(smem_slab);  // [n][bx]
  constexpr int V = kBpBlockR / 64;
  const int n = N ? N : p.n;
  const int bx = p.bx;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int vx = 2 * bx;
  const float cx = (float)bx * 0.5f, iend = (float)(bx - 1);
  const float hvx = (float)vx * 0.5f, htb = (float)p.tb * 0.5f;
  const LaneProj lp(p, n, lane);
  const bool all_fast_ok = p.cnt_all >= 0.001f;
  const int groups = cdiv(p.by, R);
  const int total = groups * p.tb;
  const int qa = blockIdx.x * per, qb = min(total, qa + per);
  Slot* slab = reinterpret_cast<Slot*>(p.slab);

  // The reference's per-voxel loop with per-row validity (range, skipped
  // samples) for the V voxels of this lane.
  auto general = [&](const float* fx, float zf, unsigned long long m_out,
                     float res[V][R]) {
    float acc[V][R], cnt[V][R];
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int r = 0; r < R; r++) acc[v][r] = cnt[v][r] = 0.f;
#pragma unroll 1
    for (int i = 0; i < n; i++) {
      if ((m_out >> i) & 1ull) continue;
      const float c = lane_f(lp.c, i), w = lane_f(lp.w, i);
      const float zsc = fmaf(zf, lane_f(lp.s, i), cx);
      const Slot* row = sq + i * bx;
#pragma unroll
      for (int v = 0; v < V; v++) {
        const float rr = fmaf(fx[v], c, zsc);
        const bool inr = (rr >= 0.f) && (rr <= iend);
        float val[R];
        FreeSlot<R>::split(row[(int)__builtin_amdgcn_fmed3f(rr, 0.f, iend)],
                           val);
#pragma unroll
        for (int r = 0; r < R; r++) {
          const bool ok = inr && (val[r] != -INFINITY);
          acc[v][r] = fmaf(ok ? val[r] : 0.f, w, acc[v][r]);
          cnt[v][r] += ok ? w : 0.f;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < V; v++)
#pragma unroll
      for (int r = 0; r < R; r++)
        res[v][r] = (cnt[v][r] < 0.001f) ? kInvalid : acc[v][r] / cnt[v][r];
  };

  int gcur = -1;
  for (int q0 = qa; q0 < qb;) {
    const int g = q0 / p.tb;
    const int za = q0 - g * p.tb;
    const int zb = min(p.tb, za + (qb - q0));
    q0 += zb - za;
    if (g != gcur) {
      __syncthreads();  // the previous group's waves are done with the slots
      int y[R];
#pragma unroll
      for (int r = 0; r < R; r++) y[r] = min(R * g + r, p.by - 1);
      load_sino_slots<R>(p, n, y, sq);
      gcur = g;
    }
    for (int iz = za + wave; iz < zb; iz += nwaves) {
      const float fz = (float)iz + 0.5f - htb;
      const float lz = fmaf(fz, lp.s, cx);  // lane i: zsc of projection i
      // With N fixed, the projections' z terms sit in registers and their
      // cos and weights are kernel arguments.
      float zs[N ? N : 1];
#pragma unroll
      for (int i = 0; i < (N ? N : 1); i++)
        zs[i] = N ? fmaf(fz, p.s[i], cx) : 0.f;
      Slot* orow = slab + ((size_t)g * p.tb + iz) * vx + lane;
      for (int ix0 = 0; ix0 < vx; ix0 += kBpBlockR) {
        const int ixl = min(ix0 + kBpBlockR - 1, vx - 1);
        const float rf = fmaf((float)ix0 + 0.5f - hvx, lp.c, lz);
        const float rl = fmaf((float)ixl + 0.5f - hvx, lp.c, lz);
        const unsigned long long m_in =
            __ballot(rf >= 0.f && rl <= iend) & lp.mask;
        const unsigned long long m_out =
            __ballot(rl < 0.f || rf > iend) & lp.mask;
        float fx[V], res[V][R];
#pragma unroll
        for (int v = 0; v < V; v++)
          fx[v] = (float)min(ix0 + lane + 64 * v, ixl) + 0.5f - hvx;
        if (m_out == lp.mask) {
#pragma unroll
          for (int v = 0; v < V; v++)
#pragma unroll
            for (int r = 0; r < R; r++) res[v][r] = kInvalid;
        } else if (m_in == lp.mask && all_fast_ok) {
          // Every projection in range over the block: no tests, one index
          // and one slot per lookup for the R rows.
          float acc[V][R];
#pragma unroll
          for (int v = 0; v < V; v++)
#pragma unroll
            for (int r = 0; r < R; r++) acc[v][r] = 0.f;
          auto lookup = [&](int i, float c, float w, float zsc) {
            const Slot* row = sq + i * bx;
#pragma unroll
            for (int v = 0; v < V; v++) {
              float val[R];
              FreeSlot<R>::split(row[(int)fmaf(fx[v], c, zsc)], val);
#pragma unroll
              for (int r = 0; r < R; r++)
                acc[v][r] = fmaf(val[r], w, acc[v][r]);
            }
          };
          if (N) {
#pragma unroll
            for (int i = 0; i < (N ? N : 1); i++)
              lookup(i, p.c[i], p.w[i], zs[i]);
          } else {
#pragma unroll 1
            for (int i = 0; i < n; i++)
              lookup(i, lane_f(lp.c, i), lane_f(lp.w, i), lane_f(lz, i));
          }
          // Valid samples give finite voxels, so a non-finite sum of the
          // lane's voxels means a skipped sample (-inf) reached one of them
          // (a false alarm only recomputes the block, exactly).
          float chk = 0.f;
#pragma unroll
          for (int v = 0; v < V; v++)
#pragma unroll
            for (int r = 0; r < R; r++) {
              res[v][r] = div_by_rcp(acc[v][r], p.cnt_all, p.rcp_all);
              chk += res[v][r];
            }
          const bool bad = !isfinite(chk);
          // A skipped sample made some voxel -inf or NaN.
          if (__ballot(bad)) general(fx, fz, m_out, res);
        } else {
          general(fx, fz, m_out, res);
        }
#pragma unroll
        for (int v = 0; v < V; v++)
          if (ix0 + lane + 64 * v < vx)
            stream_store(orow + ix0 + 64 * v, FreeSlot<R>::make(res[v]));
      }
    }
  }
}

// Ray sums over the R-row slab: one lane per detector position u and group
// g, the R rows' samples read as one slot.  The samples are visited in the
// order of k_rp_slab, so each row's output is bit-identical to it.
template <int R>
__global__ __launch_bounds__(256) void k_rp_slabr(const RpParams q) {
  typedef typename FreeSlot<R>::T Slot;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  if (u >= q.bx) return;
  const int vx = 2 * q.bx;
  const Slot* vol =
      reinterpret_cast<const Slot*>(q.slab) + (size_t)g * q.tb * vx;
  const float xp = (float)u + 0.5f - (float)q.bx * 0.5f;
  const float tx = fmaf(xp, q.c, (float)vx * 0.5f);
  const float tz = fmaf(xp, q.s, (float)q.tb * 0.5f);
  const float z0 = (-xp * q.s) / q.c - 0.5f * (float)q.L;
  const float iex = (float)(vx - 1), iez = (float)(q.tb - 1);
  int t0, t1;
  rp_sample_range(q, tx, tz, z0, &t0, &t1);
  float sum[R];
  int cnt[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    sum[r] = 0.0f;
    cnt[r] = 0;
  }
#pragma unroll 4
  for (int t = t0; t < t1; t++) {
    const float fz = (float)t + z0;
    const float X = fmaf(-fz, q.s, tx);
    const float Z = fmaf(fz, q.c, tz);
    const bool inb = (X >= 0.f) & (X < iex) & (Z >= 0.f) & (Z < iez);
    const int ix = inb ? (int)X : 0, iz = inb ? (int)Z : 0;
    float v[R];
    FreeSlot<R>::split(vol[(size_t)vx * iz + ix], v);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const bool ok = inb && (v[r] >= -1e10f);
      sum[r] += ok ? v[r] : 0.0f;
      cnt[r] += ok ? 1 : 0;
    }
  }
  const float lim = 0.8f * (float)q.L;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int y = R * g + r;
    if (y < q.by)
      q.out[(size_t)y * q.bx + u] =
          ((float)cnt[r] < lim) ? kInvalid : sum[r] / (float)cnt[r];
  }
}

__global__ void k_fill(float* p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

size_t bp_lds_bytes(int n, int bx) {
  return (size_t)n * ((size_t)bx + 1) * 4;
}

typedef void (*FusedKernel)(BpParams, RpParams, int);

template <int N>
struct KernelTable {
  static void fill(FusedKernel* fu) {
    fu[N] = k_bprp_fused<N>;
    KernelTable<N - 1>::fill(fu);
  }
};
template <>
struct KernelTable<0> {
  static void fill(FusedKernel* fu) { fu[0] = k_bprp_fused<0>; }
};

template <int N>
struct TabKernelTable {
  static void fill(FusedKernel* tk, FusedKernel* tk2, FusedKernel* f2,
                   FusedKernel* f4, FusedKernel* d2, FusedKernel* d4) {
    tk[N] = k_bprp_tab<N>;
    tk2[N] = k_bprp_rows<N, 2>;
    f2[N] = k_bprp_free<N, 2, false>;
    f4[N] = k_bprp_free<N, 4, false>;
    d2[N] = k_bprp_free<N, 2, true>;
    d4[N] = k_bprp_free<N, 4, true>;
    TabKernelTable<N - 1>::fill(tk, tk2, f2, f4, d2, d4);
  }
};
typedef void (*SlabKernel)(BpParams, int);

template <int N>
struct SlabKernelTable {
  static void fill(SlabKernel* s4, SlabKernel* s2) {
    s4[N] = k_bp_slabr<N, 4>;
    s2[N] = k_bp_slabr<N, 2>;
    SlabKernelTable<N - 1>::fill(s4, s2);
  }
};
template <>
struct SlabKernelTable<-1> {
  static void fill(SlabKernel*, SlabKernel*) {}
};

template <>
struct TabKernelTable<0> {
  static void fill(FusedKernel* tk, FusedKernel* tk2, FusedKernel* f2,
                   FusedKernel* f4, FusedKernel* d2, FusedKernel* d4) {
    tk[0] = tk2[0] = f2[0] = f4[0] = d2[0] = d4[0] = nullptr;
  }
};

struct Kernels {
  FusedKernel fu[kMaxFixedN + 1];
  FusedKernel tab[kTabMaxN + 1];
  FusedKernel rows2[kTabMaxN + 1];
  FusedKernel free2[kTabMaxN + 1];
  FusedKernel free4[kTabMaxN + 1];
  // The same with one division per ray's test-free interval (mode 8).
  FusedKernel div2[kTabMaxN + 1];
  FusedKernel div4[kTabMaxN + 1];
  // Materialised R-row backprojection, n = 1..kTabMaxN fixed, 0 runtime.
  SlabKernel slab4[kTabMaxN + 1];
  SlabKernel slab2[kTabMaxN + 1];
  // Static LDS of every kernel above and of k_bp_slab, read once (block
  // reductions such as __syncthreads_or use some).
  std::map<const void*, size_t> static_lds;
  Kernels() {
    KernelTable<kMaxFixedN>::fill(fu);
    TabKernelTable<kTabMaxN>::fill(tab, rows2, free2, free4, div2, div4);
    SlabKernelTable<kTabMaxN>::fill(slab4, slab2);
    auto note = [&](const void* k) {
      if (!k) return;
      hipFuncAttributes a;
      MF_HIP(hipFuncGetAttributes(&a, k));
      static_lds[k] = a.sharedSizeBytes;
    };
    for (FusedKernel k : fu) note((const void*)k);
    for (FusedKernel* t : {tab, rows2, free2, free4, div2, div4})
      for (int n = 0; n <= kTabMaxN; n++) note((const void*)t[n]);
    note((const void*)k_bp_slab);
    for (SlabKernel* t : {slab4, slab2})
      for (int n = 0; n <= kTabMaxN; n++) note((const void*)t[n]);
  }
};

const Kernels& kernels();

// Whether 'dyn' bytes of dynamic LDS fit in a workgroup's 160 KB next to
// kernel k's static LDS.
bool lds_fits(const void* k, size_t dyn) {
  if (!k || dyn > 160 * 1024) return false;
  const auto& m = kernels().static_lds;
  const auto it = m.find(k);
  MF_REQUIRE(it != m.end(), "kernel without an LDS record");
  return dyn + it->second <= 160 * 1024;
}

const Kernels& kernels() {
  static Kernels k;
  return k;
}

}  // namespace

void find_proj_range(const float* tilts, int nz, const unsigned char* skip,
                     int target, int nproj, int range[2]) {
  // CalcReproj.cu:206-259, float/double semantics kept.
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
  bp.rcp_all = n > 0 ? 1.0f / cnt : 0.f;

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
  if ((mode >= 1 && mode <= 5) || mode == 8) {
    if (n == 0) {
      hipLaunchKernelGGL(k_fill, dim3(cdiv((long long)bx * by, 256)),
                         dim3(256), 0, st, d_reproj, (size_t)bx * by,
                         kInvalid);
      MF_HIP(hipGetLastError());
      return;
    }
    FusedKernel k = kernels().fu[kid];
    size_t flds = lds;
    int threads = kFuThreads;
    long long units = (long long)by * cdiv(bx, kFuThreads);
    // Multi-row forms.  Mode 1 takes the table-free form with the most image
    // rows per unit whose slots fit in LDS (4, then 2; measured faster than
    // the two-row form with tables wherever both fit), then the two-row form
    // with tables, then one row; modes 3, 4 and 5 pin one row, two rows with
    // tables and two rows without.
    const bool tab = (mode == 1 || mode >= 3) && n <= kTabMaxN;
    const Kernels& ks = kernels();
    // Mode 8: mode 1's selection with the one-division forms of the
    // table-free kernels.
    const bool m1 = mode == 1 || mode == 8;
    const FusedKernel* f4s = mode == 8 ? ks.div4 : ks.free4;
    const FusedKernel* f2s = mode == 8 ? ks.div2 : ks.free2;
    int rows = 0;
    if (tab && m1 && by >= 4 &&
        lds_fits((const void*)f4s[n], free_lds_bytes(4, n, bx))) {
      rows = 4;
      k = f4s[n];
      flds = free_lds_bytes(4, n, bx);
    } else if (tab && (m1 || mode == 5) && by > 1 &&
               lds_fits((const void*)f2s[n], free_lds_bytes(2, n, bx))) {
      rows = 2;
      k = f2s[n];
      flds = free_lds_bytes(2, n, bx);
    } else if (tab && (m1 || mode == 4) && by > 1 &&
               lds_fits((const void*)ks.rows2[n], rows_lds_bytes(n, bx, tb))) {
      rows = 2;
      k = ks.rows2[n];
      flds = rows_lds_bytes(n, bx, tb);
    } else if (tab && lds_fits((const void*)ks.tab[n], tab_lds_bytes(n, bx, tb))) {
      k = ks.tab[n];
      flds = tab_lds_bytes(n, bx, tb);
    }
    if (rows > 0) {
      // Groups of image rows; ray chunks of at most kRowsThreads, evened out.
      const int nch = cdiv(bx, kRowsThreads);
      threads = cdiv(cdiv(bx, nch), 64) * 64;
      units = (long long)cdiv(by, rows) * nch;
    }
    MF_REQUIRE(lds_fits((const void*)k, flds), "sinogram rows exceed LDS");
    ensure_dyn_lds(ctx, (const void*)k, flds);
    // One resident wave of workgroups over the units.
    const int occ = occupancy(ctx, (const void*)k, threads, flds);
    const long long slots = (long long)std::max(occ, 1) * std::max(ctx->num_cu, 1);
    const int per = cdiv(units, std::min(units, slots));
    TimedLaunch tl(ctx->timer, "bprp_fused", st);
    hipLaunchKernelGGL(k, dim3(cdiv(units, per)), dim3(threads), flds, st, bp,
                       rp, per);
    MF_HIP(hipGetLastError());
    return;
  }
  if (mode == 6 || mode == 7) {
    // Materialised slab of R image rows per voxel (R = 4 or 2).
    const int R = mode == 6 ? 4 : 2;
    const int groups = cdiv(by, R);
    const size_t elems = (size_t)groups * R * tb * 2 * bx;
    float* slab = d_slab;
    if (!slab) slab = (float*)ctx->slab_ws.ensure(ctx, elems * 4);
    bp.slab = slab;
    rp.slab = slab;
    if (n == 0) {
      hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, st, slab, elems,
                         kInvalid);
    } else {
      const int sid = n <= kTabMaxN ? n : 0;
      const SlabKernel kf =
          R == 4 ? kernels().slab4[sid] : kernels().slab2[sid];
      const void* k = (const void*)kf;
      const size_t flds = free_lds_bytes(R, n, bx);
      MF_REQUIRE(lds_fits(k, flds),
                 "sinogram rows of this many image rows exceed LDS");
      ensure_dyn_lds(ctx, k, flds);
      const int occ = occupancy(ctx, k, kRowsThreads, flds);
      const long long rows = (long long)groups * tb;
      const long long slots =
          (long long)std::max(occ, 1) * std::max(ctx->num_cu, 1);
      const int per = (int)cdiv(rows, std::min(rows, slots));
      TimedLaunch tl(ctx->timer, "bp_slab_rows", st);
      hipLaunchKernelGGL(kf, dim3((int)cdiv(rows, per)), dim3(kRowsThreads),
                         flds, st, bp, per);
    }
    MF_HIP(hipGetLastError());
    {
      TimedLaunch tl(ctx->timer, "rp_slab_rows", st);
      if (R == 4)
        hipLaunchKernelGGL(k_rp_slabr<4>, dim3(cdiv(bx, 256), groups),
                           dim3(256), 0, st, rp);
      else
        hipLaunchKernelGGL(k_rp_slabr<2>, dim3(cdiv(bx, 256), groups),
                           dim3(256), 0, st, rp);
    }
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
    const void* k = (const void*)k_bp_slab;
    MF_REQUIRE(lds_fits(k, lds), "sinogram rows exceed LDS");
    ensure_dyn_lds(ctx, k, lds);
    // One resident wave of workgroups, each owning an equal share of the
    // by*tb slab rows.
    const int occ = occupancy(ctx, k, kBpThreads, lds);
    const long long rows = (long long)by * tb;
    const long long slots = (long long)std::max(occ, 1) * std::max(ctx->num_cu, 1);
    const int per = (int)cdiv(rows, std::min(rows, slots));
    const int grid = (int)cdiv(rows, per);
    TimedLaunch tl(ctx->timer, "bp_slab", st);
    hipLaunchKernelGGL(k_bp_slab, dim3(grid), dim3(kBpThreads), lds, st, bp,
                       per);
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
