// Device-side FFT building blocks (LDS Stockham stages, Bluestein, real
// packing).  Included by the .hip files whose kernels run transforms inside
// their own workgroups (correlation, R-weighting, common-line spectra).
#pragma once
#include "fft.h"

namespace mf {

// Walks the (line, index) pairs l < nl, k < n of a workgroup's lines, one
// pair per thread per step, without an integer division per step: the
// block stride is split once into whole lines and a remainder.
struct LineIter {
  int l, k, dl, dk, n;
  __device__ __forceinline__ explicit LineIter(int n_) : n(n_) {
    l = (int)threadIdx.x / n;
    k = (int)threadIdx.x - l * n;
    dl = (int)blockDim.x / n;
    dk = (int)blockDim.x - dl * n;
  }
  __device__ __forceinline__ void next() {
    l += dl;
    k += dk;
    if (k >= n) {
      k -= n;
      l++;
    }
  }
};

// floor(a / d) for 0 <= a < 2^22 and d >= 1 from r = 1.0f / d: (a + 0.5)/d
// lies at least 0.5/d from any integer, and the two roundings of the float
// evaluation err by less than (a + 0.5)/d * 2^-23 < 0.5/d, so truncation
// gives the exact quotient.
__device__ __forceinline__ int qdiv(int a, float r) {
  return (int)(((float)a + 0.5f) * r);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
// Multiply by -i (forward) or +i (inverse).
__device__ __forceinline__ float2 cmul_mi(float2 a, bool inv) {
  return inv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// One radix-R butterfly of a Stockham stage: reads R inputs src[q*ss],
// applies twiddles tw[q*tws], DFT_R, writes dst[p*ds].
__device__ __forceinline__ void fft_butterfly(int R, const float2* src, int ss,
                                              float2* dst, int ds, int tws,
                                              const float2* __restrict__ tw,
                                              int N, bool inv) {
  const float kS = 0.70710678118654752f;
  float2 x[8];
  if (R <= 8 && R != 6 && R != 7) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      if (q < R) {
        float2 v = src[q * ss];
        if (q > 0 && tws) {
          float2 w = tw[q * tws];
          v = cmul(v, inv ? cconj(w) : w);
        }
        x[q] = v;
      }
    }
  }
  switch (R) {
    case 2: {
      dst[0] = cadd(x[0], x[1]);
      dst[ds] = csub(x[0], x[1]);
      return;
    }
    case 3: {
      const float c = -0.5f, s = 0.86602540378443865f;
      float2 t = cadd(x[1], x[2]);
      float2 u = csub(x[1], x[2]);
      float2 a = make_float2(x[0].x + c * t.x, x[0].y + c * t.y);
      float2 b = cmul_mi(cscale(u, s), inv);
      dst[0] = cadd(x[0], t);
      dst[ds] = cadd(a, b);
      dst[2 * ds] = csub(a, b);
      return;
    }
    case 4: {
      float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
      float2 c = cadd(x[1], x[3]), d = cmul_mi(csub(x[1], x[3]), inv);
      dst[0] = cadd(a, c);
      dst[ds] = cadd(b, d);
      dst[2 * ds] = csub(a, c);
      dst[3 * ds] = csub(b, d);
      return;
    }
    case 5: {
      const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
      const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
      float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
      float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
      float2 r1 = make_float2(x[0].x + c1 * a1.x + c2 * a2.x,
                              x[0].y + c1 * a1.y + c2 * a2.y);
      float2 r2 = make_float2(x[0].x + c2 * a1.x + c1 * a2.x,
                              x[0].y + c2 * a1.y + c1 * a2.y);
      float2 i1 = cmul_mi(make_float2(s1 * b1.x + s2 * b2.x,
                                      s1 * b1.y + s2 * b2.y), inv);
      float2 i2 = cmul_mi(make_float2(s2 * b1.x - s1 * b2.x,
                                      s2 * b1.y - s1 * b2.y), inv);
      dst[0] = make_float2(x[0].x + a1.x + a2.x, x[0].y + a1.y + a2.y);
      dst[ds] = cadd(r1, i1);
      dst[4 * ds] = csub(r1, i1);
      dst[2 * ds] = cadd(r2, i2);
      dst[3 * ds] = csub(r2, i2);
      return;
    }
    case 8: {
      // Two radix-4 transforms on even/odd inputs, then radix-2 combine.
      float2 a = cadd(x[0], x[4]), b = csub(x[0], x[4]);
      float2 c = cadd(x[2], x[6]), d = cmul_mi(csub(x[2], x[6]), inv);
      float2 e0 = cadd(a, c), e1 = cadd(b, d), e2 = csub(a, c), e3 = csub(b, d);
      a = cadd(x[1], x[5]);
      b = csub(x[1], x[5]);
      c = cadd(x[3], x[7]);
      d = cmul_mi(csub(x[3], x[7]), inv);
      float2 o0 = cadd(a, c), o1 = cadd(b, d), o2 = csub(a, c), o3 = csub(b, d);
      // w8^1 = (s, -s) forward, w8^2 = -i, w8^3 = (-s, -s) forward.
      float2 w1 = inv ? make_float2(kS, kS) : make_float2(kS, -kS);
      float2 w3 = inv ? make_float2(-kS, kS) : make_float2(-kS, -kS);
      o1 = cmul(o1, w1);
      o2 = cmul_mi(o2, inv);
      o3 = cmul(o3, w3);
      dst[0] = cadd(e0, o0);
      dst[4 * ds] = csub(e0, o0);
      dst[ds] = cadd(e1, o1);
      dst[5 * ds] = csub(e1, o1);
      dst[2 * ds] = cadd(e2, o2);
      dst[6 * ds] = csub(e2, o2);
      dst[3 * ds] = cadd(e3, o3);
      dst[7 * ds] = csub(e3, o3);
      return;
    }
    default: {
      // Generic odd-prime radix: O(R^2) from LDS.  The stage twiddles are
      // applied to the inputs once, in place: within a stage every input
      // belongs to this butterfly alone, and the input buffer is the next
      // stage's output (with no twiddles in the first stage, Ns = 1, the
      // caller's data are never modified).  DFT_R twiddles come from the
      // length-N table, w_R^{pq} = tw[(pq mod R) * N/R], with pq mod R
      // stepped by pp.
      const int step = N / R;
      if (tws) {
        float2* s = const_cast<float2*>(src);
        for (int q = 1; q < R; q++) {
          const float2 w = tw[q * tws];
          s[q * ss] = cmul(s[q * ss], inv ? cconj(w) : w);
        }
      }
      for (int pp = 0; pp < R; pp++) {
        float2 acc = src[0];
        int k = 0;
#pragma unroll 4
        for (int q = 1; q < R; q++) {
          k += pp;
          if (k >= R) k -= R;
          const float2 wr = tw[k * step];
          acc = cadd(acc, cmul(src[q * ss], inv ? cconj(wr) : wr));
        }
        dst[pp * ds] = acc;
      }
      return;
    }
  }
}

// One Stockham stage of radix R over nl lines: butterfly j = (q, k) reads
// x[j + p*N/R] and writes y[q*Ns*R + k + p*Ns].  R > 0 is fixed at compile
// time (its butterfly unrolled, no per-butterfly radix dispatch); R = 0 takes
// the radix from 'rr' at run time.
template <int R>
__device__ __forceinline__ void stockham_stage(const float2* x, float2* y,
                                               int N, int ls, int Ns, int rr,
                                               const float2* __restrict__ tw,
                                               int nl, bool inv) {
  const int r = R ? R : rr;
  const int NR = N / r;
  const int tstride = N / (Ns * r);
  const float rNs = 1.0f / (float)Ns;
  for (LineIter it(NR); it.l < nl; it.next()) {
    const int j = it.k;
    const int q = qdiv(j, rNs), k = j - q * Ns;
    fft_butterfly(r, x + it.l * ls + j, NR, y + it.l * ls + q * Ns * r + k,
                  Ns, k * tstride, tw, N, inv);
  }
}

// Direct Stockham transform of nl lines of length N (line stride ls) held in
// x; y is scratch of the same shape.  Returns the buffer holding the result.
// Caller must __syncthreads() before (data loaded) -- stages sync themselves.
__device__ __forceinline__ float2* fft_stockham(float2* x, float2* y, int N,
                                                int ls, int nrad,
                                                const int* __restrict__ rad,
                                                const float2* __restrict__ tw,
                                                int nl, bool inv) {
  int Ns = 1;
  for (int s = 0; s < nrad; s++) {
    const int R = rad[s];
    switch (R) {
      case 2: stockham_stage<2>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 3: stockham_stage<3>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 4: stockham_stage<4>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 5: stockham_stage<5>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 7: stockham_stage<7>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 8: stockham_stage<8>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 11: stockham_stage<11>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      case 13: stockham_stage<13>(x, y, N, ls, Ns, R, tw, nl, inv); break;
      default: stockham_stage<0>(x, y, N, ls, Ns, R, tw, nl, inv); break;
    }
    __syncthreads();
    float2* t = x;
    x = y;
    y = t;
    Ns *= R;
  }
  return x;
}

// ------------------------------------------------ power-of-two transforms
// Lines of N = 2^k points (256..4096) take a Stockham schedule fixed at
// compile time: radix-8 stages, then one radix-2 or radix-4 stage for the
// remainder.  Every index is a shift or mask, every stride an immediate
// offset, and the complex products are contracted to FMAs; no runtime
// radix dispatch remains.  Same twiddle table and Stockham ordering as the
// generic path (results agree to FP32 rounding, not bit for bit).

__host__ __device__ constexpr int pow2_log(int n) {
  return n <= 1 ? 0 : 1 + pow2_log(n / 2);
}
// Radix of stage s for N: eights, then 2^(log N mod 3).
__host__ __device__ constexpr int pow2_radix(int N, int s) {
  return s < pow2_log(N) / 3 ? 8 : (1 << (pow2_log(N) % 3));
}
__host__ __device__ constexpr int pow2_stages(int N) {
  return pow2_log(N) / 3 + (pow2_log(N) % 3 ? 1 : 0);
}

template <bool INV>
__device__ __forceinline__ float2 tw_mul(float2 v, float2 w) {
#pragma clang fp contract(fast)
  return INV ? make_float2(v.x * w.x + v.y * w.y, v.y * w.x - v.x * w.y)
             : make_float2(v.x * w.x - v.y * w.y, v.y * w.x + v.x * w.y);
}

// One radix-R butterfly (R = 2, 4, 8): inputs src[q*SS] times twiddles
// tw[q*tws], DFT_R, outputs dst[p*DS].
template <int R, int SS, int DS, bool INV>
__device__ __forceinline__ void pow2_butterfly(const float2* src, float2* dst,
                                               int tws,
                                               const float2* __restrict__ tw) {
#pragma clang fp contract(fast)
  float2 x[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    x[q] = src[q * SS];
    if (q > 0 && tws) x[q] = tw_mul<INV>(x[q], tw[q * tws]);
  }
  if (R == 2) {
    dst[0] = cadd(x[0], x[1]);
    dst[DS] = csub(x[0], x[1]);
  } else if (R == 4) {
    const float2 a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
    const float2 c = cadd(x[1], x[3]), d = cmul_mi(csub(x[1], x[3]), INV);
    dst[0] = cadd(a, c);
    dst[DS] = cadd(b, d);
    dst[2 * DS] = csub(a, c);
    dst[3 * DS] = csub(b, d);
  } else {
    const float kS = 0.70710678118654752f;
    float2 a = cadd(x[0], x[4]), b = csub(x[0], x[4]);
    float2 c = cadd(x[2], x[6]), d = cmul_mi(csub(x[2], x[6]), INV);
    const float2 e0 = cadd(a, c), e1 = cadd(b, d), e2 = csub(a, c),
                 e3 = csub(b, d);
    a = cadd(x[1], x[5]);
    b = csub(x[1], x[5]);
    c = cadd(x[3], x[7]);
    d = cmul_mi(csub(x[3], x[7]), INV);
    float2 o0 = cadd(a, c), o1 = cadd(b, d), o2 = csub(a, c), o3 = csub(b, d);
    // w8^1 o1 and w8^3 o3: (x, y) * (s, -s) forward, (s, s) inverse.
    o1 = INV ? make_float2(kS * (o1.x - o1.y), kS * (o1.x + o1.y))
             : make_float2(kS * (o1.x + o1.y), kS * (o1.y - o1.x));
    o2 = cmul_mi(o2, INV);
    o3 = INV ? make_float2(-kS * (o3.x + o3.y), kS * (o3.x - o3.y))
             : make_float2(kS * (o3.y - o3.x), -kS * (o3.x + o3.y));
    dst[0] = cadd(e0, o0);
    dst[4 * DS] = csub(e0, o0);
    dst[DS] = cadd(e1, o1);
    dst[5 * DS] = csub(e1, o1);
    dst[2 * DS] = cadd(e2, o2);
    dst[6 * DS] = csub(e2, o2);
    dst[3 * DS] = cadd(e3, o3);
    dst[7 * DS] = csub(e3, o3);
  }
}

// The two halves of pow2_butterfly, for in-place stages: load R inputs
// src[q*SS] times twiddles tw[q*tws] into x, then DFT_R of x to dst[p*DS].
template <int R, int SS, bool INV>
__device__ __forceinline__ void pow2_load(const float2* src, int tws,
                                          const float2* __restrict__ tw,
                                          float2* x) {
#pragma clang fp contract(fast)
#pragma unroll
  for (int q = 0; q < R; q++) {
    x[q] = src[q * SS];
    if (q > 0 && tws) x[q] = tw_mul<INV>(x[q], tw[q * tws]);
  }
}
// Stage S of the N-point schedule over nl lines (stride N) from x to y.
template <int N, int S, bool INV>
__device__ __forceinline__ void pow2_stage(const float2* x, float2* y, int nl,
                                           const float2* __restrict__ tw) {
  constexpr int R = pow2_radix(N, S);
  constexpr int NS = S == 0 ? 1 : (1 << (3 * S));  // product of earlier radices
  constexpr int NR = N / R;
  constexpr int TSTRIDE = N / (NS * R);
  for (LineIter it(NR); it.l < nl; it.next()) {
    const int j = it.k;
    const int q = j / NS, k = j & (NS - 1);
    pow2_butterfly<R, NR, NS, INV>(x + it.l * N + j,
                                   y + it.l * N + q * NS * R + k,
                                   k * TSTRIDE, tw);
  }
  __syncthreads();
}

template <int N, int S, bool INV>
__device__ __forceinline__ float2* pow2_stages_from(float2* x, float2* y,
                                                    int nl,
                                                    const float2* tw) {
  if constexpr (S == pow2_stages(N)) {
    return x;
  } else {
    pow2_stage<N, S, INV>(x, y, nl, tw);
    return pow2_stages_from<N, S + 1, INV>(y, x, nl, tw);
  }
}

// Stage S in place over LINES lines (stride LS) with THREADS threads: every
// thread loads the inputs of its butterflies (at most MAXB) into registers,
// the workgroup syncs, then each writes its outputs over the same buffer.
// Same arithmetic as pow2_stage, half the LDS.
template <int N, int S, bool INV, int LINES, int THREADS, int LS>
__device__ __forceinline__ void pow2_stage_inplace(float2* x,
                                                   const float2* __restrict__ tw) {
  constexpr int R = pow2_radix(N, S);
  constexpr int NS = S == 0 ? 1 : (1 << (3 * S));
  constexpr int NR = N / R;
  constexpr int TSTRIDE = N / (NS * R);
  constexpr int MAXB = (LINES * NR + THREADS - 1) / THREADS;
  float2 v[MAXB][R];
  int dst[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; b++) {
    const int t = (int)threadIdx.x + b * THREADS;  // butterfly index
    dst[b] = -1;
    if (t < LINES * NR) {
      const int l = t / NR, j = t - l * NR;
      const int q = j / NS, k = j & (NS - 1);
      pow2_load<R, NR, INV>(x + l * LS + j, k * TSTRIDE, tw, v[b]);
      dst[b] = l * LS + q * NS * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; b++)
    if (dst[b] >= 0) pow2_butterfly<R, 1, NS, INV>(v[b], x + dst[b], 0, tw);
  __syncthreads();
}

template <int N, int S, bool INV, int LINES, int THREADS, int LS>
__device__ __forceinline__ void pow2_inplace_from(float2* x,
                                                  const float2* tw) {
  if constexpr (S < pow2_stages(N)) {
    pow2_stage_inplace<N, S, INV, LINES, THREADS, LS>(x, tw);
    pow2_inplace_from<N, S + 1, INV, LINES, THREADS, LS>(x, tw);
  }
}

// In-place transform of LINES lines of N points (stride LS) in x with a
// THREADS-thread workgroup; bit-identical to fft_pow2.  Caller syncs after
// loading; the result is in x on return (synchronised).
template <int N, bool INV, int LINES, int THREADS, int LS>
__device__ __forceinline__ void fft_pow2_inplace(float2* x, const float2* tw) {
  pow2_inplace_from<N, 0, INV, LINES, THREADS, LS>(x, tw);
}

// Complex transform of nl lines of N points (stride N) in x, y scratch;
// returns the buffer holding the result.  Caller syncs after loading.
template <int N, bool INV>
__device__ __forceinline__ float2* fft_pow2(float2* x, float2* y, int nl,
                                           const float2* tw) {
  return pow2_stages_from<N, 0, INV>(x, y, nl, tw);
}

// Line stride in LDS for a plan.
__host__ __device__ __forceinline__ int fft_lds_len(const FftDevPlan& P) {
  return P.blue ? P.m : P.n;
}

// Copies the twiddle table of P (fft_lds_len(P) entries) into LDS; the
// caller synchronises before the first transform.
// Twiddles to LDS, four per thread per round with the round's loads all in
// flight before its LDS writes (one memory round trip for up to 4 blockDim.x
// entries; a load-then-write loop waits once per entry).
__device__ __forceinline__ void fft_load_tw(const FftDevPlan& P, float2* tw) {
  const int L = fft_lds_len(P);
  const int bd = blockDim.x;
  for (int i0 = threadIdx.x; i0 < L; i0 += 4 * bd) {
    float2 t[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      t[j] = i0 + j * bd < L ? P.tw[i0 + j * bd] : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (i0 + j * bd < L) tw[i0 + j * bd] = t[j];
  }
}

// Complex transform of nl lines (length P.n, stride fft_lds_len(P)) held in
// buf; tmp is scratch of the same size; tw is the LDS copy of the plan's
// twiddles (fft_load_tw).  Input in the first n entries of each line;
// returns the buffer holding the n outputs.  Syncs internally; the caller
// must sync after loading.
__device__ __forceinline__ float2* fft_lines(float2* buf, float2* tmp,
                                             const FftDevPlan& P, int nl,
                                             bool inv, const float2* tw) {
  if (!P.blue) {
    switch (P.n) {
      case 256:
        return inv ? fft_pow2<256, true>(buf, tmp, nl, tw)
                   : fft_pow2<256, false>(buf, tmp, nl, tw);
      case 512:
        return inv ? fft_pow2<512, true>(buf, tmp, nl, tw)
                   : fft_pow2<512, false>(buf, tmp, nl, tw);
      case 1024:
        return inv ? fft_pow2<1024, true>(buf, tmp, nl, tw)
                   : fft_pow2<1024, false>(buf, tmp, nl, tw);
      case 2048:
        return inv ? fft_pow2<2048, true>(buf, tmp, nl, tw)
                   : fft_pow2<2048, false>(buf, tmp, nl, tw);
      case 4096:
        return inv ? fft_pow2<4096, true>(buf, tmp, nl, tw)
                   : fft_pow2<4096, false>(buf, tmp, nl, tw);
      default:
        return fft_stockham(buf, tmp, P.n, P.n, P.nrad, P.drad, tw, nl, inv);
    }
  }
  const int n = P.n, m = P.m;
  for (LineIter it(m); it.l < nl; it.next()) {
    const int l = it.l, k = it.k;
    float2 v = make_float2(0.f, 0.f);
    if (k < n) {
      float2 c = P.chirp[k];
      v = cmul(buf[l * m + k], inv ? cconj(c) : c);
    }
    buf[l * m + k] = v;
  }
  __syncthreads();
  float2* r = fft_stockham(buf, tmp, m, m, P.nrad, P.drad, tw, nl, false);
  float2* o = (r == buf) ? tmp : buf;
  const float2* f = inv ? P.bi : P.bf;
  for (LineIter it(m); it.l < nl; it.next())
    r[it.l * m + it.k] = cmul(r[it.l * m + it.k], f[it.k]);
  __syncthreads();
  r = fft_stockham(r, o, m, m, P.nrad, P.drad, tw, nl, true);
  for (LineIter it(m); it.l < nl; it.next()) {
    const int l = it.l, k = it.k;
    if (k < n) {
      float2 c = P.chirp[k];
      r[l * m + k] = cmul(r[l * m + k], inv ? cconj(c) : c);
    }
  }
  __syncthreads();
  return r;
}

// Real-to-complex post-processing of one line: Z (N complex, in LDS) holds
// the length-N transform of z_n = x_2n + i x_2n+1; writes X_k, k = 0..N, of
// the length-2N real input, times 'scale', through out(k, X).
template <typename Out>
__device__ __forceinline__ void r2c_unpack(const float2* Z, int N,
                                           const float2* __restrict__ half,
                                           float scale, int k, Out out) {
  const float2 zk = Z[k == N ? 0 : k];
  const float2 zn = cconj(Z[k == 0 ? 0 : N - k]);
  const float2 e = cscale(cadd(zk, zn), 0.5f);
  const float2 o = cscale(csub(zk, zn), 0.5f);
  // X_k = E + (-i) w_k O,  w_k = exp(-i pi k / N)
  const float2 t = cmul_mi(cmul(half[k], o), false);
  out(k, cscale(cadd(e, t), scale));
}

// Complex-to-real pre-processing: given X_k (k = 0..N) returns Z'_k (k < N)
// whose unnormalised inverse length-N transform holds x_2m in .x and x_2m+1
// in .y (cufftExecC2R convention: no 1/n).
__device__ __forceinline__ float2 c2r_pack(float2 xk, float2 xnk, float2 hk) {
  const float2 c = cconj(xnk);
  const float2 a = cadd(xk, c);
  const float2 b = csub(xk, c);
  // + i * conj(w_k) * b
  const float2 t = cmul(cconj(hk), b);
  return make_float2(a.x - t.y, a.y + t.x);
}

}  // namespace mf
