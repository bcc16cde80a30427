// Small device helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

// Workgroup reduction (64-wide waves, up to 16 waves): shuffle within each
// wave, then the first wave combines the per-wave partials.  Lanes beyond
// the wave count contribute 'identity'.  The result is valid in thread 0.
// 'sh' must hold 16 elements.
template <typename T, typename Op>
__device__ T block_reduce(T v, T* sh, Op op, T identity) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
  if (lane == 0) sh[w] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  if (w == 0) {
    v = lane < nw ? sh[lane] : identity;
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
  }
  __syncthreads();
  return v;
}

// Workgroup index remapped so that consecutive logical blocks run on the same
// XCD: the dispatcher hands workgroup g to XCD g mod 8, so logical blocks
// [x*per, (x+1)*per) (per = nb/8, the first nb mod 8 XCDs taking one more)
// all land on XCD x.  A bijection on [0, nb); only locality depends on it.
__device__ __forceinline__ int xcd_block(int g, int nb) {
  constexpr int kXcd = 8;
  const int x = g % kXcd, slot = g / kXcd;
  const int per = nb / kXcd, rem = nb % kXcd;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + slot;
}

}  // namespace mf
