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

}  // namespace mf
