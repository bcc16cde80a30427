// Backprojection / reprojection launchers (CalcReprojFBP counterpart).
#pragma once
#include "common.h"

namespace mf {

// Most reference projections one target may use (FindProjRange caps the
// range at NPROJ and at +-20.5 degrees; 64 covers any practical NPROJ).
constexpr int kMaxRefProj = 64;

void find_proj_range(const float* tilts, int nz, const unsigned char* skip,
                     int target, int nproj, int range[2]);

// mode 0: materialised slab (d_slab or the context workspace), mode 1: fused
// (as many rows per unit as LDS holds), mode 2: fused one row without
// tables, modes 3-5: pinned fused forms, modes 6 and 7: materialised slab of
// 4 or 2 image rows per voxel ([cdiv(by, R)][tb][2bx][R]).  Identical
// reprojections in every mode.
void reproject_target(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                      int bx, int by, int nz, int tb, const float* tilts,
                      const unsigned char* skip, int target, int nproj,
                      int mode, float* d_reproj, float* d_slab);

// Floats of the slab a materialised mode (0, 6, 7) writes.
inline size_t slab_elems(int bx, int by, int tb, int mode) {
  const int r = mode == 6 ? 4 : mode == 7 ? 2 : 1;
  return (size_t)((by + r - 1) / r) * r * tb * 2 * bx;
}

}  // namespace mf
