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
// (table-driven where the tables fit in LDS), mode 2: fused without tables.
void reproject_target(mf_ctx* ctx, hipStream_t st, const float* d_stack,
                      int bx, int by, int nz, int tb, const float* tilts,
                      const unsigned char* skip, int target, int nproj,
                      int mode, float* d_reproj, float* d_slab);

}  // namespace mf
