// N_r plan: ball weights laid out per (dx, dy) column for the direct
// convolution, plus the tile marks that select where N_r is evaluated.
#pragma once

#include "common.h"

namespace svimx {

// N_r is evaluated on tiles of kTX x kTY x kTZ nodes: one workgroup per tile,
// one wave (64 lanes = kTY x kTZ) per x-layer.
constexpr int kTX = 4, kTY = 8, kTZ = 8;

struct NrColumn {
  int dy;         // column offset in y (dx is implied by the column group)
  int dz0, nz;    // first dz and number of nonzero taps
  int woff;       // offset into the weight array
};

struct NrPlan {
  int m;
  double h, radius;
  Dims d;
  int ntx, nty, ntz;
  int64_t ntiles;
  int ncols;
  int nweights;
  float scale;        // float32(h^3)
  float negbv;        // float32(-1 / (4/3 pi r^3))
  NrColumn* cols;     // device, grouped by dx
  int* col_start;     // device, 2m+2 entries: columns of dx = g - m are [col_start[g], col_start[g+1])
  float* weights;     // device
  uint8_t* marks;     // device, ntiles
  int32_t* list;      // device, ntiles
  int32_t* count;     // device, 1
};

// Mark the tiles covering the tricubic stencil [base, base+3]^3 of a point.
__device__ __forceinline__ void mark_stencil_tiles(uint8_t* marks, int nty, int ntz, int bx, int by, int bz) {
  int tx0 = bx / kTX, tx1 = (bx + 3) / kTX;
  int ty0 = by / kTY, ty1 = (by + 3) / kTY;
  int tz0 = bz / kTZ, tz1 = (bz + 3) / kTZ;
  for (int tx = tx0; tx <= tx1; ++tx)
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tz = tz0; tz <= tz1; ++tz) marks[((int64_t)tx * nty + ty) * ntz + tz] = 1;
}

}  // namespace svimx
