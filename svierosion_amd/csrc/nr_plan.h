// N_r plan: ball weights laid out per (dx, dy) column for the direct
// convolution, plus the marks that select where N_r is evaluated.
//
// Two evaluation modes:
//  * runs  (m <= kRunMaxM): the unit of work is a z-run of run_len nodes
//    (4, 8 or 16: `run_mode` holds the run length, 0 means tile mode) on
//    `ylines` adjacent y-lines (1, 2 or 4: the lines share every occupancy
//    window they read); the fractions pass marks the run holding each closest
//    point's tricubic stencil base, N_r selection dilates the marks by the
//    stencil (+0..3 in x and y) and compacts the units holding a needed run
//    on the device, and one lane sums one unit.
//  * tiles (larger balls): 4x8x8 tiles streamed through LDS x-plane by x-plane.
#pragma once

#include <vector>

#include "common.h"
#include "fft.h"

namespace svimx {

constexpr int kTX = 4, kTY = 8, kTZ = 8;   // tile mode: one wave per x-layer of 8x8 nodes
constexpr int kRunLen = 8;                 // run mode: default nodes per z-run
constexpr int kRunLenMin = 4, kRunLenMax = 16;  // run lengths the tuner chooses from
constexpr int kNrTuneCands = 11;           // (run length, lines, LDS) shapes the autotuner times
struct NrShape {
  int rl, lines, lds;
};
constexpr NrShape kNrTuneShapes[kNrTuneCands] = {{4, 1, 0}, {8, 1, 0}, {16, 1, 0}, {4, 2, 0}, {8, 2, 0}, {16, 2, 0},
                                                 {4, 4, 0}, {8, 4, 0}, {4, 1, 1}, {8, 1, 1}, {16, 1, 1}};
constexpr int kRunMaxM = 16;               // largest ball half-width handled in run mode (= kOccPad)
// Column record of the multi-line kernels: the column's k (int bits) and its
// weights at |dz| = 0..k (the taps are mirror-symmetric in z), so one scalar
// load brings a column's tap count and weights together.
constexpr int kNrRecStride = kRunMaxM + 2;

struct NrColumn {
  int dx, dy;     // column offset
  int dz0, nz;    // first dz and number of nonzero taps (tile mode); run mode uses k = -dz0
  int woff;       // offset into the weight array
  int pad;
};

struct NrPlan {
  int m;
  double h, radius;
  Dims d;
  int run_mode;       // 0: tile mode; else the run length (4, 8 or 16)
  int lds;            // run mode: 1 stages the occupancy through LDS (same sums, same order)
  int lds_ok;         // the occupancy's float4 count fits the LDS kernel's int32 indexing
  int ylines;         // run mode: y-lines per unit (1, 2, 4)
  int nyu;            // units per x-plane column: ceil(ny / ylines)
  int64_t nunits;     // nx * nyu * nzr (the compacted list holds unit indices)
  int cols_contig;    // each dx group's columns have consecutive dy (needed for ylines > 1)
  // tile mode
  int ntx, nty, ntz;
  int64_t ntiles;
  // run mode
  int nzr;            // runs per z-line
  int64_t nruns;      // nx * ny * nzr
  int64_t nmarks;     // entries of `marks` (ntiles or nruns)
  int ncols;
  int nweights;
  float scale;        // float32(h^3)
  float negbv;        // float32(-1 / (4/3 pi r^3))
  NrColumn* cols;     // device, grouped by dx
  int* col_start;     // device, 2m+2 entries: columns of dx = g - m are [col_start[g], col_start[g+1])
  // run-mode sum order (every L1 shape): |dx| groups a = 0..m, group a > 0
  // summing the windows of planes x + a and x - a before the taps of its
  // columns (those of dx = +a); lo folds the groups a < asplit, hi the rest
  int asplit;
  int cfirst;         // col_start[m]: the first column of dx = 0
  int csplit;         // col_start[m + asplit]: the first column of the hi groups
  float* weights;     // device; run mode: per column the taps dz = -k..k
  float* colrec;      // device, ncols x kNrRecStride column records (run mode)
  uint8_t* marks;     // device, nmarks
  uint8_t* dil;       // device, nruns bytes (run mode: need flag of each unit)
  int32_t* list;      // device, nmarks
  int32_t* count;     // device, 1 (tile mode)
  int64_t* ccounts;   // device, compaction chunk counts + total (run mode)
  int64_t nchunks;
  int fft;            // 1: N_r by FFT over the whole grid (no marks, no selection)
  FftPlan fftp;
  std::vector<double> weights_host;  // the (2m+1)^3 weights, for switching to the FFT
};

// Run length and lines per unit of a run-mode plan (its arrays are sized for
// kRunLenMin and one line); nr_plan_set_run_len keeps the lines.
void nr_plan_set_shape(NrPlan* p, int rl, int ylines);
void nr_plan_set_run_len(NrPlan* p, int rl);
bool nr_shape_allowed(const NrPlan* p, const NrShape& s);

// Tile mode: mark the tiles covering the tricubic stencil [base, base+3]^3.
__device__ __forceinline__ void mark_stencil_tiles(uint8_t* marks, int nty, int ntz, int bx, int by, int bz) {
  int tx0 = bx / kTX, tx1 = (bx + 3) / kTX;
  int ty0 = by / kTY, ty1 = (by + 3) / kTY;
  int tz0 = bz / kTZ, tz1 = (bz + 3) / kTZ;
  for (int tx = tx0; tx <= tx1; ++tx)
    for (int ty = ty0; ty <= ty1; ++ty)
      for (int tz = tz0; tz <= tz1; ++tz) marks[((int64_t)tx * nty + ty) * ntz + tz] = 1;
}

// Run mode: mark the runs (of rl nodes) of the stencil base line (x = bx,
// y = by) that the stencil's z-range [bz, bz+3] touches; selection dilates in
// x and y.
__device__ __forceinline__ void mark_stencil_runs(uint8_t* marks, int ny, int nzr, int rl, int bx, int by, int bz) {
  int64_t line = ((int64_t)bx * ny + by) * nzr;
  marks[line + bz / rl] = 1;
  marks[line + (bz + 3) / rl] = 1;
}

}  // namespace svimx
