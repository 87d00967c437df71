// Device body of the run-mode N_r summation, shared by the single-job kernel
// (nr.hip) and the job-table kernel of the batch engine (engine.hip).
#pragma once

#include "common.h"
#include "nr_plan.h"

namespace svimx {

// ---------------------------------------------------------------- run selection
// dil(x, y, zr) = OR_{b=0..3} marks(x, y-b, zr)
__device__ __forceinline__ void dilate_y_run(const uint8_t* __restrict__ marks, uint8_t* __restrict__ dil, int ny,
                                             int nzr, int64_t r) {
  const int y = (int)((r / nzr) % ny);
  uint8_t v = 0;
  for (int b = 0; b <= 3 && b <= y; ++b) v |= marks[r - (int64_t)b * nzr];
  dil[r] = v;
}

// need(x, y, zr) = OR_{a=0..3} dil(x-a, y, zr)
struct RunNeed {
  const uint8_t* dil;
  int64_t plane;  // ny * nzr
  __device__ bool operator()(int64_t r) const {
    const int64_t x = r / plane;
    uint8_t v = 0;
    for (int a = 0; a <= 3 && a <= x; ++a) v |= dil[r - a * plane];
    return v != 0;
  }
};

// ---------------------------------------------------------------- run-mode summation
template <int K, int MP, int RL, int WIN>
__device__ __forceinline__ void column_taps(const float* __restrict__ wc, const float (&win)[WIN],
                                            float (&part)[RL]) {
  if constexpr (K <= MP) {
#pragma unroll
    for (int t = 0; t <= 2 * K; ++t) {
      const float wt = wc[t];
#pragma unroll
      for (int j = 0; j < RL; ++j) part[j] += wt * win[j + MP - K + t];
    }
  }
}

// One lane per z-run (RL nodes) of the run list; lb is the (XCD-remapped)
// block index among G blocks of blockDim.x threads.
template <int MP, int RL>
__device__ __forceinline__ void nr_runs_body(const float* __restrict__ occ, const OccLayout& OL, int nzr,
                                             const int32_t* __restrict__ list, int64_t nrun,
                                             const NrColumn* __restrict__ cols, int ncols,
                                             const float* __restrict__ w, float scale, float negbv,
                                             float* __restrict__ out, int64_t lb, int64_t G) {
  constexpr int WIN = RL + 2 * MP;
  const int64_t stride = G * blockDim.x;
  for (int64_t first = lb * blockDim.x; first < nrun; first += stride) {
    const int64_t it = first + threadIdx.x;
    const bool live = it < nrun;
    const int32_t r = list[live ? it : first];
    const int zr = r % nzr;
    const int64_t xy = r / nzr;
    const int y = (int)(xy % OL.ny), x = (int)(xy / OL.ny);
    const int z0 = zr * RL;
    float acc[RL], part[RL];
#pragma unroll
    for (int j = 0; j < RL; ++j) acc[j] = part[j] = 0.0f;
    int cur_dx = cols[0].dx;
    for (int c = 0; c < ncols; ++c) {
      const NrColumn col = cols[c];
      if (col.dx != cur_dx) {
#pragma unroll
        for (int j = 0; j < RL; ++j) {
          acc[j] += part[j];
          part[j] = 0.0f;
        }
        cur_dx = col.dx;
      }
      const float4* src = reinterpret_cast<const float4*>(occ + OL.at(x + col.dx, y + col.dy, z0 - MP));
      float win[WIN];
#pragma unroll
      for (int q = 0; q < WIN / 4; ++q) {
        const float4 v = src[q];
        win[4 * q] = v.x;
        win[4 * q + 1] = v.y;
        win[4 * q + 2] = v.z;
        win[4 * q + 3] = v.w;
      }
      const float* wc = w + col.woff;
      switch (-col.dz0) {
        case 0: column_taps<0, MP, RL>(wc, win, part); break;
        case 1: column_taps<1, MP, RL>(wc, win, part); break;
        case 2: column_taps<2, MP, RL>(wc, win, part); break;
        case 3: column_taps<3, MP, RL>(wc, win, part); break;
        case 4: column_taps<4, MP, RL>(wc, win, part); break;
        case 5: column_taps<5, MP, RL>(wc, win, part); break;
        case 6: column_taps<6, MP, RL>(wc, win, part); break;
        case 7: column_taps<7, MP, RL>(wc, win, part); break;
        case 8: column_taps<8, MP, RL>(wc, win, part); break;
        case 9: column_taps<9, MP, RL>(wc, win, part); break;
        case 10: column_taps<10, MP, RL>(wc, win, part); break;
        case 11: column_taps<11, MP, RL>(wc, win, part); break;
        case 12: column_taps<12, MP, RL>(wc, win, part); break;
        case 13: column_taps<13, MP, RL>(wc, win, part); break;
        case 14: column_taps<14, MP, RL>(wc, win, part); break;
        case 15: column_taps<15, MP, RL>(wc, win, part); break;
        case 16: column_taps<16, MP, RL>(wc, win, part); break;
        default: break;
      }
    }
    if (!live) continue;
    const int64_t base = xy * OL.nz + z0;
#pragma unroll
    for (int j = 0; j < RL; ++j)
      if (z0 + j < OL.nz) out[base + j] = ((acc[j] + part[j]) * scale) * negbv + 1.0f;
  }
}

// Block remap giving XCD x (blocks go round-robin over the 8 XCDs, each with
// its own L2) a contiguous slice of a G-block range: G must be a multiple of 8
// and the range must start at a block index that is a multiple of 8.
__device__ __forceinline__ int64_t xcd_slice(int64_t b, int64_t G) { return (b % kXcds) * (G / kXcds) + b / kXcds; }

}  // namespace svimx
