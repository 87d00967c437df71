// Device body of the run-mode N_r summation, shared by the single-job kernel
// (nr.hip) and the job-table kernel of the batch engine (engine.hip).
#pragma once

#include <climits>

#include "common.h"
#include "nr_plan.h"

namespace svimx {

// ---------------------------------------------------------------- run selection
// Need flag of unit u = (x, yu, zr): some run (x - a, y', zr) with a in [0, 3]
// and y' in [yu*Y - 3, yu*Y + Y - 1] is marked, i.e. a marked tricubic
// stencil base (+0..3 in x and y) reaches one of the unit's Y lines.  Unit
// and run indices fit 31 bits: 32-bit divisions.
__device__ __forceinline__ void unit_need(const uint8_t* __restrict__ marks, uint8_t* __restrict__ need, int ny,
                                          int nzr, int nyu, int Y, int64_t u) {
  const uint32_t xyu = (uint32_t)u / (uint32_t)nzr;
  const int zr = (int)((uint32_t)u - xyu * (uint32_t)nzr);
  const int x = (int)(xyu / (uint32_t)nyu);
  const int yu = (int)(xyu - (uint32_t)x * (uint32_t)nyu);
  const int ylo = max(yu * Y - 3, 0), yhi = min(yu * Y + Y - 1, ny - 1);
  uint8_t v = 0;
  for (int a = 0; a <= 3 && a <= x; ++a) {
    const uint8_t* row = marks + (int64_t)(x - a) * ny * nzr + zr;
    for (int y = ylo; y <= yhi; ++y) v |= row[(int64_t)y * nzr];
  }
  need[u] = v ? 1 : 0;
}

// The same flags for the four units 4q .. 4q+3 of one (x, yu) row, with
// 32-bit loads (nzr a multiple of 4; marks and need 4-byte aligned): the 4·(Y+3)
// stencil loads are fixed-trip and predicated, so all of them are in flight
// at once; the word OR is the byte-wise OR, and each byte becomes 0/1.
template <int Y>
__device__ __forceinline__ void unit_need4(const uint8_t* __restrict__ marks, uint8_t* __restrict__ need, int ny,
                                           int nzr, int nyu, int64_t q) {
  const uint32_t u = (uint32_t)q * 4u;
  const uint32_t xyu = u / (uint32_t)nzr;
  const int zr = (int)(u - xyu * (uint32_t)nzr);
  const int x = (int)(xyu / (uint32_t)nyu);
  const int yu = (int)(xyu - (uint32_t)x * (uint32_t)nyu);
  uint32_t v = 0;
#pragma unroll
  for (int a = 0; a <= 3; ++a) {
    const uint8_t* row = marks + (int64_t)(x - a) * ny * nzr + zr;
#pragma unroll
    for (int k = 0; k < Y + 3; ++k) {
      const int y = yu * Y - 3 + k;
      if (a <= x && y >= 0 && y < ny) v |= *reinterpret_cast<const uint32_t*>(row + (int64_t)y * nzr);
    }
  }
  const uint32_t t = ((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v;
  *reinterpret_cast<uint32_t*>(need + u) = (t >> 7) & 0x01010101u;
}

// Need flags of units [4 q0, ...) for quads q = q0 + i·stride < nunits / 4
// when nzr % 4 == 0, else of single units u = q0 + i·stride < nunits.
__device__ __forceinline__ void unit_need_range(const uint8_t* __restrict__ marks, uint8_t* __restrict__ need,
                                                int ny, int nzr, int nyu, int Y, int64_t nunits, int64_t q0,
                                                int64_t stride) {
  if ((nzr & 3) == 0) {
    const int64_t nq = nunits >> 2;
    if (Y == 4)
      for (int64_t q = q0; q < nq; q += stride) unit_need4<4>(marks, need, ny, nzr, nyu, q);
    else if (Y == 2)
      for (int64_t q = q0; q < nq; q += stride) unit_need4<2>(marks, need, ny, nzr, nyu, q);
    else
      for (int64_t q = q0; q < nq; q += stride) unit_need4<1>(marks, need, ny, nzr, nyu, q);
  } else {
    for (int64_t u = q0; u < nunits; u += stride) unit_need(marks, need, ny, nzr, nyu, Y, u);
  }
}

// Wave-level split of an n-entry work list over a launch of G blocks: the
// work is cut into 64-entry wave items, each XCD (block lb runs on XCD lb mod
// 8; G is a multiple of 8) takes a contiguous eighth of the items actually
// present (balanced whatever the launch was sized for, and neighbouring
// items share that XCD's L2), and the XCD's waves stride over its share.
struct WaveSplit {
  int64_t first, end, step;  // item range [first, end) of this wave, stride step
};
__device__ __forceinline__ WaveSplit wave_split(int64_t n, int64_t lb, int64_t G) {
  const int64_t wpb = blockDim.x >> 6;              // waves per block
  const int64_t items = (n + 63) >> 6;
  const int64_t per = (items + kXcds - 1) / kXcds;  // items per XCD
  const int64_t x = lb % kXcds, lbx = lb / kXcds;
  WaveSplit s;
  s.first = x * per + lbx * wpb + (threadIdx.x >> 6);
  s.end = min(x * per + per, items);
  s.step = (G / kXcds) * wpb;
  return s;
}

// ---------------------------------------------------------------- run-mode summation
// A column's sum over its 2k+1 z taps in the symmetric order that every
// run-mode shape uses: the weights are mirror-symmetric in z (checked when the
// plan is built), so for d = k .. 1 the pair win[j - d] + win[j + d] takes
// one FMA with the weight at |dz| = d, and the centre tap comes last: k adds
// and k + 1 FMAs per output instead of 2k + 1 FMAs.  Kernels that apply
// several columns to one window form the pair sums once (column_taps_ps);
// the sums are the same floats either way, so every shape agrees bit for bit.
template <int K, int MP, int RL, int WIN>
__device__ __forceinline__ void column_taps(const float* __restrict__ wc, const float (&win)[WIN],
                                            float (&part)[RL]) {
  if constexpr (K <= MP) {
#pragma unroll
    for (int d = K; d >= 1; --d) {
      const float wt = wc[K + d];
#pragma unroll
      for (int j = 0; j < RL; ++j) part[j] = __builtin_fmaf(wt, win[j + MP - d] + win[j + MP + d], part[j]);
    }
    const float w0 = wc[K];
#pragma unroll
    for (int j = 0; j < RL; ++j) part[j] = __builtin_fmaf(w0, win[j + MP], part[j]);
  }
}

// All taps of one window in the multi-line kernels: the window feeds the
// columns of lines yy = ylo .. yhi (one column each, records r[yy]), whose
// half-heights are at most K.  For d = K .. 1 the pair sums of the window are
// formed once and each column takes its FMA with the weight at |dz| = d --
// 0 where d exceeds the column's own half-height (records are zero-filled
// past k), and fma(0, p, s) = s exactly for the finite, non-negative pair
// sums and partials here -- then the centre taps.  Per column these are the
// operations of column_taps in its order, so the sums equal the one-line
// kernels' bit for bit.  ALL: every line of the window has a column (the
// windows inside a group), no per-line test.
template <int K, bool ALL, int MP, int RL, int Y, int WIN>
__device__ __forceinline__ void window_taps_k(const float (&r)[Y][MP + 2], int ylo, int yhi,
                                              const float (&win)[WIN], float (&part)[Y][RL]) {
  if constexpr (K <= MP) {
#pragma unroll
    for (int d = K; d >= 1; --d) {
      float pr[RL];
#pragma unroll
      for (int j = 0; j < RL; ++j) pr[j] = win[j + MP - d] + win[j + MP + d];
#pragma unroll
      for (int yy = 0; yy < Y; ++yy)
        if (ALL || (yy >= ylo && yy <= yhi))
#pragma unroll
          for (int j = 0; j < RL; ++j) part[yy][j] = __builtin_fmaf(r[yy][1 + d], pr[j], part[yy][j]);
    }
#pragma unroll
    for (int yy = 0; yy < Y; ++yy)
      if (ALL || (yy >= ylo && yy <= yhi))
#pragma unroll
        for (int j = 0; j < RL; ++j) part[yy][j] = __builtin_fmaf(r[yy][1], win[j + MP], part[yy][j]);
  }
}

template <bool ALL, int MP, int RL, int Y>
__device__ __forceinline__ void window_taps(int k, const float (&r)[Y][MP + 2], int ylo, int yhi,
                                            const float (&win)[RL + 2 * MP], float (&part)[Y][RL]) {
  switch (k) {
#define SVIMX_WT(K) case K: window_taps_k<K, ALL, MP, RL, Y>(r, ylo, yhi, win, part); break;
    SVIMX_WT(0) SVIMX_WT(1) SVIMX_WT(2) SVIMX_WT(3) SVIMX_WT(4) SVIMX_WT(5) SVIMX_WT(6) SVIMX_WT(7) SVIMX_WT(8)
    SVIMX_WT(9) SVIMX_WT(10) SVIMX_WT(11) SVIMX_WT(12) SVIMX_WT(13) SVIMX_WT(14) SVIMX_WT(15) SVIMX_WT(16)
#undef SVIMX_WT
    default: break;
  }
}

constexpr int kNrLineLanes = 2;  // waves per unit group of the multi-line kernels (1 or 2)

// Order of the run-mode sums read through the L1 (every such shape, so they
// agree bit for bit and the autotuner's pick never changes a result): the
// columns of dx = +a (a = |dx| = 0..m) form group a; for a > 0 each window of
// plane x + a is added element-wise to the same window of plane x - a before
// the group's taps (the ball is mirror-symmetric in x, so the +a columns'
// weights serve both planes: half the FMAs); a group's columns are summed
// into a fresh partial, lo folds the partials of the groups a < asplit in
// order starting from 0.0, hi those of the groups a >= asplit, and N_r comes
// from lo + hi.

// Y lines per unit: a lane sums the RL-node z-runs of Y adjacent y-lines.
// Window line y0 + dy' of plane x + dx feeds column (dx, dy' - yy) of line yy,
// so every line's columns are applied in the order of the one-line kernel
// (bit-identical sums) while each window is loaded once for the Y lines.
// Needs each dx group's columns to have consecutive dy.  S = 2 gives each
// unit two waves, one per half of the order (twice the waves for the same
// loads, for lists that leave most SIMDs with 2-3 waves).
// The window z0 - MP .. z0 + RL + MP - 1 of line y in plane x + a, plus (for
// a > 0) the same window of plane x - a, as aligned float4 loads.
template <int MP, int RL>
__device__ __forceinline__ void nr_window(const float* __restrict__ occ, const OccLayout& OL, int x, int a, int y,
                                          int z0, float (&win)[RL + 2 * MP]) {
  const float4* src = reinterpret_cast<const float4*>(occ + OL.at(x + a, y, z0 - MP));
#pragma unroll
  for (int q = 0; q < (RL + 2 * MP) / 4; ++q) {
    const float4 v = src[q];
    win[4 * q] = v.x;
    win[4 * q + 1] = v.y;
    win[4 * q + 2] = v.z;
    win[4 * q + 3] = v.w;
  }
  if (a > 0) {
    const float4* mir = reinterpret_cast<const float4*>(occ + OL.at(x - a, y, z0 - MP));
#pragma unroll
    for (int q = 0; q < (RL + 2 * MP) / 4; ++q) {
      const float4 v = mir[q];
      win[4 * q] += v.x;
      win[4 * q + 1] += v.y;
      win[4 * q + 2] += v.z;
      win[4 * q + 3] += v.w;
    }
  }
}

// Sum of the |dx| groups [ab, ae) of the unit's Y lines, folded group by
// group from 0.0 (one half of the canonical order).  The windows of the
// groups form one sequence (group a, window t = 0 .. ncol + Y - 2).  With PF
// the loads of the next window are in flight while the taps of the current
// one run (two register sets, the loop unrolled twice so their roles
// alternate without copies): -5% on 4-node runs of 4 lines at m = 8 (0.218 ->
// 0.205 ms at the 1.3M-triangle scan), where the second set still leaves 3
// waves per SIMD; wider windows would spill, so the kernels enable it for
// MP <= 8 and 4-node runs only.  The column records of a window's lines are
// loaded before any of their taps.  Window offsets are 32-bit float4
// indices (plans whose occupancy has 2^31 float4s or more take the one-line
// shapes).
struct NrPos {
  int a, t, c0, ncol, dyf;
};

template <int MP, int RL, int Y, bool PF>
__device__ __forceinline__ void nr_lines_half(const float* __restrict__ occ, const OccLayout& OL,
                                              const NrColumn* __restrict__ cols, const int* __restrict__ col_start,
                                              const float* __restrict__ colrec, int x, int y0, int z0, int m, int ab,
                                              int ae, float (&sum)[Y][RL]) {
  constexpr int WIN = RL + 2 * MP, Q = WIN / 4;
  const float4* __restrict__ occ4 = reinterpret_cast<const float4*>(occ);
  const int32_t sx4 = (int32_t)(OL.pitch_y * OL.pitch_z / 4), sz4 = (int32_t)(OL.pitch_z / 4);
  const int32_t ub4 = (int32_t)(OL.at(x, y0, z0 - MP) / 4);
  float acc[Y][RL], part[Y][RL];
#pragma unroll
  for (int yy = 0; yy < Y; ++yy)
#pragma unroll
    for (int j = 0; j < RL; ++j) acc[yy][j] = part[yy][j] = 0.0f;
  auto group = [&](NrPos& p) {  // first window of the first group >= p.a with columns
    while (p.a < ae && col_start[m + p.a + 1] == col_start[m + p.a]) ++p.a;
    p.t = 0;
    if (p.a < ae) {
      p.c0 = col_start[m + p.a];
      p.ncol = col_start[m + p.a + 1] - p.c0;
      p.dyf = cols[p.c0].dy;
    }
  };
  auto advance = [&](NrPos p) {
    if (++p.t >= p.ncol + Y - 1) {
      ++p.a;
      group(p);
    }
    return p;
  };
  auto issue = [&](const NrPos& p, float4 (&P)[Q], float4 (&M)[Q]) {
    const int32_t o = ub4 + (p.dyf + p.t) * sz4;
    const float4* __restrict__ src = occ4 + (uint32_t)(o + p.a * sx4);  // the window's float4s at immediate offsets
#pragma unroll
    for (int q = 0; q < Q; ++q) P[q] = src[q];
    if (p.a > 0) {
      const float4* __restrict__ mir = occ4 + (uint32_t)(o - p.a * sx4);
#pragma unroll
      for (int q = 0; q < Q; ++q) M[q] = mir[q];
    }
  };
  bool first_group = true;
  auto taps = [&](const NrPos& p, const float4 (&P)[Q], const float4 (&M)[Q]) {
    float win[WIN];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      win[4 * q] = P[q].x;
      win[4 * q + 1] = P[q].y;
      win[4 * q + 2] = P[q].z;
      win[4 * q + 3] = P[q].w;
    }
    if (p.a > 0) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        win[4 * q] += M[q].x;
        win[4 * q + 1] += M[q].y;
        win[4 * q + 2] += M[q].z;
        win[4 * q + 3] += M[q].w;
      }
    }
    if (p.t == 0) {
      if (!first_group) {
#pragma unroll
        for (int yy = 0; yy < Y; ++yy)
#pragma unroll
          for (int j = 0; j < RL; ++j) {
            acc[yy][j] += part[yy][j];
            part[yy][j] = 0.0f;
          }
      }
      first_group = false;
    }
    // the records of the window's columns (lines ylo .. yhi, one column
    // each), all loads issued first
    const int ylo = max(0, p.t - p.ncol + 1), yhi = min(Y - 1, p.t);
    float r[Y][MP + 2];
    int kmax = 0;
#pragma unroll
    for (int yy = 0; yy < Y; ++yy) {
      if (yy >= ylo && yy <= yhi) {
        const float* rec = colrec + (int64_t)(p.c0 + p.t - yy) * kNrRecStride;
#pragma unroll
        for (int i = 0; i < MP + 2; ++i) r[yy][i] = rec[i];
        kmax = max(kmax, __float_as_int(r[yy][0]));
      }
    }
    if (ylo == 0 && yhi == Y - 1)
      window_taps<true, MP, RL, Y>(kmax, r, ylo, yhi, win, part);
    else
      window_taps<false, MP, RL, Y>(kmax, r, ylo, yhi, win, part);
  };
  NrPos cur{ab, 0, 0, 0, 0};
  group(cur);
  if constexpr (!PF) {
    for (; cur.a < ae; ++cur.a) {
      cur.c0 = col_start[m + cur.a];
      cur.ncol = col_start[m + cur.a + 1] - cur.c0;
      if (cur.ncol == 0) continue;
      cur.dyf = cols[cur.c0].dy;
      for (cur.t = 0; cur.t < cur.ncol + Y - 1; ++cur.t) {
        float4 PA[Q], MA[Q];
        issue(cur, PA, MA);
        taps(cur, PA, MA);
      }
    }
  } else if (cur.a < ae) {
    float4 PA[Q], MA[Q], PB[Q], MB[Q];
    issue(cur, PA, MA);
    while (true) {
      NrPos nxt = advance(cur);
      bool more = nxt.a < ae;
      if (more) issue(nxt, PB, MB);
      taps(cur, PA, MA);
      if (!more) break;
      cur = nxt;
      nxt = advance(cur);
      more = nxt.a < ae;
      if (more) issue(nxt, PA, MA);
      taps(cur, PB, MB);
      if (!more) break;
      cur = nxt;
    }
  }
#pragma unroll
  for (int yy = 0; yy < Y; ++yy)
#pragma unroll
    for (int j = 0; j < RL; ++j) sum[yy][j] = acc[yy][j] + part[yy][j];
}

// LDS floats the multi-line body needs per block of 256 threads.
template <int RL, int Y, int S>
constexpr int nr_lines_xch() { return S == 2 ? Y * RL * 128 : 1; }

// PF: software-pipelined window loads (see nr_lines_half), for kernels whose
// register budget allows them.
template <int MP, int RL>
constexpr bool nr_prefetch() { return MP <= 8 && RL == 4; }

template <int MP, int RL, int Y, int S, bool PF = nr_prefetch<MP, RL>()>
__device__ __forceinline__ void nr_lines_body(const float* __restrict__ occ, const OccLayout& OL, int nzr,
                                              const int32_t* __restrict__ list, int64_t nunit,
                                              const NrColumn* __restrict__ cols, const int* __restrict__ col_start,
                                              int m, int asplit, const float* __restrict__ colrec, float scale,
                                              float negbv, float* __restrict__ out, float* __restrict__ xch,
                                              int64_t lb, int64_t G) {
  static_assert(S == 1 || (S == 2 && Y % 2 == 0), "the two halves of a unit store half of its lines each");
  const int lane = threadIdx.x & 63;
  const uint32_t nyu = (uint32_t)((OL.ny + Y - 1) / Y);
  if constexpr (S == 1) {
    const WaveSplit ws = wave_split(nunit, lb, G);
    for (int64_t item = ws.first; item < ws.end; item += ws.step) {
      const int64_t it = item * 64 + lane;
      const bool live = it < nunit;
      const uint32_t u = (uint32_t)list[live ? it : item * 64];
      const uint32_t xyu = u / (uint32_t)nzr;
      const int zr = (int)(u - xyu * (uint32_t)nzr);
      const int x = (int)(xyu / nyu), y0 = (int)(xyu - (uint32_t)x * nyu) * Y;
      const int z0 = zr * RL;
      float lo[Y][RL], hi[Y][RL];
      nr_lines_half<MP, RL, Y, PF>(occ, OL, cols, col_start, colrec, x, y0, z0, m, 0, asplit, lo);
      nr_lines_half<MP, RL, Y, PF>(occ, OL, cols, col_start, colrec, x, y0, z0, m, asplit, m + 1, hi);
      if (!live) continue;
#pragma unroll
      for (int yy = 0; yy < Y; ++yy) {
        if (y0 + yy >= OL.ny) break;
        const int64_t base = ((int64_t)x * OL.ny + y0 + yy) * OL.nz + z0;
#pragma unroll
        for (int j = 0; j < RL; ++j)
          if (z0 + j < OL.nz) out[base + j] = ((lo[yy][j] + hi[yy][j]) * scale) * negbv + 1.0f;
      }
    }
  } else {
    // A block (4 waves) takes 128 units: waves 0-1 sum their lo halves, waves
    // 2-3 the hi halves of the same units (each wave's loop bounds stay
    // uniform).  The lo wave stores lines 0..YH-1, the hi wave lines YH..Y-1;
    // each hands the other half's sums of the partner's lines over LDS, laid
    // out [half][k][unit] (consecutive lanes, consecutive banks).
    constexpr int YH = Y / 2, K = YH * RL;
    // readfirstlane: the compiler cannot see that the wave index is uniform,
    // and the half decides the (scalar) column loop
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int half = wave >> 1, slot = (wave & 1) * 64 + lane;
    const int64_t items = (nunit + 127) >> 7;
    const int64_t per = (items + kXcds - 1) / kXcds;  // contiguous items per XCD
    const int64_t xcd = lb % kXcds;
    const int64_t end = min(xcd * per + per, items);
    for (int64_t item = xcd * per + lb / kXcds; item < end; item += G / kXcds) {  // block-uniform
      const int64_t it = item * 128 + slot;
      const bool live = it < nunit;
      const uint32_t u = (uint32_t)list[live ? it : item * 128];
      const uint32_t xyu = u / (uint32_t)nzr;
      const int zr = (int)(u - xyu * (uint32_t)nzr);
      const int x = (int)(xyu / nyu), y0 = (int)(xyu - (uint32_t)x * nyu) * Y;
      const int z0 = zr * RL;
      float own[Y][RL];
      nr_lines_half<MP, RL, Y, PF>(occ, OL, cols, col_start, colrec, x, y0, z0, m, half ? asplit : 0,
                               half ? m + 1 : asplit, own);
#pragma unroll
      for (int yl = 0; yl < YH; ++yl)
#pragma unroll
        for (int j = 0; j < RL; ++j)
          xch[(half * K + yl * RL + j) * 128 + slot] = half ? own[yl][j] : own[YH + yl][j];
      __syncthreads();
      float tot[YH][RL];
#pragma unroll
      for (int yl = 0; yl < YH; ++yl)
#pragma unroll
        for (int j = 0; j < RL; ++j) {
          const float other = xch[((1 - half) * K + yl * RL + j) * 128 + slot];
          tot[yl][j] = half ? other + own[YH + yl][j] : own[yl][j] + other;  // lo + hi
        }
      __syncthreads();  // the exchange buffer is rewritten by the next item
      if (!live) continue;
#pragma unroll
      for (int yl = 0; yl < YH; ++yl) {
        const int yy = half * YH + yl;
        if (y0 + yy >= OL.ny) continue;
        const int64_t base = ((int64_t)x * OL.ny + y0 + yy) * OL.nz + z0;
#pragma unroll
        for (int j = 0; j < RL; ++j)
          if (z0 + j < OL.nz) out[base + j] = (tot[yl][j] * scale) * negbv + 1.0f;
      }
    }
  }
}

// One lane per z-run (RL nodes) of the run list; lb is the block index among
// G blocks of blockDim.x threads (work split by wave_split).
template <int MP, int RL>
__device__ __forceinline__ void nr_runs_body(const float* __restrict__ occ, const OccLayout& OL, int nzr,
                                             const int32_t* __restrict__ list, int64_t nrun,
                                             const NrColumn* __restrict__ cols, int ncols, int cfirst,
                                             int csplit, const float* __restrict__ w, float scale, float negbv,
                                             float* __restrict__ out, int64_t lb, int64_t G) {
  constexpr int WIN = RL + 2 * MP;
  const WaveSplit ws = wave_split(nrun, lb, G);
  for (int64_t item = ws.first; item < ws.end; item += ws.step) {
    const int64_t it = item * 64 + (threadIdx.x & 63);
    const bool live = it < nrun;
    const int32_t r = list[live ? it : item * 64];
    const int zr = r % nzr;
    const int64_t xy = r / nzr;
    const int x = (int)((uint32_t)xy / (uint32_t)OL.ny), y = (int)((uint32_t)xy - (uint32_t)x * (uint32_t)OL.ny);
    const int z0 = zr * RL;
    float lo[RL], acc[RL], part[RL];
#pragma unroll
    for (int j = 0; j < RL; ++j) lo[j] = acc[j] = part[j] = 0.0f;
    int cur_dx = cols[cfirst].dx;
    for (int c = cfirst; c < ncols; ++c) {  // the columns of dx >= 0
      const NrColumn col = cols[c];
      if (c == csplit) {  // first hi group: lo is complete
#pragma unroll
        for (int j = 0; j < RL; ++j) {
          lo[j] = acc[j] + part[j];
          acc[j] = part[j] = 0.0f;
        }
        cur_dx = col.dx;
      } else if (col.dx != cur_dx) {
#pragma unroll
        for (int j = 0; j < RL; ++j) {
          acc[j] += part[j];
          part[j] = 0.0f;
        }
        cur_dx = col.dx;
      }
      float win[WIN];
      nr_window<MP, RL>(occ, OL, x, col.dx, y + col.dy, z0, win);
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
      if (z0 + j < OL.nz) out[base + j] = ((lo[j] + (acc[j] + part[j])) * scale) * negbv + 1.0f;
  }
}


// ---------------------------------------------------------------- LDS-staged run summation
// The one-line run sums with the occupancy staged through LDS, one dx plane at
// a time, so in the plain order: the dx groups -m..m folded in turn from 0.0
// (the same value as the L1 shapes up to rounding; an explicit option that
// the autotuner does not pick).  Per wave of 64 runs, the runs are
// grouped into up to kNrLdsClusters clusters (same x, z-runs within +-Dz and
// lines within +-reach of the first pending run; the first cluster of a pass
// shrinks its reach until it fits, down to one line), each with a bounding
// slab of rows (y - m .. y + m) x (z windows).  Every lane stages up to
// kNrLdsSlots float4s of the slabs per dx plane: the next plane's loads are
// issued into registers before the current plane is summed from LDS.  Within
// a slab row, float4 chunk c sits at (c mod S) * half + c / S (S = float4s
// per run), so the lanes of one line read consecutive chunks (no bank
// conflicts).  Runs that do not fit in this pass's slabs are taken next.
constexpr int kNrLdsCap4 = 1000;                     // float4 per wave (16 KB): 9 waves per CU
constexpr int kNrLdsSlots = 16;                      // float4s each lane stages per plane
static_assert(kNrLdsSlots * 64 >= kNrLdsCap4, "the slots must cover the slab");
#define NR_LDS_SLOTS(F) F(0) F(1) F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13) F(14) F(15)
constexpr int kNrLdsClusters = 4;
constexpr int kNrLdsDy = 40;                   // y reach of a cluster from its first run
constexpr int nr_lds_dz(int rl) { return rl >= 24 ? 1 : 24 / rl; }  // z reach in runs
// LDS float4s of a slab row holding sz4 chunks in the swizzled layout
__host__ __device__ constexpr int nr_lds_pitch(int sz4, int s) { return (s * ((sz4 + s - 1) / s)) | 1; }

struct NrLdsClu {
  int g4;    // float4 index of the slab origin (x, ylo - m, zlo * RL - MP) in the occupancy
  int sz4;   // float4 chunks per row
  int p4;    // LDS row pitch (float4)
  int half;  // chunks per swizzle group
  int base;  // LDS offset (float4)
  int q0;    // first staging index of this cluster
  int pad0, pad1;
};

__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Orders this wave's LDS accesses across lanes (a wavefront executes its LDS
// instructions in order; the fences stop the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Taps of one column with half-height K from a swizzled slab row: only the
// float4s that the taps touch are read.
template <int K, int MP, int RL>
__device__ __forceinline__ void lds_column(const float4* __restrict__ row, int half, const float* __restrict__ wc,
                                           float (&part)[RL]) {
  if constexpr (K <= MP) {
    constexpr int WIN = RL + 2 * MP, S = RL / 4;
    constexpr int q0 = (MP - K) / 4, q1 = (MP + RL + K - 1) / 4;
    float win[WIN];
#pragma unroll
    for (int q = q0; q <= q1; ++q) {
      const float4 v = row[(q % S) * half + q / S];
      win[4 * q] = v.x;
      win[4 * q + 1] = v.y;
      win[4 * q + 2] = v.z;
      win[4 * q + 3] = v.w;
    }
    column_taps<K, MP, RL>(wc, win, part);
  }
}

// slab: this wave's kNrLdsCap4 float4s; clu: its kNrLdsClusters records.
// wv / NW: this one-wave block's index and the block count of the launch
// (work split by wave_split).  The
// occupancy's float4 count must be below 2^31.
template <int MP, int RL>
__device__ __forceinline__ void nr_lds_body(const float* __restrict__ occ, const OccLayout& OL, int nzr,
                                            const int32_t* __restrict__ list, int64_t nrun,
                                            const NrColumn* __restrict__ cols, const int* __restrict__ col_start,
                                            int m, const float* __restrict__ w, float scale, float negbv,
                                            float* __restrict__ out, float4* __restrict__ slab,
                                            NrLdsClu* __restrict__ clu, int64_t wv, int64_t NW) {
  constexpr int DZ = nr_lds_dz(RL), S = RL / 4;
  static_assert((2 * MP + 1) * nr_lds_pitch(((2 * DZ + 1) * RL + 2 * MP) / 4, S) <= kNrLdsCap4,
                "a one-line cluster must fit the slab");
  const int lane = threadIdx.x & 63;
  const int gp4 = (int)(OL.pitch_z / 4);                  // float4s per occupancy row
  const int plane4 = (int)(OL.pitch_y * OL.pitch_z / 4);  // float4s per occupancy x-plane
  const float4* __restrict__ occ4 = reinterpret_cast<const float4*>(occ);
  int gfirst = 0;
  while (col_start[gfirst] == col_start[gfirst + 1]) ++gfirst;  // first plane with columns
  const WaveSplit ws = wave_split(nrun, wv, NW);
  for (int64_t item = ws.first; item < ws.end; item += ws.step) {
    const int64_t it = item * 64 + lane;
    const bool live = it < nrun;
    const int32_t r = list[live ? it : item * 64];
    const int zr = r % nzr;
    const int64_t xy = r / nzr;
    const int x = (int)((uint32_t)xy / (uint32_t)OL.ny), y = (int)((uint32_t)xy - (uint32_t)x * (uint32_t)OL.ny);
    bool pending = live;
    while (__any(pending)) {
      // ---- clusters of this pass
      int cid = -1, used = 0, nclu = 0, qtot = 0;
      int qs1 = INT_MAX, qs2 = INT_MAX, qs3 = INT_MAX;  // staging starts of clusters 1..3
      int my_ylo = 0, my_zlo = 0, my_p4 = 1, my_base = 0, my_half = 0;
      while (nclu < kNrLdsClusters) {
        const bool fr = pending && cid < 0;
        const unsigned long long fm = __ballot(fr);
        if (fm == 0ull) break;
        const int p0 = __ffsll(fm) - 1;
        const int x0 = __shfl(x, p0, 64), y0 = __shfl(y, p0, 64), z0 = __shfl(zr, p0, 64);
        int reach = kNrLdsDy;
        bool cand;
        int ylo, yhi, zlo, sz4, p4, sy;
        for (;;) {
          cand = fr && x == x0 && abs(zr - z0) <= DZ && abs(y - y0) <= reach;
          ylo = wave_min_i(cand ? y : INT_MAX);
          yhi = wave_max_i(cand ? y : INT_MIN);
          zlo = wave_min_i(cand ? zr : INT_MAX);
          const int zhi = wave_max_i(cand ? zr : INT_MIN);
          sz4 = ((zhi - zlo + 1) * RL + 2 * MP) / 4;
          p4 = nr_lds_pitch(sz4, S);
          sy = yhi - ylo + 1 + 2 * m;
          if (used + sy * p4 <= kNrLdsCap4 || nclu > 0 || reach == 0) break;
          reach >>= 1;
        }
        if (used + sy * p4 > kNrLdsCap4) break;  // only after the first cluster
        const int half = (sz4 + S - 1) / S;
        if (cand) {
          cid = nclu;
          my_ylo = ylo;
          my_zlo = zlo;
          my_p4 = p4;
          my_base = used;
          my_half = half;
        }
        if (nclu == 1) qs1 = qtot;
        if (nclu == 2) qs2 = qtot;
        if (nclu == 3) qs3 = qtot;
        if (lane == 0)
          clu[nclu] = NrLdsClu{(int)(OL.at(x0, ylo - m, zlo * RL - MP) / 4), sz4, p4, half, used, qtot, 0, 0};
        used += sy * p4;
        qtot += sy * sz4;
        ++nclu;
      }
      wave_lds_sync();  // the cluster records are visible to every lane
      // ---- staging addresses of this pass (the plane offset is added per dx);
      // slots past the slabs load cluster 0's origin, which is in bounds, and
      // store nothing.  Named registers (not an array) keep the staged float4s
      // out of scratch memory.
      const int g4_0 = clu[0].g4;
#define NR_SLOT_DECL(i) \
  float4 stg##i;        \
  int src##i, dst##i;
      NR_LDS_SLOTS(NR_SLOT_DECL)
#undef NR_SLOT_DECL
#define NR_SLOT_ADDR(i)                                                                   \
  {                                                                                       \
    const int q = lane + 64 * i;                                                          \
    const int c = (q >= qs1 ? 1 : 0) + (q >= qs2 ? 1 : 0) + (q >= qs3 ? 1 : 0);           \
    const NrLdsClu C = clu[c];                                                            \
    const int local = q - C.q0;                                                           \
    const int row = (int)(((float)local + 0.5f) * (1.0f / (float)C.sz4));                 \
    const int col = local - row * C.sz4;                                                  \
    const bool valid = q < qtot;                                                          \
    src##i = valid ? C.g4 + row * gp4 + col : g4_0;                                       \
    dst##i = valid ? C.base + row * C.p4 + (col % S) * C.half + col / S : -1;             \
  }
      NR_LDS_SLOTS(NR_SLOT_ADDR)
#undef NR_SLOT_ADDR
#define NR_SLOT_ISSUE(i) stg##i = occ4[src##i + off];
#define NR_SLOT_COMMIT(i) \
  if (dst##i >= 0) slab[dst##i] = stg##i;
      {
        const int off = (gfirst - m) * plane4;
        NR_LDS_SLOTS(NR_SLOT_ISSUE)
      }
      // ---- sums over the dx planes
      const bool act = cid >= 0;
      const float4* my_row = slab + my_base + (y - my_ylo + m) * my_p4 + (zr - my_zlo);
      float acc[RL], part[RL];
#pragma unroll
      for (int j = 0; j < RL; ++j) acc[j] = part[j] = 0.0f;
      for (int g = gfirst; g <= 2 * m;) {
        const int c0 = col_start[g], c1 = col_start[g + 1];
        int gn = g + 1;
        while (gn <= 2 * m && col_start[gn] == col_start[gn + 1]) ++gn;
        wave_lds_sync();  // the previous plane's reads are done
        NR_LDS_SLOTS(NR_SLOT_COMMIT)
        wave_lds_sync();
        if (gn <= 2 * m) {  // next plane's loads fly while this one is summed
          const int off = (gn - m) * plane4;
          NR_LDS_SLOTS(NR_SLOT_ISSUE)
        }
        if (g != gfirst) {
#pragma unroll
          for (int j = 0; j < RL; ++j) {
            acc[j] += part[j];
            part[j] = 0.0f;
          }
        }
        if (act) {
          for (int c = c0; c < c1; ++c) {
            const NrColumn col = cols[c];
            const float4* rowp = my_row + col.dy * my_p4;
            const float* wc = w + col.woff;
            switch (-col.dz0) {
              case 0: lds_column<0, MP, RL>(rowp, my_half, wc, part); break;
              case 1: lds_column<1, MP, RL>(rowp, my_half, wc, part); break;
              case 2: lds_column<2, MP, RL>(rowp, my_half, wc, part); break;
              case 3: lds_column<3, MP, RL>(rowp, my_half, wc, part); break;
              case 4: lds_column<4, MP, RL>(rowp, my_half, wc, part); break;
              case 5: lds_column<5, MP, RL>(rowp, my_half, wc, part); break;
              case 6: lds_column<6, MP, RL>(rowp, my_half, wc, part); break;
              case 7: lds_column<7, MP, RL>(rowp, my_half, wc, part); break;
              case 8: lds_column<8, MP, RL>(rowp, my_half, wc, part); break;
              case 9: lds_column<9, MP, RL>(rowp, my_half, wc, part); break;
              case 10: lds_column<10, MP, RL>(rowp, my_half, wc, part); break;
              case 11: lds_column<11, MP, RL>(rowp, my_half, wc, part); break;
              case 12: lds_column<12, MP, RL>(rowp, my_half, wc, part); break;
              case 13: lds_column<13, MP, RL>(rowp, my_half, wc, part); break;
              case 14: lds_column<14, MP, RL>(rowp, my_half, wc, part); break;
              case 15: lds_column<15, MP, RL>(rowp, my_half, wc, part); break;
              case 16: lds_column<16, MP, RL>(rowp, my_half, wc, part); break;
              default: break;
            }
          }
        }
        g = gn;
      }
      if (act) {
        const int z0 = zr * RL;
        const int64_t base = xy * OL.nz + z0;
#pragma unroll
        for (int j = 0; j < RL; ++j)
          if (z0 + j < OL.nz) out[base + j] = ((acc[j] + part[j]) * scale) * negbv + 1.0f;
      }
      pending = pending && !act;
#undef NR_SLOT_ISSUE
#undef NR_SLOT_COMMIT
    }
  }
}

}  // namespace svimx
