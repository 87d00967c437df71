// N_r by direct summation of the ball-weight convolution.
//
// N_r(x) = 1 - h^3 * sum_d w(d) f(x + d) / |B_r|, w the exact ball/cell
// overlaps (svierosion/svi.py:24-55), f the occupancy (1 inside, cell fraction
// in the band).  The reference evaluates the same sum as a circular FFT
// convolution over the whole grid (svierosion/svi.py:84-93); the margin check
// keeps the solid m+1 cells from the boundary, so the circular and the linear
// sum agree wherever f can be nonzero, and the zero padding of the occupancy
// layout is exact.  N_r is only needed at the tricubic stencils of the closest
// points, so only those nodes are evaluated.
//
// Run mode (ball half-width m <= 16).  Work unit: a z-run of 8 nodes.  Each
// lane owns one run: for every (dx, dy) column of the ball (uniform across the
// wave, read through scalar loads) it loads the 8+2M' occupancy values of the
// column window as aligned float4s (M' = m rounded up to 4) and accumulates
// the 2k+1 nonzero taps (the ball is symmetric, so a column's taps are
// dz = -k..k) into 8 accumulators with a fully unrolled sequence per k.  Column
// sums are grouped per dx before entering the total, keeping float32 rounding
// ~1e-7 relative.
//
// Tile mode (larger balls).  One workgroup of 4 waves per 4x8x8 tile, wave w
// owning x-layer x0+w; source x-planes stream through double-buffered LDS and
// each wave sums the columns of its (wave-uniform) offset dx = X - (x0+w).
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "compact.h"
#include "nr_dev.h"
#include "nr_plan.h"

namespace svimx {

// ---------------------------------------------------------------- selection
__global__ void k_nr_zero_count(int32_t* count) { *count = 0; }

__global__ void k_nr_clear_marks(uint8_t* marks, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) marks[i] = 0;
}

__global__ void k_nr_compact_tiles(const uint8_t* __restrict__ marks, int64_t ntiles, int32_t* __restrict__ list,
                                   int32_t* __restrict__ count) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles && marks[t]) list[atomicAdd(count, 1)] = (int32_t)t;
}

__global__ void k_nr_mark_nodes(uint8_t* marks, int run_mode, int nty, int ntz, int nzr, Dims d,
                                const int32_t* nodes, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int i, j, k;
  unflat(d, nodes[q], i, j, k);
  if (run_mode)
    marks[((int64_t)i * d.ny + j) * nzr + k / run_mode] = 1;
  else
    marks[((int64_t)(i / kTX) * nty + j / kTY) * ntz + k / kTZ] = 1;
}

// Marks of the tricubic stencils of index-space points: base b = floor(p) - 1
// clamped to [0, n - 4] per axis, so the stencil [b, b + 3] covers every node
// an axis-clamped interpolation reads (selection dilates the run marks in x
// and y by the stencil, as for the closest points of the fractions pass).
__global__ void k_nr_mark_points(uint8_t* marks, int run_mode, int nty, int ntz, int nzr, Dims d,
                                 const float* __restrict__ p, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int dims[3] = {d.nx, d.ny, d.nz};
  int b[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float v = p[3 * q + a];
    const int f = v == v ? (int)fminf(fmaxf(floorf(v), -2.0f), (float)dims[a]) : 0;  // NaN -> 0
    b[a] = min(max(f - 1, 0), max(dims[a] - 4, 0));
  }
  if (run_mode)
    mark_stencil_runs(marks, d.ny, nzr, run_mode, b[0], b[1], b[2]);
  else
    mark_stencil_tiles(marks, nty, ntz, b[0], b[1], b[2]);
}

__global__ void k_nr_unit_need(const uint8_t* __restrict__ marks, uint8_t* __restrict__ need, int ny, int nzr,
                               int nyu, int Y, int64_t nunits) {
  unit_need_range(marks, need, ny, nzr, nyu, Y, nunits, (int64_t)blockIdx.x * blockDim.x + threadIdx.x,
                  (int64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------- run-mode summation
template <int MP, int RL>
__global__ void __launch_bounds__(256)
    k_nr_runs(const float* __restrict__ occ, OccLayout OL, int nzr, const int32_t* __restrict__ list,
              const int64_t* __restrict__ total, const NrColumn* __restrict__ cols, int ncols, int cfirst,
              int csplit, const float* __restrict__ w, float scale, float negbv, float* __restrict__ out,
              double* __restrict__ results) {
  const int64_t nrun = *total;
  if (blockIdx.x == 0 && threadIdx.x == 0 && results) results[SVIMX_R_TILES] = (double)nrun;
  // neighbouring runs read the same occupancy columns: keep them on one XCD's L2
  nr_runs_body<MP, RL>(occ, OL, nzr, list, nrun, cols, ncols, cfirst, csplit, w, scale, negbv, out, blockIdx.x,
                       gridDim.x);
}

// Units of Y adjacent y-lines sharing their occupancy windows.
template <int MP, int RL, int Y>
__global__ void __launch_bounds__(256)
    k_nr_lines(const float* __restrict__ occ, OccLayout OL, int nzr, const int32_t* __restrict__ list,
               const int64_t* __restrict__ total, const NrColumn* __restrict__ cols, const int* __restrict__ col_start,
               int m, int asplit, const float* __restrict__ colrec, float scale, float negbv, float* __restrict__ out,
               double* __restrict__ results) {
  const int64_t nunit = *total;
  if (blockIdx.x == 0 && threadIdx.x == 0 && results) results[SVIMX_R_TILES] = (double)nunit;
  __shared__ float xch[nr_lines_xch<RL, Y, kNrLineLanes>()];
  nr_lines_body<MP, RL, Y, kNrLineLanes>(occ, OL, nzr, list, nunit, cols, col_start, m, asplit, colrec, scale, negbv,
                                         out, xch, blockIdx.x, gridDim.x);
}

// The same sums with the occupancy staged through LDS: one wave per block.
template <int MP, int RL>
__global__ void __launch_bounds__(64)
    k_nr_lds(const float* __restrict__ occ, OccLayout OL, int nzr, const int32_t* __restrict__ list,
             const int64_t* __restrict__ total, const NrColumn* __restrict__ cols, const int* __restrict__ col_start,
             int m, const float* __restrict__ w, float scale, float negbv, float* __restrict__ out,
             double* __restrict__ results) {
  __shared__ float4 slab[kNrLdsCap4];
  __shared__ NrLdsClu clu[kNrLdsClusters];
  const int64_t nrun = *total;
  if (blockIdx.x == 0 && threadIdx.x == 0 && results) results[SVIMX_R_TILES] = (double)nrun;
  nr_lds_body<MP, RL>(occ, OL, nzr, list, nrun, cols, col_start, m, w, scale, negbv, out, slab, clu,
                      blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- tile-mode summation
// Row pitch of an LDS plane: >= 8+2m and = 8 (mod 32), so the 4 y-rows of a
// half-wave fall in disjoint bank groups.
__host__ __device__ inline int nr_pitch(int m) {
  int p = kTZ + 2 * m;
  return p + ((8 - p % 32) + 32) % 32;
}

__global__ void __launch_bounds__(256)
    k_nr_planes(const float* __restrict__ occ, OccLayout OL, const int32_t* __restrict__ list,
                const int32_t* __restrict__ count, int nty, int ntz, int m, const NrColumn* __restrict__ cols,
                const int* __restrict__ col_start, const float* __restrict__ w, float scale, float negbv,
                float* __restrict__ out, double* __restrict__ results) {
  extern __shared__ float planes[];
  const int PY = kTY + 2 * m, PZ = nr_pitch(m), P = PY * PZ;
  const int ntile = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && results) results[SVIMX_R_TILES] = (double)ntile;
  // the wave index is uniform per wave; readfirstlane lets the compiler keep dx,
  // the column table and the weights in scalar registers (s_load)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ly = lane >> 3, lz = lane & 7;
  const int span = kTZ + 2 * m;
  for (int it = blockIdx.x; it < ntile; it += gridDim.x) {
    const int t = list[it];
    const int tz = t % ntz;
    const int ty = (t / ntz) % nty;
    const int tx = t / (ntz * nty);
    const int x0 = tx * kTX, y0 = ty * kTY - m, z0 = tz * kTZ - m;
    float acc = 0.0f, comp = 0.0f;
    __syncthreads();
    for (int X = x0 - m; X <= x0 + kTX - 1 + m; ++X) {
      float* buf = planes + (X & 1) * P;
      const bool xin = X >= 0 && X < OL.nx;
      for (int q = threadIdx.x; q < PY * span; q += 256) {
        int a = q / span, c = q - a * span;
        int gy = y0 + a, gz = z0 + c;
        float v = 0.0f;
        if (xin && gy >= 0 && gy < OL.ny && gz >= 0 && gz < OL.nz) v = occ[OL.at(X, gy, gz)];
        buf[a * PZ + c] = v;
      }
      __syncthreads();
      const int dx = X - (x0 + wave);
      if (dx >= -m && dx <= m) {
        const int c0 = col_start[dx + m], c1 = col_start[dx + m + 1];
        for (int ci = c0; ci < c1; ++ci) {
          const NrColumn col = cols[ci];
          const float* src = buf + (ly + m + col.dy) * PZ + (lz + m + col.dz0);
          const float* wc = w + col.woff;
          float s = 0.0f;
#pragma unroll 4
          for (int q = 0; q < col.nz; ++q) s += wc[q] * src[q];
          // Neumaier compensated accumulation of the column sums
          const float tsum = acc + s;
          comp += fabsf(acc) >= fabsf(s) ? (acc - tsum) + s : (s - tsum) + acc;
          acc = tsum;
        }
      }
    }
    const int gx = x0 + wave, gy = ty * kTY + ly, gz = tz * kTZ + lz;
    if (gx < OL.nx && gy < OL.ny && gz < OL.nz)
      out[((int64_t)gx * OL.ny + gy) * OL.nz + gz] = ((acc + comp) * scale) * negbv + 1.0f;
  }
}

// Run length and lines per unit of a run-mode plan (the arrays are sized for
// kRunLenMin and one line per unit).
void nr_plan_set_shape(NrPlan* p, int rl, int ylines) {
  p->run_mode = rl;
  p->ylines = ylines;
  p->nzr = (p->d.nz + rl - 1) / rl;
  p->nruns = (int64_t)p->d.nx * p->d.ny * p->nzr;
  p->nyu = (p->d.ny + ylines - 1) / ylines;
  p->nunits = (int64_t)p->d.nx * p->nyu * p->nzr;
  p->nmarks = p->nruns;
  p->nchunks = (p->nunits + kCompactChunk - 1) / kCompactChunk;
}

void nr_plan_set_run_len(NrPlan* p, int rl) { nr_plan_set_shape(p, rl, p->ylines > 0 ? p->ylines : 1); }

// Shapes a run-mode plan can take: LDS staging needs one line and an
// occupancy indexable in int32; several lines need consecutive-dy columns,
// windows within the occupancy's y padding (m + lines - 1 <= kOccPad) and
// runs of at most 8 nodes for four lines.
bool nr_shape_allowed(const NrPlan* p, const NrShape& s) {
  if (!p->run_mode || p->fft) return false;
  if (s.rl != 4 && s.rl != 8 && s.rl != 16) return false;
  if (s.lines != 1 && s.lines != 2 && s.lines != 4) return false;
  if (s.lds) return s.lines == 1 && p->lds_ok;
  // the multi-line kernels index the occupancy by 32-bit float4 offsets (lds_ok)
  if (s.lines > 1)
    return p->cols_contig && p->lds_ok && p->m + s.lines - 1 <= kOccPad && !(s.lines == 4 && s.rl > 8);
  return true;
}

}  // namespace svimx

using namespace svimx;

extern "C" {

int64_t svimx_occupancy_elems(int nx, int ny, int nz) { return OccLayout::of(nx, ny, nz).size(); }

int svimx_nr_plan_create(const double* weights, int m, double h, double radius, int nx, int ny, int nz,
                         void** plan_out) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(weights && plan_out && m >= 0 && h > 0 && radius > 0, SVIMX_EBADARG, "bad N_r plan arguments");
  SVIMX_REQUIRE(2 * m + 1 <= nx && 2 * m + 1 <= ny && 2 * m + 1 <= nz, SVIMX_EBADARG,
                "ball kernel larger than the grid");
  {
    // every evaluation path relies on the ball's mirror symmetry in x, y and z
    // (paired +-dx windows, symmetric z taps, a real kernel spectrum)
    const int W = 2 * m + 1;
    auto wf = [&](int a, int b, int c) { return (float)weights[((int64_t)a * W + b) * W + c]; };
    bool sym = true;
    for (int a = 0; a < W && sym; ++a)
      for (int b = 0; b < W && sym; ++b)
        for (int c = 0; c < W && sym; ++c)
          sym = wf(a, b, c) == wf(W - 1 - a, b, c) && wf(a, b, c) == wf(a, W - 1 - b, c) &&
                wf(a, b, c) == wf(a, b, W - 1 - c);
    SVIMX_REQUIRE(sym, SVIMX_EBADARG, "N_r weights must be mirror-symmetric in x, y and z (as FP32 values)");
  }
  const bool run_mode = m <= kRunMaxM;
  // balls wider than the LDS plane of the tile kernel (m > 33) go to the FFT
  const bool use_fft = !run_mode && sizeof(float) * 2 * (size_t)(kTY + 2 * m) * nr_pitch(m) > 64 * 1024;
  const int W = 2 * m + 1;
  std::vector<NrColumn> cols;
  std::vector<int> col_start(W + 1, 0);
  std::vector<float> wts;
  for (int a = 0; a < W; ++a) {
    col_start[a] = (int)cols.size();
    for (int b = 0; b < W; ++b) {
      int first = -1, last = -1;
      for (int c = 0; c < W; ++c) {
        float v = (float)weights[((int64_t)a * W + b) * W + c];
        if (v != 0.0f) {
          if (first < 0) first = c;
          last = c;
        }
      }
      if (first < 0) continue;
      if (run_mode) {  // symmetric tap range dz = -k..k covering every nonzero tap
        int k = std::max(m - first, last - m);
        first = m - k;
        last = m + k;
      }
      NrColumn col{a - m, b - m, first - m, last - first + 1, (int)wts.size(), 0};
      for (int c = first; c <= last; ++c) wts.push_back((float)weights[((int64_t)a * W + b) * W + c]);
      cols.push_back(col);
    }
  }
  col_start[W] = (int)cols.size();
  NrPlan* p = new NrPlan();
  p->m = m;
  // the lo / hi split of the |dx| groups, as even as possible in a cost of
  // taps plus window loads (the two waves that share a unit group in the
  // multi-line kernels take one half each); it depends on the ball only, so
  // every shape sums in the same order
  {
    std::vector<int64_t> cost(m + 1, 0);
    int64_t all = 0;
    for (int a = 0; a <= m; ++a) {
      for (int c = col_start[m + a]; c < col_start[m + a + 1]; ++c) cost[a] += 2 * (int64_t)(-cols[c].dz0) + 1;
      cost[a] += 8 * (int64_t)(col_start[m + a + 1] - col_start[m + a]) * (a > 0 ? 2 : 1);
      all += cost[a];
    }
    int64_t lo = 0, best = INT64_MAX;
    p->asplit = m + 1;
    for (int a = 0; a <= m + 1; ++a) {
      const int64_t diff = std::llabs(all - 2 * lo);
      if (diff < best) best = diff, p->asplit = a;
      if (a <= m) lo += cost[a];
    }
    p->cfirst = col_start[m];
    p->csplit = col_start[m + p->asplit];
  }
  p->weights_host.assign(weights, weights + (size_t)W * W * W);
  p->h = h;
  p->radius = radius;
  p->d = Dims{nx, ny, nz};
  p->run_mode = 0;
  p->lds = 0;
  p->lds_ok = OccLayout::of(nx, ny, nz).size() / 4 < (int64_t)INT32_MAX ? 1 : 0;
  p->ylines = 1;
  p->cols_contig = 1;  // every dx group's columns are a consecutive dy range
  for (int g = 0; g < W; ++g)
    for (int c = col_start[g] + 1; c < col_start[g + 1]; ++c)
      if (cols[c].dy != cols[c - 1].dy + 1) p->cols_contig = 0;
  p->ntx = (nx + kTX - 1) / kTX;
  p->nty = (ny + kTY - 1) / kTY;
  p->ntz = (nz + kTZ - 1) / kTZ;
  p->ntiles = (int64_t)p->ntx * p->nty * p->ntz;
  // run-mode arrays are sized for the shortest run length the tuner may pick
  const int64_t max_runs = (int64_t)nx * ny * ((nz + kRunLenMin - 1) / kRunLenMin);
  const int64_t cap_marks = run_mode ? max_runs : p->ntiles;
  const int64_t cap_chunks = (max_runs + kCompactChunk - 1) / kCompactChunk;
  p->nmarks = p->ntiles;
  if (run_mode) nr_plan_set_run_len(p, kRunLen);
  p->ncols = (int)cols.size();
  p->nweights = (int)wts.size();
  p->scale = (float)(h * h * h);
  p->negbv = (float)(-1.0 / (4.0 / 3.0 * M_PI * radius * radius * radius));
  std::vector<float> rec((cols.size() + 1) * kNrRecStride, 0.0f);
  if (run_mode)
    for (size_t c = 0; c < cols.size(); ++c) {
      const int k = -cols[c].dz0;
      const int kbits = k;
      std::memcpy(&rec[c * kNrRecStride], &kbits, sizeof(int));
      for (int d = 0; d <= k; ++d) rec[c * kNrRecStride + 1 + d] = wts[cols[c].woff + k + d];
    }
  hipError_t e = hipMalloc(&p->cols, sizeof(NrColumn) * (cols.size() + 1));
  if (e == hipSuccess) e = hipMalloc(&p->colrec, sizeof(float) * rec.size());
  if (e == hipSuccess) e = hipMalloc(&p->col_start, sizeof(int) * col_start.size());
  if (e == hipSuccess) e = hipMalloc(&p->weights, sizeof(float) * (wts.size() + 1));
  if (e == hipSuccess) e = hipMalloc(&p->marks, cap_marks);
  if (e == hipSuccess) e = hipMalloc(&p->dil, run_mode ? max_runs : 1);
  if (e == hipSuccess) e = hipMalloc(&p->list, sizeof(int32_t) * cap_marks);
  if (e == hipSuccess) e = hipMalloc(&p->count, sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&p->ccounts, sizeof(int64_t) * (cap_chunks + 2));
  if (e == hipSuccess && !cols.empty())
    e = hipMemcpy(p->cols, cols.data(), sizeof(NrColumn) * cols.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->col_start, col_start.data(), sizeof(int) * col_start.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !wts.empty())
    e = hipMemcpy(p->weights, wts.data(), sizeof(float) * wts.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->colrec, rec.data(), sizeof(float) * rec.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->marks, 0, cap_marks);
  if (e == hipSuccess) e = hipMemset(p->list, 0, sizeof(int32_t) * cap_marks);
  if (e == hipSuccess) e = hipMemset(p->ccounts, 0, sizeof(int64_t) * (cap_chunks + 2));
  if (e != hipSuccess) {
    set_error("N_r plan allocation failed: %s", hipGetErrorString(e));
    hipFree(p->cols); hipFree(p->col_start); hipFree(p->weights); hipFree(p->colrec); hipFree(p->marks); hipFree(p->dil);
    hipFree(p->list); hipFree(p->count); hipFree(p->ccounts);
    delete p;
    return SVIMX_EDEVICE;
  }
  if (use_fft) {
    p->fft = 1;
    p->nmarks = 0;
    int rc = fft_plan_create(p->fftp, weights, m, p->d);
    if (rc) {
      svimx_nr_plan_destroy(p);
      return rc;
    }
  }
  *plan_out = p;
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_nr_plan_set_method(void* plan, int method) {
  SVIMX_TRY_BEGIN
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && (method == 1 || method == 2), SVIMX_EBADARG, "method must be 1 (direct) or 2 (FFT)");
  if (method == 2 && !p->fft) {
    SVIMX_REQUIRE(p->weights_host.size() == (size_t)(2 * p->m + 1) * (2 * p->m + 1) * (2 * p->m + 1), SVIMX_EBADARG,
                  "plan has no weights");
    int rc = fft_plan_create(p->fftp, p->weights_host.data(), p->m, p->d);
    if (rc) return rc;
    p->fft = 1;
    p->nmarks = 0;
  } else if (method == 1 && p->fft) {
    SVIMX_REQUIRE(p->m <= 33, SVIMX_EBADARG, "ball radius of %d cells needs the FFT", p->m);
    fft_plan_destroy(p->fftp);
    p->fft = 0;
    if (p->run_mode)
      nr_plan_set_run_len(p, p->run_mode);
    else
      p->nmarks = p->ntiles;
  }
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_nr_plan_method(void* plan) {
  const NrPlan* p = static_cast<const NrPlan*>(plan);
  return p ? (p->fft ? 2 : 1) : 0;
}

int svimx_nr_plan_destroy(void* plan) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  if (!p) return SVIMX_OK;
  hipFree(p->cols);
  hipFree(p->col_start);
  hipFree(p->weights);
  hipFree(p->colrec);
  hipFree(p->marks);
  hipFree(p->dil);
  hipFree(p->list);
  hipFree(p->count);
  hipFree(p->ccounts);
  if (p->fft) fft_plan_destroy(p->fftp);
  delete p;
  return SVIMX_OK;
}

int svimx_nr_reset(void* stream, void* plan) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p, SVIMX_EBADARG, "null plan");
  if (p->fft) return SVIMX_OK;
  hipLaunchKernelGGL(k_nr_clear_marks, dim3(blocks_for(p->nmarks, 256)), dim3(256), 0, as_stream(stream), p->marks,
                     p->nmarks);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int svimx_nr_mark_nodes(void* stream, void* plan, const int32_t* nodes, int64_t n) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p, SVIMX_EBADARG, "null plan");
  if (n <= 0 || p->fft) return SVIMX_OK;
  hipLaunchKernelGGL(k_nr_mark_nodes, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), p->marks,
                     p->run_mode, p->nty, p->ntz, p->nzr, p->d, nodes, n);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int svimx_nr_mark_points(void* stream, void* plan, const float* points, int64_t n) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && (points || n <= 0), SVIMX_EBADARG, "null argument");
  if (n <= 0 || p->fft) return SVIMX_OK;
  hipLaunchKernelGGL(k_nr_mark_points, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), p->marks,
                     p->run_mode, p->nty, p->ntz, p->nzr, p->d, points, n);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int svimx_nr_select(void* stream, void* plan) {
  SVIMX_TRY_BEGIN
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p, SVIMX_EBADARG, "null plan");
  if (p->fft) return SVIMX_OK;  // the FFT evaluates every node
  hipStream_t st = as_stream(stream);
  if (p->run_mode) {
    SVIMX_REQUIRE(((reinterpret_cast<uintptr_t>(p->marks) | reinterpret_cast<uintptr_t>(p->dil)) & 3) == 0,
                  SVIMX_EBADARG, "misaligned N_r selection buffers");
    const int64_t nthr = (p->nzr & 3) == 0 ? p->nunits / 4 : p->nunits;
    hipLaunchKernelGGL(k_nr_unit_need, dim3(std::max(blocks_for(nthr, 256), 1)), dim3(256), 0, st, p->marks, p->dil,
                       p->d.ny, p->nzr, p->nyu, p->ylines, p->nunits);
    SVIMX_CHECK_LAUNCH();
    FlagPred pred{p->dil};
    hipLaunchKernelGGL(k_compact_count<FlagPred>, dim3((unsigned)p->nchunks), dim3(kCompactThreads), 0, st, pred,
                       p->nunits, p->ccounts);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, p->ccounts, p->nchunks);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_compact_scatter<FlagPred, IdentityMap>), dim3((unsigned)p->nchunks),
                       dim3(kCompactThreads), 0, st, pred, p->nunits, p->ccounts, p->list, p->nunits, IdentityMap());
    SVIMX_CHECK_LAUNCH();
    return SVIMX_OK;
  }
  hipLaunchKernelGGL(k_nr_zero_count, dim3(1), dim3(1), 0, st, p->count);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_nr_compact_tiles, dim3(blocks_for(p->ntiles, 256)), dim3(256), 0, st, p->marks, p->ntiles,
                     p->list, p->count);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_nr_sum(void* stream, void* plan, const float* occupancy, float* nr_out, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && occupancy && nr_out, SVIMX_EBADARG, "null argument");
  SVIMX_REQUIRE((reinterpret_cast<uintptr_t>(occupancy) & 15) == 0, SVIMX_EBADARG,
                "occupancy must be 16-byte aligned");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, p->d.size(), w)) return SVIMX_EBADARG;
  hipStream_t st = as_stream(stream);
  const OccLayout OL = OccLayout::of(p->d.nx, p->d.ny, p->d.nz);
  if (p->fft) return fft_nr(st, p->fftp, occupancy, nr_out, p->scale, p->negbv);
  if (p->run_mode) {
    const int64_t* total = p->ccounts + p->nchunks;
    // grid: a multiple of the XCD count (the kernel's block remapping needs it)
    const int grid = (int)std::min<int64_t>((p->nunits + 255) / 256 + kXcds - 1, 256 * 8) / kXcds * kXcds;
    const int mp = (p->m + 3) / 4 * 4;
    // LDS variant: one wave per block, at most 10 resident per CU (LDS-bound)
    const int grid_lds = (int)std::min<int64_t>((p->nruns + 63) / 64 + kXcds - 1, 256 * 10) / kXcds * kXcds;
#define SVIMX_LINES(MPV, RLV, YV)                                                                              \
  hipLaunchKernelGGL((k_nr_lines<MPV, RLV, YV>), dim3(grid), dim3(256), 0, st, occupancy, OL, p->nzr, p->list, \
                     total, p->cols, p->col_start, p->m, p->asplit, p->colrec, p->scale, p->negbv, nr_out,    \
                     w.results)
#define SVIMX_RUNS(MPV, RLV)                                                                                       \
  if (p->lds)                                                                                                      \
    hipLaunchKernelGGL((k_nr_lds<MPV, RLV>), dim3(grid_lds), dim3(64), 0, st, occupancy, OL, p->nzr, p->list,      \
                       total, p->cols, p->col_start, p->m, p->weights, p->scale, p->negbv, nr_out, w.results);    \
  else if (p->ylines == 2)                                                                                         \
    SVIMX_LINES(MPV, RLV, 2);                                                                                      \
  else if (p->ylines == 4 && RLV < 16)                                                                             \
    SVIMX_LINES(MPV, (RLV < 16 ? RLV : 8), 4);                                                                     \
  else                                                                                                             \
    hipLaunchKernelGGL((k_nr_runs<MPV, RLV>), dim3(grid), dim3(256), 0, st, occupancy, OL, p->nzr, p->list, total, \
                       p->cols, p->ncols, p->cfirst, p->csplit, p->weights, p->scale, p->negbv, nr_out, w.results)
#define SVIMX_RUNS_MP(RLV)                  \
  switch (mp) {                             \
    case 0:                                 \
    case 4: SVIMX_RUNS(4, RLV); break;      \
    case 8: SVIMX_RUNS(8, RLV); break;      \
    case 12: SVIMX_RUNS(12, RLV); break;    \
    default: SVIMX_RUNS(16, RLV); break;    \
  }
    switch (p->run_mode) {
      case 4: SVIMX_RUNS_MP(4); break;
      case 16: SVIMX_RUNS_MP(16); break;
      default: SVIMX_RUNS_MP(8); break;
    }
#undef SVIMX_RUNS_MP
#undef SVIMX_RUNS
#undef SVIMX_LINES
    SVIMX_CHECK_LAUNCH();
    return SVIMX_OK;
  }
  const size_t lds = sizeof(float) * 2 * (size_t)(kTY + 2 * p->m) * nr_pitch(p->m);
  int grid = (int)std::min<int64_t>(p->ntiles, 256 * 8);
  hipLaunchKernelGGL(k_nr_planes, dim3(grid), dim3(256), lds, st, occupancy, OL, p->list, p->count, p->nty, p->ntz,
                     p->m, p->cols, p->col_start, p->weights, p->scale, p->negbv, nr_out, w.results);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_nr_plan_tune(void* stream, void* plan, const float* phi, const int32_t* band, int64_t nb, int nx, int ny,
                       int nz, double h, double cut_hi, float* occupancy, float* nr_out, void* ws, int64_t ws_bytes,
                       double* ms_out) {
  SVIMX_TRY_BEGIN
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && phi && band && occupancy && nr_out, SVIMX_EBADARG, "null argument");
  if (!p->run_mode || p->fft) return SVIMX_OK;  // tile mode and the FFT have a single shape
  hipStream_t st = as_stream(stream);
  hipEvent_t ev[2];
  SVIMX_CHECK_HIP(hipEventCreate(&ev[0]));
  SVIMX_CHECK_HIP(hipEventCreate(&ev[1]));
  NrShape best_shape{p->run_mode, p->ylines, p->lds};
  int rc = SVIMX_OK;
  double best = INFINITY;
  for (int i = 0; i < kNrTuneCands && !rc; ++i) {
    if (ms_out) ms_out[i] = NAN;
    const NrShape sh = kNrTuneShapes[i];
    // the LDS staging keeps its own summation order: the tuner picks among
    // the shapes whose sums agree bit for bit, so a job's results never
    // depend on the timings
    if (sh.lds || !nr_shape_allowed(p, sh)) continue;
    nr_plan_set_shape(p, sh.rl, sh.lines);
    p->lds = sh.lds;
    rc = svimx_fractions(stream, phi, band, nb, nx, ny, nz, h, occupancy, p, cut_hi, ws, ws_bytes);
    if (!rc) rc = svimx_nr_select(stream, p);
    if (!rc) rc = svimx_nr_sum(stream, p, occupancy, nr_out, ws, ws_bytes);  // warm-up
    if (rc) break;
    hipEventRecord(ev[0], st);
    for (int r = 0; r < 3 && !rc; ++r) rc = svimx_nr_sum(stream, p, occupancy, nr_out, ws, ws_bytes);
    hipEventRecord(ev[1], st);
    if (hipEventSynchronize(ev[1]) != hipSuccess) rc = SVIMX_EDEVICE;
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (ms_out) ms_out[i] = ms / 3.0;
    if (ms / 3.0 < best) {
      best = ms / 3.0;
      best_shape = sh;
    }
  }
  hipEventDestroy(ev[0]);
  hipEventDestroy(ev[1]);
  // leave the plan's marks and run list those of the chosen shape
  nr_plan_set_shape(p, best_shape.rl, best_shape.lines);
  p->lds = best_shape.lds;
  if (!rc) rc = svimx_fractions(stream, phi, band, nb, nx, ny, nz, h, occupancy, p, cut_hi, ws, ws_bytes);
  if (!rc) rc = svimx_nr_select(stream, p);
  return rc;
  SVIMX_TRY_END
}

int svimx_nr_plan_run_length(void* plan) {
  const NrPlan* p = static_cast<const NrPlan*>(plan);
  return p ? p->run_mode : 0;
}

int svimx_nr_plan_set_run_length(void* plan, int rl) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && nr_shape_allowed(p, NrShape{rl, p->ylines, p->lds}), SVIMX_EBADARG,
                "run length %d is not available for this plan", rl);
  nr_plan_set_run_len(p, rl);
  return SVIMX_OK;
}

int svimx_nr_plan_set_lines(void* plan, int lines) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && nr_shape_allowed(p, NrShape{p->run_mode, lines, p->lds}), SVIMX_EBADARG,
                "%d lines per unit are not available for this plan", lines);
  nr_plan_set_shape(p, p->run_mode, lines);
  return SVIMX_OK;
}

int svimx_nr_plan_lines(void* plan) {
  const NrPlan* p = static_cast<const NrPlan*>(plan);
  return p && p->run_mode ? p->ylines : 0;
}

int svimx_nr_tune_shape(int i, int* run_length, int* lines, int* lds) {
  SVIMX_REQUIRE(i >= 0 && i < kNrTuneCands && run_length && lines && lds, SVIMX_EBADARG, "bad tuning index");
  *run_length = kNrTuneShapes[i].rl;
  *lines = kNrTuneShapes[i].lines;
  *lds = kNrTuneShapes[i].lds;
  return SVIMX_OK;
}

int svimx_nr_plan_set_staging(void* plan, int lds) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && (lds == 0 || lds == 1), SVIMX_EBADARG, "bad N_r staging argument");
  SVIMX_REQUIRE(!p->run_mode || nr_shape_allowed(p, NrShape{p->run_mode, p->ylines, lds}), SVIMX_EBADARG,
                "LDS staging needs one line per unit and a grid whose occupancy fits int32 indexing");
  p->lds = lds;
  return SVIMX_OK;
}

int svimx_nr_plan_staging(void* plan) {
  const NrPlan* p = static_cast<const NrPlan*>(plan);
  return p ? p->lds : 0;
}

int svimx_nr_exec(void* stream, void* plan, const float* occupancy, float* nr_out, void* ws, int64_t ws_bytes) {
  int rc = svimx_nr_select(stream, plan);
  if (rc) return rc;
  return svimx_nr_sum(stream, plan, occupancy, nr_out, ws, ws_bytes);
}

}  // extern "C"
