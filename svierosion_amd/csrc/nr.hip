// N_r by direct summation of the ball-weight convolution over marked tiles.
//
// N_r(x) = 1 - h^3 * sum_d w(d) f(x + d) / |B_r|, w the exact ball/cell
// overlaps (svierosion/svi.py:24-55), f the occupancy (1 inside, cell fraction
// in the band).  The reference evaluates the same sum as a circular FFT
// convolution over the whole grid (svierosion/svi.py:84-93); the margin check
// keeps the solid m+1 cells from the boundary, so the circular and the linear
// sum agree wherever f can be nonzero, and reading f = 0 outside the grid is
// exact.
//
// Work decomposition: one workgroup (one wave of 64 lanes) per 4x4x4 output
// tile; the tile plus an m-cell halo of f is staged in LDS, then each lane sums
// the nonzero taps column by column ((dx, dy) columns with contiguous dz runs,
// read through scalar loads).  Column partial sums keep the float32
// accumulation error ~1e-7 relative.
#include <cmath>
#include <vector>

#include "common.h"
#include "nr_plan.h"

namespace svimx {

__global__ void k_nr_compact_tiles(const uint8_t* __restrict__ marks, int64_t ntiles, int32_t* __restrict__ list,
                                   int32_t* __restrict__ count) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntiles && marks[t]) list[atomicAdd(count, 1)] = (int32_t)t;
}

__global__ void k_nr_zero_count(int32_t* count) { *count = 0; }

__global__ void k_nr_clear_marks(uint8_t* marks, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) marks[i] = 0;
}

__global__ void k_nr_mark_nodes(uint8_t* marks, int nty, int ntz, Dims d, const int32_t* nodes, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int i, j, k;
  unflat(d, nodes[q], i, j, k);
  marks[((int64_t)(i / kNrTile) * nty + j / kNrTile) * ntz + k / kNrTile] = 1;
}

__global__ void __launch_bounds__(64)
    k_nr_direct(const float* __restrict__ occ, Dims d, const int32_t* __restrict__ list,
                const int32_t* __restrict__ count, int nty, int ntz, int m, const NrColumn* __restrict__ cols,
                int ncols, const float* __restrict__ w, float scale, float negbv, float* __restrict__ out,
                double* __restrict__ results) {
  extern __shared__ float region[];
  const int R = kNrTile + 2 * m;
  const int RR = R * R;
  const int total = RR * R;
  const int ntile = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && results) results[SVIMX_R_TILES] = (double)ntile;
  const int lx = threadIdx.x >> 4, ly = (threadIdx.x >> 2) & 3, lz = threadIdx.x & 3;
  for (int it = blockIdx.x; it < ntile; it += gridDim.x) {
    const int t = list[it];
    const int tz = t % ntz;
    const int ty = (t / ntz) % nty;
    const int tx = t / (ntz * nty);
    const int x0 = tx * kNrTile - m, y0 = ty * kNrTile - m, z0 = tz * kNrTile - m;
    __syncthreads();
    for (int q = threadIdx.x; q < total; q += 64) {
      int a = q / RR, r = q - a * RR;
      int b = r / R, c = r - b * R;
      int gx = x0 + a, gy = y0 + b, gz = z0 + c;
      float v = 0.0f;
      if (gx >= 0 && gx < d.nx && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz)
        v = occ[((int64_t)gx * d.ny + gy) * d.nz + gz];
      region[q] = v;
    }
    __syncthreads();
    float acc = 0.0f;
    for (int ci = 0; ci < ncols; ++ci) {
      const NrColumn col = cols[ci];
      const float* src = region + (lx + m + col.dx) * RR + (ly + m + col.dy) * R + (lz + m + col.dz0);
      const float* wc = w + col.woff;
      float s = 0.0f;
      for (int q = 0; q < col.nz; ++q) s += wc[q] * src[q];
      acc += s;
    }
    const int gx = tx * kNrTile + lx, gy = ty * kNrTile + ly, gz = tz * kNrTile + lz;
    if (gx < d.nx && gy < d.ny && gz < d.nz) out[((int64_t)gx * d.ny + gy) * d.nz + gz] = (acc * scale) * negbv + 1.0f;
  }
}

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_nr_plan_create(const double* weights, int m, double h, double radius, int nx, int ny, int nz,
                         void** plan_out) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(weights && plan_out && m >= 0 && h > 0 && radius > 0, SVIMX_EBADARG, "bad N_r plan arguments");
  SVIMX_REQUIRE(2 * m + 1 <= nx && 2 * m + 1 <= ny && 2 * m + 1 <= nz, SVIMX_EBADARG,
                "ball kernel larger than the grid");
  const int R = kNrTile + 2 * m;
  SVIMX_REQUIRE((int64_t)R * R * R * 4 <= 64 * 1024, SVIMX_EBADARG,
                "ball radius of %d cells exceeds the direct-summation tile limit (m <= 10)", m);
  const int W = 2 * m + 1;
  std::vector<NrColumn> cols;
  std::vector<float> wts;
  for (int a = 0; a < W; ++a)
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
      NrColumn col{a - m, b - m, first - m, last - first + 1, (int)wts.size()};
      for (int c = first; c <= last; ++c) wts.push_back((float)weights[((int64_t)a * W + b) * W + c]);
      cols.push_back(col);
    }
  NrPlan* p = new NrPlan();
  p->m = m;
  p->h = h;
  p->radius = radius;
  p->d = Dims{nx, ny, nz};
  p->ntx = (nx + kNrTile - 1) / kNrTile;
  p->nty = (ny + kNrTile - 1) / kNrTile;
  p->ntz = (nz + kNrTile - 1) / kNrTile;
  p->ntiles = (int64_t)p->ntx * p->nty * p->ntz;
  p->ncols = (int)cols.size();
  p->nweights = (int)wts.size();
  p->scale = (float)(h * h * h);
  p->negbv = (float)(-1.0 / (4.0 / 3.0 * M_PI * radius * radius * radius));
  hipError_t e = hipSuccess;
  e = hipMalloc(&p->cols, sizeof(NrColumn) * (cols.size() + 1));
  if (e == hipSuccess) e = hipMalloc(&p->weights, sizeof(float) * (wts.size() + 1));
  if (e == hipSuccess) e = hipMalloc(&p->marks, p->ntiles);
  if (e == hipSuccess) e = hipMalloc(&p->list, sizeof(int32_t) * p->ntiles);
  if (e == hipSuccess) e = hipMalloc(&p->count, sizeof(int32_t));
  if (e == hipSuccess && !cols.empty())
    e = hipMemcpy(p->cols, cols.data(), sizeof(NrColumn) * cols.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !wts.empty())
    e = hipMemcpy(p->weights, wts.data(), sizeof(float) * wts.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(p->marks, 0, p->ntiles);
  if (e != hipSuccess) {
    set_error("N_r plan allocation failed: %s", hipGetErrorString(e));
    hipFree(p->cols); hipFree(p->weights); hipFree(p->marks); hipFree(p->list); hipFree(p->count);
    delete p;
    return SVIMX_EDEVICE;
  }
  *plan_out = p;
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_nr_plan_destroy(void* plan) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  if (!p) return SVIMX_OK;
  hipFree(p->cols);
  hipFree(p->weights);
  hipFree(p->marks);
  hipFree(p->list);
  hipFree(p->count);
  delete p;
  return SVIMX_OK;
}

int svimx_nr_reset(void* stream, void* plan) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p, SVIMX_EBADARG, "null plan");
  hipLaunchKernelGGL(k_nr_clear_marks, dim3(blocks_for(p->ntiles, 256)), dim3(256), 0, as_stream(stream), p->marks,
                     p->ntiles);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int svimx_nr_mark_nodes(void* stream, void* plan, const int32_t* nodes, int64_t n) {
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p, SVIMX_EBADARG, "null plan");
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_nr_mark_nodes, dim3(blocks_for(n, 256)), dim3(256), 0, as_stream(stream), p->marks, p->nty,
                     p->ntz, p->d, nodes, n);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int svimx_nr_exec(void* stream, void* plan, const float* occupancy, float* nr_out, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p && occupancy && nr_out, SVIMX_EBADARG, "null argument");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, p->d.size(), w)) return SVIMX_EBADARG;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_nr_zero_count, dim3(1), dim3(1), 0, st, p->count);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_nr_compact_tiles, dim3(blocks_for(p->ntiles, 256)), dim3(256), 0, st, p->marks, p->ntiles,
                     p->list, p->count);
  SVIMX_CHECK_LAUNCH();
  const int R = kNrTile + 2 * p->m;
  const size_t lds = sizeof(float) * (size_t)R * R * R;
  int grid = (int)std::min<int64_t>(p->ntiles, 256 * 16);
  hipLaunchKernelGGL(k_nr_direct, dim3(grid), dim3(64), lds, st, occupancy, p->d, p->list, p->count, p->nty,
                     p->ntz, p->m, p->cols, p->ncols, p->weights, p->scale, p->negbv, nr_out, w.results);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
