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
// Work decomposition: one workgroup of 4 waves per 4x8x8 output tile; wave w
// owns the x-layer x0+w (64 lanes = 8x8 (y, z) nodes).  The source x-planes
// x0-m .. x0+3+m stream through a double-buffered LDS plane of (8+2m) x
// (8+2m) values; for each plane every wave sums the columns of its offset
// dx = X - (x0+w), which is uniform across the wave, so the per-(dx, dy)
// column table is read through scalar loads and no lane diverges.  Column
// partial sums (<= 2m+1 taps) are accumulated with Neumaier compensation, so
// the float32 sum of ~4/3 pi R^3 terms keeps ~1e-8 relative accuracy.
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
  marks[((int64_t)(i / kTX) * nty + j / kTY) * ntz + k / kTZ] = 1;
}

// Row pitch of an LDS plane: >= 8+2m and = 8 (mod 32), so the 4 y-rows of a
// half-wave fall in disjoint bank groups.
__host__ __device__ inline int nr_pitch(int m) {
  int p = kTZ + 2 * m;
  return p + ((8 - p % 32) + 32) % 32;
}

__global__ void __launch_bounds__(256)
    k_nr_planes(const float* __restrict__ occ, Dims d, const int32_t* __restrict__ list,
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
      const bool xin = X >= 0 && X < d.nx;
      for (int q = threadIdx.x; q < PY * span; q += 256) {
        int a = q / span, c = q - a * span;
        int gy = y0 + a, gz = z0 + c;
        float v = 0.0f;
        if (xin && gy >= 0 && gy < d.ny && gz >= 0 && gz < d.nz) v = occ[((int64_t)X * d.ny + gy) * d.nz + gz];
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
          const float t = acc + s;
          comp += fabsf(acc) >= fabsf(s) ? (acc - t) + s : (s - t) + acc;
          acc = t;
        }
      }
    }
    const int gx = x0 + wave, gy = ty * kTY + ly, gz = tz * kTZ + lz;
    if (gx < d.nx && gy < d.ny && gz < d.nz) out[((int64_t)gx * d.ny + gy) * d.nz + gz] = ((acc + comp) * scale) * negbv + 1.0f;
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
  SVIMX_REQUIRE(sizeof(float) * 2 * (size_t)(kTY + 2 * m) * nr_pitch(m) <= 64 * 1024, SVIMX_EBADARG,
                "ball radius of %d cells exceeds the LDS plane limit (m <= 33)", m);
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
      NrColumn col{b - m, first - m, last - first + 1, (int)wts.size()};
      for (int c = first; c <= last; ++c) wts.push_back((float)weights[((int64_t)a * W + b) * W + c]);
      cols.push_back(col);
    }
  }
  col_start[W] = (int)cols.size();
  NrPlan* p = new NrPlan();
  p->m = m;
  p->h = h;
  p->radius = radius;
  p->d = Dims{nx, ny, nz};
  p->ntx = (nx + kTX - 1) / kTX;
  p->nty = (ny + kTY - 1) / kTY;
  p->ntz = (nz + kTZ - 1) / kTZ;
  p->ntiles = (int64_t)p->ntx * p->nty * p->ntz;
  p->ncols = (int)cols.size();
  p->nweights = (int)wts.size();
  p->scale = (float)(h * h * h);
  p->negbv = (float)(-1.0 / (4.0 / 3.0 * M_PI * radius * radius * radius));
  hipError_t e = hipSuccess;
  e = hipMalloc(&p->cols, sizeof(NrColumn) * (cols.size() + 1));
  if (e == hipSuccess) e = hipMalloc(&p->col_start, sizeof(int) * col_start.size());
  if (e == hipSuccess)
    e = hipMemcpy(p->col_start, col_start.data(), sizeof(int) * col_start.size(), hipMemcpyHostToDevice);
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
    hipFree(p->cols); hipFree(p->col_start); hipFree(p->weights); hipFree(p->marks); hipFree(p->list); hipFree(p->count);
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
  hipFree(p->col_start);
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

int svimx_nr_select(void* stream, void* plan) {
  SVIMX_TRY_BEGIN
  NrPlan* p = static_cast<NrPlan*>(plan);
  SVIMX_REQUIRE(p, SVIMX_EBADARG, "null plan");
  hipStream_t st = as_stream(stream);
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
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, p->d.size(), w)) return SVIMX_EBADARG;
  const size_t lds = sizeof(float) * 2 * (size_t)(kTY + 2 * p->m) * nr_pitch(p->m);
  int grid = (int)std::min<int64_t>(p->ntiles, 256 * 8);
  hipLaunchKernelGGL(k_nr_planes, dim3(grid), dim3(256), lds, as_stream(stream), occupancy, p->d, p->list, p->count,
                     p->nty, p->ntz, p->m, p->cols, p->col_start, p->weights, p->scale, p->negbv, nr_out, w.results);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_nr_exec(void* stream, void* plan, const float* occupancy, float* nr_out, void* ws, int64_t ws_bytes) {
  int rc = svimx_nr_select(stream, plan);
  if (rc) return rc;
  return svimx_nr_sum(stream, plan, occupancy, nr_out, ws, ws_bytes);
}

}  // extern "C"
