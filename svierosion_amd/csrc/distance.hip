// Unsigned distance from points to a triangle mesh on the device: the depth
// annotation of eroded surfaces (svierosion/api.py:104-168, :171-194).
//
// The reference builds a k-d tree of triangle centroids on the host and
// evaluates the k nearest candidates, enlarging k until a certificate holds.
// Here the triangles (split to edges <= 8 cells and stored in triangle-local
// float32 frames, as for the SDF) are binned into a uniform grid of buckets of
// `bs` cells by their bounding boxes (count, scan, fill), and every point
// searches rings of buckets around its own: after ring s, any triangle not
// yet tested lies only in buckets at Chebyshev distance >= s+1, so it is at
// least s*bs + delta away (delta = distance from the point to the faces of
// its own bucket).  The search stops once the best squared distance is at
// most that bound squared, so the result is the exact minimum (up to float32
// rounding in the local frames), not a k-nearest approximation.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "compact.h"

namespace svimx {

struct BucketGrid {
  int bx, by, bz;  // buckets per axis
  float bs;        // bucket edge in cells
  __device__ __forceinline__ int64_t id(int i, int j, int k) const { return ((int64_t)i * by + j) * bz + k; }
};

__device__ __forceinline__ void tri_bucket_range(const float* __restrict__ local, const int32_t* __restrict__ org,
                                                 int64_t t, const BucketGrid& B, int lo[3], int hi[3]) {
  const float* v = local + 9 * t;
  const int32_t* o = org + 3 * t;
  const int nb[3] = {B.bx, B.by, B.bz};
  for (int a = 0; a < 3; ++a) {
    const float mn = fminf(fminf(v[a], v[3 + a]), v[6 + a]) + (float)o[a];
    const float mx = fmaxf(fmaxf(v[a], v[3 + a]), v[6 + a]) + (float)o[a];
    lo[a] = min(max((int)floorf(mn / B.bs), 0), nb[a] - 1);
    hi[a] = min(max((int)floorf(mx / B.bs), 0), nb[a] - 1);
  }
}

__global__ void k_bucket_zero(int32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_bucket_count(const float* __restrict__ local, const int32_t* __restrict__ org, int64_t ntri,
                               BucketGrid B, int32_t* counts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  int lo[3], hi[3];
  tri_bucket_range(local, org, t, B, lo, hi);
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k) atomicAdd(&counts[B.id(i, j, k)], 1);
}


__global__ void k_bucket_fill(const float* __restrict__ local, const int32_t* __restrict__ org, int64_t ntri,
                              BucketGrid B, const int64_t* __restrict__ offs, int32_t* cursor, int32_t* items) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  int lo[3], hi[3];
  tri_bucket_range(local, org, t, B, lo, hi);
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k) {
        const int64_t b = B.id(i, j, k);
        items[offs[b] + atomicAdd(&cursor[b], 1)] = (int32_t)t;
      }
}

__device__ __forceinline__ float bucket_min_d2(const float* __restrict__ local, const int32_t* __restrict__ org,
                                               const int64_t* __restrict__ offs, const int32_t* __restrict__ items,
                                               int64_t b, const float p[3], float best) {
  const int64_t e = offs[b + 1];
  for (int64_t q = offs[b]; q < e; ++q) {
    const int32_t t = items[q];
    const float* v = local + 9 * (int64_t)t;
    const int32_t* o = org + 3 * (int64_t)t;
    const float pl[3] = {p[0] - (float)o[0], p[1] - (float)o[1], p[2] - (float)o[2]};
    best = fminf(best, point_triangle_d2(pl, v, v + 3, v + 6));
  }
  return best;
}

__global__ void k_point_mesh_distance(const float* __restrict__ local, const int32_t* __restrict__ org,
                                      BucketGrid B, const int64_t* __restrict__ offs,
                                      const int32_t* __restrict__ items, const float* __restrict__ pts, int64_t npts,
                                      float* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const float p[3] = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  const int nb[3] = {B.bx, B.by, B.bz};
  int c[3];
  float delta = INFINITY;
  for (int a = 0; a < 3; ++a) {
    c[a] = min(max((int)floorf(p[a] / B.bs), 0), nb[a] - 1);
    const float lo = c[a] * B.bs, hi = lo + B.bs;
    delta = fminf(delta, fmaxf(fminf(p[a] - lo, hi - p[a]), 0.0f));
  }
  // rings beyond this radius leave the bucket grid on every side
  int smax = 0;
  for (int a = 0; a < 3; ++a) smax = max(smax, max(c[a], nb[a] - 1 - c[a]));
  float best = INFINITY;
  for (int s = 0; s <= smax; ++s) {
    for (int i = c[0] - s; i <= c[0] + s; ++i) {
      if (i < 0 || i >= nb[0]) continue;
      const bool xface = i == c[0] - s || i == c[0] + s;
      for (int j = c[1] - s; j <= c[1] + s; ++j) {
        if (j < 0 || j >= nb[1]) continue;
        const bool face = xface || j == c[1] - s || j == c[1] + s;
        if (face) {
          for (int k = max(c[2] - s, 0); k <= min(c[2] + s, nb[2] - 1); ++k)
            best = bucket_min_d2(local, org, offs, items, B.id(i, j, k), p, best);
        } else {
          if (c[2] - s >= 0) best = bucket_min_d2(local, org, offs, items, B.id(i, j, c[2] - s), p, best);
          if (s > 0 && c[2] + s < nb[2]) best = bucket_min_d2(local, org, offs, items, B.id(i, j, c[2] + s), p, best);
        }
      }
    }
    const float bound = s * B.bs + delta;
    if (best <= bound * bound) break;
  }
  out[q] = sqrtf(best);
}

template <class T>
struct DBuf {
  T* p = nullptr;
  hipStream_t st = nullptr;
  ~DBuf() {
    if (p) hipFreeAsync(p, st);
  }
  hipError_t alloc(int64_t n, hipStream_t s) {
    st = s;
    return hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(T) * (n > 0 ? n : 1), s);
  }
};

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_point_mesh_distance(void* stream, const float* tri_local, const int32_t* tri_org, int64_t ntri, int nx,
                              int ny, int nz, double bucket_cells, const float* points, int64_t npts,
                              float* dist_out) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(tri_local && tri_org && points && dist_out && ntri > 0 && npts >= 0 && bucket_cells >= 1.0,
                SVIMX_EBADARG, "bad point-mesh distance arguments");
  if (npts == 0) return SVIMX_OK;
  hipStream_t st = as_stream(stream);
  BucketGrid B;
  B.bs = (float)bucket_cells;
  B.bx = std::max(1, (int)std::ceil(nx / bucket_cells));
  B.by = std::max(1, (int)std::ceil(ny / bucket_cells));
  B.bz = std::max(1, (int)std::ceil(nz / bucket_cells));
  const int64_t nbk = (int64_t)B.bx * B.by * B.bz;
  DBuf<int32_t> counts, cursor, items;
  DBuf<int64_t> offs;
  SVIMX_CHECK_HIP(counts.alloc(nbk, st));
  SVIMX_CHECK_HIP(cursor.alloc(nbk, st));
  SVIMX_CHECK_HIP(offs.alloc(nbk + 1, st));
  const dim3 gb(blocks_for(nbk, 256)), gt(blocks_for(ntri, 256)), tb(256);
  hipLaunchKernelGGL(k_bucket_zero, gb, tb, 0, st, counts.p, nbk);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bucket_zero, gb, tb, 0, st, cursor.p, nbk);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bucket_count, gt, tb, 0, st, tri_local, tri_org, ntri, B, counts.p);
  SVIMX_CHECK_LAUNCH();
  {
    const int rc = scan_to_offsets(st, counts.p, nbk, offs.p);
    if (rc) return rc;
  }
  int64_t nitems = 0;
  SVIMX_CHECK_HIP(hipMemcpyAsync(&nitems, offs.p + nbk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  SVIMX_CHECK_HIP(items.alloc(nitems, st));
  hipLaunchKernelGGL(k_bucket_fill, gt, tb, 0, st, tri_local, tri_org, ntri, B, offs.p, cursor.p, items.p);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_point_mesh_distance, dim3(blocks_for(npts, 256)), tb, 0, st, tri_local, tri_org, B, offs.p,
                     items.p, points, npts, dist_out);
  SVIMX_CHECK_LAUNCH();
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
