// Order-preserving stream compaction of a dense predicate into a sorted int32
// index list: per-chunk counts -> one-block exclusive scan -> per-chunk scatter
// with wave ballots.  Chunks are kCompactChunk consecutive nodes; thread t of a
// chunk visits nodes base + r*kCompactThreads + t for r = 0..kCompactPerThread-1,
// so every global read is coalesced and the output stays in node order.
#pragma once

#include "common.h"

namespace svimx {

// Pred: __device__ bool operator()(int64_t node) const.
template <class Pred>
__global__ void __launch_bounds__(kCompactThreads) k_compact_count(Pred pred, int64_t n,
                                                                   int64_t* chunk_counts) {
  __shared__ double red[4];
  const int64_t base = (int64_t)blockIdx.x * kCompactChunk;
  int64_t c = 0;
#pragma unroll 4
  for (int r = 0; r < kCompactPerThread; ++r) {
    int64_t e = base + (int64_t)r * kCompactThreads + threadIdx.x;
    if (e < n && pred(e)) ++c;
  }
  double tot = block_reduce<0>((double)c, red);
  if (threadIdx.x == 0) chunk_counts[blockIdx.x] = (int64_t)tot;
}

// Exclusive scan of `n` counts in place into offs[0..n), total into offs[n].
__global__ void __launch_bounds__(1024) k_scan_counts(int64_t* counts, int64_t n);

// Output value of a selected position: the position itself, or map[position].
struct IdentityMap {
  __device__ int32_t operator()(int64_t e) const { return (int32_t)e; }
};
struct ListMap {
  const int32_t* list;
  __device__ int32_t operator()(int64_t e) const { return list[e]; }
};

template <class Pred, class Map = IdentityMap>
__global__ void __launch_bounds__(kCompactThreads) k_compact_scatter(Pred pred, int64_t n,
                                                                     const int64_t* offs,
                                                                     int32_t* out, int64_t cap, Map map = Map()) {
  __shared__ int wave_tot[kCompactThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = (int64_t)blockIdx.x * kCompactChunk;
  int64_t pos = offs[blockIdx.x];
  for (int r = 0; r < kCompactPerThread; ++r) {
    int64_t e = base + (int64_t)r * kCompactThreads + threadIdx.x;
    bool f = e < n && pred(e);
    uint64_t m = __ballot(f);
    int before = __popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int woff = 0, tot = 0;
    for (int w = 0; w < kCompactThreads / 64; ++w) {
      int t = wave_tot[w];
      woff += (w < wave) ? t : 0;
      tot += t;
    }
    if (f) {
      int64_t o = pos + woff + before;
      if (o < cap) out[o] = map(e);
    }
    pos += tot;
  }
}

// Runs the three passes; synchronises and returns the selected count.
template <class Pred, class Map = IdentityMap>
int compact(hipStream_t st, Pred pred, int64_t n, int32_t* out, int64_t cap, Workspace& ws,
            int64_t* count, Map map = Map()) {
  int64_t chunks = (n + kCompactChunk - 1) / kCompactChunk;
  SVIMX_REQUIRE(chunks + 1 <= ws.max_chunks, SVIMX_EBADARG, "workspace too small for compaction");
  if (chunks == 0) {
    *count = 0;
    return SVIMX_OK;
  }
  hipLaunchKernelGGL(k_compact_count<Pred>, dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, pred,
                     n, ws.chunk_counts);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, ws.chunk_counts, chunks);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_compact_scatter<Pred, Map>), dim3((unsigned)chunks), dim3(kCompactThreads), 0, st,
                     pred, n, ws.chunk_counts, out, cap, map);
  SVIMX_CHECK_LAUNCH();
  int64_t total = 0;
  SVIMX_CHECK_HIP(hipMemcpyAsync(&total, ws.chunk_counts + chunks, sizeof(int64_t),
                                 hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  *count = total;
  SVIMX_REQUIRE(total <= cap, SVIMX_ECAPACITY, "compaction output needs %lld entries, capacity %lld",
                (long long)total, (long long)cap);
  return SVIMX_OK;
}

}  // namespace svimx
