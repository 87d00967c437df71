// Order-preserving stream compaction of a dense predicate into a sorted int32
// index list: per-chunk counts -> one-block exclusive scan -> per-chunk scatter
// with wave ballots.  Chunks are kCompactChunk consecutive nodes; thread t of a
// chunk visits nodes base + r*kCompactThreads + t for r = 0..kCompactPerThread-1,
// so every global read is coalesced and the output stays in node order.
#pragma once

#include "common.h"

namespace svimx {

// Pred: __device__ bool operator()(int64_t node) const.  Chunk c of the
// count pass; the kernels below run chunk blockIdx.x, the batch engine runs
// its jobs' chunks through the same bodies.
template <class Pred>
__device__ __forceinline__ void compact_count_body(const Pred& pred, int64_t n, int64_t* chunk_counts, int64_t c) {
  __shared__ double red[4];
  const int64_t base = c * kCompactChunk;
  int64_t cnt = 0;
#pragma unroll 4
  for (int r = 0; r < kCompactPerThread; ++r) {
    int64_t e = base + (int64_t)r * kCompactThreads + threadIdx.x;
    if (e < n && pred(e)) ++cnt;
  }
  double tot = block_reduce<0>((double)cnt, red);
  if (threadIdx.x == 0) chunk_counts[c] = (int64_t)tot;
}

template <class Pred>
__global__ void __launch_bounds__(kCompactThreads) k_compact_count(Pred pred, int64_t n,
                                                                   int64_t* chunk_counts) {
  compact_count_body(pred, n, chunk_counts, blockIdx.x);
}

// Exclusive scan of n counts in place into counts[0..n), total into
// counts[n]; one block of 1024 threads, thread t owning a contiguous range.
__device__ __forceinline__ void scan_counts_body(int64_t* counts, int64_t n) {
  __shared__ int64_t sums[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t lo = t * per < n ? t * per : n;
  const int64_t hi = lo + per < n ? lo + per : n;
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += counts[i];
  sums[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int64_t v = t >= o ? sums[t - o] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int64_t run = sums[t] - s;  // exclusive prefix of this thread's range
  for (int64_t i = lo; i < hi; ++i) {
    int64_t c = counts[i];
    counts[i] = run;
    run += c;
  }
  if (t == 1023) counts[n] = sums[1023];
  __syncthreads();
}

// Exclusive scan of `n` counts in place into offs[0..n), total into offs[n].
__global__ void __launch_bounds__(1024) k_scan_counts(int64_t* counts, int64_t n);

// Selects the positions whose byte flag is nonzero.
struct FlagPred {
  const uint8_t* f;
  __device__ bool operator()(int64_t e) const { return f[e] != 0; }
};

// Output value of a selected position: the position itself, or map[position].
struct IdentityMap {
  __device__ int32_t operator()(int64_t e) const { return (int32_t)e; }
};
struct ListMap {
  const int32_t* list;
  __device__ int32_t operator()(int64_t e) const { return list[e]; }
};

template <class Pred, class Map>
__device__ __forceinline__ void compact_scatter_body(const Pred& pred, int64_t n, const int64_t* offs, int32_t* out,
                                                     int64_t cap, const Map& map, int64_t c) {
  __shared__ int wave_tot[kCompactThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = c * kCompactChunk;
  int64_t pos = offs[c];
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
  __syncthreads();
}

template <class Pred, class Map = IdentityMap>
__global__ void __launch_bounds__(kCompactThreads) k_compact_scatter(Pred pred, int64_t n,
                                                                     const int64_t* offs,
                                                                     int32_t* out, int64_t cap, Map map = Map()) {
  compact_scatter_body(pred, n, offs, out, cap, map, blockIdx.x);
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

// ---------------------------------------------------------------- device-count form
// The same three passes over a list whose length lives on the device (n_dev,
// clamped to nmax; the launches are sized for nmax): no host synchronisation.
// When *skip_dev exceeds skip_cap the passes do nothing (an earlier stage of
// the caller overflowed its buffers).  The count is left at
// ws.chunk_counts[chunks(nmax)], returned in *count_dev.
template <class Pred>
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_count_dev(Pred pred, const int64_t* n_dev, int64_t nmax, const int64_t* skip_dev, int64_t skip_cap,
                        int64_t* chunk_counts) {
  const int64_t n = *skip_dev > skip_cap ? 0 : min(*n_dev, nmax);
  compact_count_body(pred, n, chunk_counts, blockIdx.x);
}
template <class Pred, class Map>
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_scatter_dev(Pred pred, const int64_t* n_dev, int64_t nmax, const int64_t* skip_dev, int64_t skip_cap,
                          const int64_t* offs, int32_t* out, int64_t cap, Map map) {
  const int64_t n = *skip_dev > skip_cap ? 0 : min(*n_dev, nmax);
  if ((int64_t)blockIdx.x * kCompactChunk >= n) return;  // block-uniform
  compact_scatter_body(pred, n, offs, out, cap, map, blockIdx.x);
}

template <class Pred, class Map = IdentityMap>
int compact_async(hipStream_t st, Pred pred, const int64_t* n_dev, int64_t nmax, const int64_t* skip_dev,
                  int64_t skip_cap, int32_t* out, int64_t cap, Workspace& ws, const int64_t** count_dev,
                  Map map = Map()) {
  const int64_t chunks = (nmax + kCompactChunk - 1) / kCompactChunk;
  SVIMX_REQUIRE(chunks + 1 <= ws.max_chunks, SVIMX_EBADARG, "workspace too small for compaction");
  *count_dev = ws.chunk_counts + chunks;
  if (chunks == 0) {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, ws.chunk_counts, (int64_t)0);
    SVIMX_CHECK_LAUNCH();
    return SVIMX_OK;
  }
  hipLaunchKernelGGL(k_compact_count_dev<Pred>, dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, pred, n_dev,
                     nmax, skip_dev, skip_cap, ws.chunk_counts);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, ws.chunk_counts, chunks);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_compact_scatter_dev<Pred, Map>), dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, pred,
                     n_dev, nmax, skip_dev, skip_cap, ws.chunk_counts, out, cap, map);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

// ---------------------------------------------------------------- counts -> offsets
// Exclusive scan of n per-item counts (int32 or int64) into int64 offsets,
// offs[n] = the total, in the same chunks as the compaction: each block sums
// its chunk with coalesced loads, one block scans the chunk sums, each block
// then scans its chunk from its base (wave prefix sums and one LDS exchange
// per round).  Every pass is parallel over the chunks.
template <typename T>
__global__ void __launch_bounds__(kCompactThreads) k_chunk_sum(const T* __restrict__ cnt, int64_t n,
                                                                int64_t* __restrict__ sums) {
  __shared__ long long wsum[kCompactThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kCompactChunk;
  long long s = 0;
  for (int r = 0; r < kCompactPerThread; ++r) {
    const int64_t e = base + (int64_t)r * kCompactThreads + threadIdx.x;
    if (e < n) s += (long long)cnt[e];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < kCompactThreads / 64; ++w) t += wsum[w];
    sums[blockIdx.x] = (int64_t)t;
  }
}

template <typename T>
__global__ void __launch_bounds__(kCompactThreads) k_chunk_scan(const T* __restrict__ cnt, int64_t n,
                                                                 const int64_t* __restrict__ base_of,
                                                                 int64_t* __restrict__ offs) {
  __shared__ long long wtot[kCompactThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = (int64_t)blockIdx.x * kCompactChunk;
  long long run = (long long)base_of[blockIdx.x];
  for (int r = 0; r < kCompactPerThread; ++r) {
    const int64_t e = base + (int64_t)r * kCompactThreads + threadIdx.x;
    const long long v = e < n ? (long long)cnt[e] : 0;
    long long inc = v;
    for (int o = 1; o < 64; o <<= 1) {
      const long long u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    __syncthreads();  // wtot of the previous round is read
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    long long before = 0, tot = 0;
    for (int w = 0; w < kCompactThreads / 64; ++w) {
      before += w < wave ? wtot[w] : 0;
      tot += wtot[w];
    }
    if (e < n) offs[e] = (int64_t)(run + before + inc - v);
    run += tot;
  }
}

// offs[0..n) = exclusive prefix of cnt, offs[n] = total (stream-ordered, no
// host synchronisation; a temporary of (chunks + 1) int64 on the stream).
template <typename T>
int scan_to_offsets(hipStream_t st, const T* cnt, int64_t n, int64_t* offs) {
  const int64_t chunks = (n + kCompactChunk - 1) / kCompactChunk;
  int64_t* sums = nullptr;
  SVIMX_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&sums), sizeof(int64_t) * (chunks + 1), st));
  if (chunks > 0) {
    hipLaunchKernelGGL(k_chunk_sum<T>, dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, cnt, n, sums);
    SVIMX_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, sums, chunks);
  SVIMX_CHECK_LAUNCH();
  if (chunks > 0) {
    hipLaunchKernelGGL(k_chunk_scan<T>, dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, cnt, n, sums, offs);
    SVIMX_CHECK_LAUNCH();
  }
  SVIMX_CHECK_HIP(hipMemcpyAsync(offs + n, sums + chunks, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
  SVIMX_CHECK_HIP(hipFreeAsync(sums, st));
  return SVIMX_OK;
}

}  // namespace svimx
