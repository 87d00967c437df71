// Host-side plumbing shared by the libsvimx translation units: error state,
// launch checks, reduction helpers and the workspace layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "../../include/svimx.h"
#include "ops.h"

namespace svimx {

void set_error(const char* fmt, ...);

#define SVIMX_CHECK_HIP(expr)                                                       \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      ::svimx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),     \
                         __FILE__, __LINE__);                                       \
      return SVIMX_EDEVICE;                                                         \
    }                                                                               \
  } while (0)

#define SVIMX_CHECK_LAUNCH() SVIMX_CHECK_HIP(hipGetLastError())

#define SVIMX_REQUIRE(cond, code, ...)         \
  do {                                         \
    if (!(cond)) {                             \
      ::svimx::set_error(__VA_ARGS__);         \
      return (code);                           \
    }                                          \
  } while (0)

#define SVIMX_TRY_BEGIN try {
#define SVIMX_TRY_END                                          \
  }                                                            \
  catch (const std::exception& ex) {                           \
    ::svimx::set_error("exception: %s", ex.what());            \
    return SVIMX_EDEVICE;                                      \
  }                                                            \
  catch (...) {                                                \
    ::svimx::set_error("unknown exception");                   \
    return SVIMX_EDEVICE;                                      \
  }

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int blocks_for(int64_t n, int threads) { return (int)((n + threads - 1) / threads); }

// ---------------------------------------------------------------- workspace
// One per job, caller-allocated (svimx_workspace_bytes).  Fixed carve-up:
//   [0, 4 KiB)            results (doubles, SVIMX_R_* slots)
//   partials              kPartialSlots doubles per block of kBandThreads (slot-major)
//   compaction counts     one int64 per compaction chunk (+ scan)
struct Workspace {
  double* results;
  long long* counters;  // 256 integer counters (header bytes 2048..4095)
  double* partials;
  int64_t max_partial_blocks;
  int64_t* chunk_counts;
  int64_t max_chunks;
};

constexpr int kBandThreads = 256;       // threads per block of band-list kernels
constexpr int kPartialSlots = 16;       // doubles written per block by reduction kernels
constexpr int kMaxReduceBlocks = 2048;  // grid cap of reducing kernels (grid-stride loops)
constexpr int kXcds = 8;                // MI355X: workgroups go round-robin over 8 XCDs

// Grid of a reducing kernel over n items: grid-stride with at most kMaxReduceBlocks
// blocks, so the finalisation reads a bounded number of partials.  Always a
// multiple of the XCD count: the velocity pass gives each XCD a contiguous
// slice of the band only on such grids, and in the batch engine every job's
// segment of these grids then starts on a multiple of 8, so a virtual block's
// local index mod 8 is its XCD (surplus blocks write identity partials).
__host__ __device__ inline int reduce_grid(int64_t n) {
  int64_t b = (n + kBandThreads - 1) / kBandThreads;
  b = b < 1 ? 1 : b;
  b = (b + kXcds - 1) / kXcds * kXcds;
  return (int)(b > kMaxReduceBlocks ? kMaxReduceBlocks : b);
}

// Block remap giving XCD x (blocks go round-robin over the 8 XCDs, each with
// its own L2) a contiguous slice of a G-block range: G must be a multiple of 8
// and the range must start at a block index that is a multiple of 8.
__host__ __device__ inline int64_t xcd_slice(int64_t b, int64_t G) { return (b % kXcds) * (G / kXcds) + b / kXcds; }
constexpr int kCompactThreads = 256;
constexpr int kCompactPerThread = 16;
constexpr int64_t kCompactChunk = (int64_t)kCompactThreads * kCompactPerThread;

int64_t workspace_bytes_for(int64_t n_nodes);
bool carve_workspace(void* ws, int64_t ws_bytes, int64_t n_nodes, Workspace& out);

// ---------------------------------------------------------------- device reductions
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int64_t wave_sum_i(int64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reduction through LDS for blockDim.x == kBandThreads (4 waves).
// op: 0 sum, 1 max, 2 min.  Result valid in thread 0.
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* lds4) {
  for (int o = 32; o > 0; o >>= 1) {
    double u = __shfl_xor(v, o, 64);
    v = OP == 0 ? v + u : (OP == 1 ? fmax(v, u) : fmin(v, u));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds4[wave] = v;
  __syncthreads();
  double r = lds4[0];
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 1; w < nw; ++w) {
    double u = lds4[w];
    r = OP == 0 ? r + u : (OP == 1 ? fmax(r, u) : fmin(r, u));
  }
  return r;
}

// Launch-wide finalisation: reduce kPartialSlots x nblocks partials in a fixed
// order (deterministic) and store slot s into results[dst[s]] with op[s].
// ops[slot]: 0 sum, 1 max, 2 min, -1 skip.
struct ReduceSpec {
  int8_t op[kPartialSlots];
  int8_t dst[kPartialSlots];
};
int launch_finalize(hipStream_t st, const double* partials, int64_t nblocks, double* results,
                    const ReduceSpec& spec);

__device__ __forceinline__ double fold_op(int op, double a, double b) {
  return op == 0 ? a + b : (op == 1 ? fmax(a, b) : fmin(a, b));
}

// Block partials are slot-major: the reducing kernel of a G-block grid stores
// slot s of block b at partials[s * G + b], and the finalisation reads them
// back with the same G (= its nblocks).  Slot s is folded by one wave: lane l
// loads rows l, l + 64, ... (each load instruction reads 512 contiguous
// bytes, 16 loads in flight before their sequential fold), then the wave
// reduces with the xor tree 32, 16, ..., 1.  Unused slots cost nothing.  The
// order depends on neither the block size nor the number of active slots, so
// a job's results are the same from the single-job kernels and from the batch
// engine.  (Row-major 128-byte rows read by one lane each, the earlier
// layout, made every load instruction touch 64 cache lines: 13-19 us per
// finalisation against the few us of its data.  Folding the rows in the
// producer's last block instead of a separate launch was measured slower as
// well: every block's agent-scope release writes back its XCD's L2.)
constexpr int kFinalizeThreads = 1024;
constexpr int kFinalizeRows = kMaxReduceBlocks / 64;  // rows per lane
__device__ __forceinline__ int spec_byte(const int8_t (&a)[kPartialSlots], int s) {
  uint64_t w[2];
  __builtin_memcpy(w, a, sizeof(w));
  return (int)(int8_t)((s < 8 ? w[0] : w[1]) >> (8 * (s & 7)));
}
__device__ __forceinline__ void finalize_body(const double* __restrict__ partials, int64_t nblocks,
                                              double* __restrict__ results, const ReduceSpec& spec) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwr = (int)(blockDim.x >> 6);
  for (int s = wave; s < kPartialSlots; s += nwr) {
    const int op = spec_byte(spec.op, s);
    if (op < 0) continue;  // wave-uniform
    const double id = op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY);
    const double* col = partials + (int64_t)s * nblocks;
    double v = id;
    // rows in batches of 16 per lane (all loads of a batch in flight); a
    // batch past nblocks would fold only identities and is skipped
    for (int i0 = 0; i0 < kFinalizeRows && 64 * i0 < nblocks; i0 += 16) {
      double x[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int64_t b = lane + 64 * (i0 + i);
        x[i] = b < nblocks ? col[b] : id;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) v = fold_op(op, v, x[i]);
    }
    for (int o = 32; o > 0; o >>= 1) v = fold_op(op, v, __shfl_xor(v, o, 64));
    if (lane == 0) results[spec_byte(spec.dst, s)] = v;
  }
  __syncthreads();  // callers read results[] from other threads of the block
}

}  // namespace svimx
