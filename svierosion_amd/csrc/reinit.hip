// PDE reinitialisation (Min 2010) on the dilated band, followed by the band
// rebuild (svierosion/levelset.py:154-183).
//
// Work set: the band scattered into a bit mask (one bit per node; z-line
// (x, y) holds wpl = ceil(nz/32) words), dilated by a (2r+1)^3 cube with three
// separable word passes (shifts and ORs along z, ORs of whole words along x
// and y), minus `margin` layers at the grid faces, then compacted in node
// order straight from the set bits -- svierosion/levelset.py:166-174.
//
// Geometry (subcell crossings and local dtau, svierosion/kernels.py:335-359)
// is computed once from phi0.  Only nodes next to the interface have a
// crossing; they keep dtau and the six subcell distances in their slot of
// `geo` (slot = work index), every other node only its 1-byte info word (sign,
// no crossing, dtau = cfl*h), so a sweep streams ~9 bytes for most nodes.
// Measured and dropped (res-400 scan, per sweep): splitting the work set into
// a plain and a crossing list so plain waves skip the subcell arithmetic
// (25 -> 38 us: consecutive crossing nodes sit on different cache lines, so
// their waves' loads stop coalescing); 32-bit load offsets with the six
// reciprocals precomputed in 16-float slots (25 -> 30 us: twice the slot
// bytes and 62 instead of 29 VGPRs for 2% fewer VALU instructions).
// Every count stays on the device (kernels read it, launches are sized for the
// caller's capacity): one host synchronisation per reinitialisation, at the
// end, for the new band size and extent.
//
// Each TVD-RK2 iteration is two Jacobi sweeps ping-ponging between phi and a
// scratch field:  scratch = phi + dtau R(phi);  phi = (phi + scratch +
// dtau R(scratch)) / 2  (svierosion/levelset.py:178-182).  The second sweep
// reads phi only at its own node, so writing phi in place is race-free.
//
// Rebuild: every band node lies in the work set (the work set is the band
// dilated) and no node outside the work set changed, so clamping, the copies
// into the other two stage buffers, the occupancy reset and the new band are
// all computed over the work set alone; the count of inside nodes outside the
// band follows from counts over the old band and the work set.  Counts and the
// band extent use per-block partials and the deterministic finalisation.
#include "common.h"
#include "compact.h"

namespace svimx {

__global__ void k_bits_fill(uint32_t* m, int64_t nw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nw) m[i] = 0u;
}

// Sets the bit of every listed node.  The list is sorted, so the lanes of a
// wave that hit one word are adjacent: a segmented OR over the wave leaves one
// atomic per word and wave instead of one per node.
__global__ void k_bits_scatter(uint32_t* m, const int32_t* __restrict__ idx, int64_t n, int nz, int wpl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = i < n;
  const uint32_t node = ok ? (uint32_t)idx[i] : 0u;
  const uint32_t line = node / (uint32_t)nz, z = node - line * (uint32_t)nz;
  const int w = ok ? (int)(line * (uint32_t)wpl + (z >> 5)) : -1;  // word count < 2^31
  uint32_t bits = ok ? 1u << (z & 31u) : 0u;
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t b2 = __shfl_down(bits, o, 64);
    const int w2 = __shfl_down(w, o, 64);
    if (lane + o < 64 && w2 == w) bits |= b2;
  }
  const int wprev = __shfl_up(w, 1, 64);
  if (ok && (lane == 0 || wprev != w)) atomicOr(&m[w], bits);
}

// bits of word wi of a z-line that are nodes (z < nz)
__device__ __forceinline__ uint32_t valid_bits(int wi, int nz) {
  const int left = nz - 32 * wi;
  return left >= 32 ? 0xffffffffu : ((1u << left) - 1u);
}

// out = OR over dz in [-r, r] (r <= 31) of in, within each z-line.
__global__ void k_bits_dilate_z(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int64_t nw, int wpl,
                                int nz, int r) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  const int wi = (int)((uint64_t)w % (uint32_t)wpl);
  const uint32_t cur = in[w];
  const uint32_t prev = wi > 0 ? in[w - 1] : 0u;
  const uint32_t next = wi < wpl - 1 ? in[w + 1] : 0u;
  uint32_t acc = cur;
  for (int s = 1; s <= r; ++s) {
    acc |= (cur >> s) | (next << (32 - s));  // bit i <- bit i + s
    acc |= (cur << s) | (prev >> (32 - s));  // bit i <- bit i - s
  }
  out[w] = acc & valid_bits(wi, nz);
}

// out = OR over [-r, r] along x (axis 0) or y (axis 1) of whole words; with
// margin > 0 the `margin` layers at every face are cleared.
__global__ void k_bits_dilate_xy(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int nx, int ny,
                                 int wpl, int nz, int axis, int r, int margin) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = (int64_t)nx * ny * wpl;
  if (w >= nw) return;
  const int64_t xy = w / wpl;
  const int wi = (int)(w - xy * wpl);
  const int y = (int)(xy % ny), x = (int)(xy / ny);
  const int c = axis == 0 ? x : y, n = axis == 0 ? nx : ny;
  const int64_t st = axis == 0 ? (int64_t)ny * wpl : (int64_t)wpl;
  const int lo = c - r < 0 ? -c : -r, hi = c + r > n - 1 ? n - 1 - c : r;
  uint32_t acc = 0u;
  for (int q = lo; q <= hi; ++q) acc |= in[w + q * st];
  if (margin > 0) {
    if (x < margin || x > nx - 1 - margin || y < margin || y > ny - 1 - margin) acc = 0u;
    // keep z in [margin, nz - 1 - margin]
    const int z0 = 32 * wi;
    for (int b = 0; b < 32; ++b) {
      const int z = z0 + b;
      if (z < margin || z > nz - 1 - margin) acc &= ~(1u << b);
    }
  }
  out[w] = acc & valid_bits(wi, nz);
}

// The bit lists are compacted in chunks of kBitsChunk words: a few hundred
// words per block, so even a res-400 grid (1.2M words) spreads over thousands
// of blocks (chunks of kCompactChunk words gave 284 blocks doing 16 serial
// rounds each).
constexpr int kBitsRounds = 2;
constexpr int64_t kBitsChunk = (int64_t)kCompactThreads * kBitsRounds;

// Set bits per compaction chunk (kBitsChunk words).
__global__ void __launch_bounds__(kCompactThreads) k_bits_count(const uint32_t* __restrict__ m, int64_t nw,
                                                                int64_t* chunk_counts) {
  __shared__ double red[4];
  const int64_t base = (int64_t)blockIdx.x * kBitsChunk;
  int64_t cnt = 0;
  for (int r = 0; r < kBitsRounds; ++r) {
    const int64_t w = base + (int64_t)r * kCompactThreads + threadIdx.x;
    if (w < nw) cnt += __popc(m[w]);
  }
  const double tot = block_reduce<0>((double)cnt, red);
  if (threadIdx.x == 0) chunk_counts[blockIdx.x] = (int64_t)tot;
}

// Exclusive prefix of v over the block (kCompactThreads threads); *total = sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds_waves, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  __syncthreads();
  if (lane == 63) lds_waves[wave] = inc;
  __syncthreads();
  int before = 0, tot = 0;
  for (int q = 0; q < kCompactThreads / 64; ++q) {
    const int t = lds_waves[q];
    before += q < wave ? t : 0;
    tot += t;
  }
  *total = tot;
  return before + inc - v;
}

// Node indices of the set bits, in node order, at the scanned chunk offsets.
// Per round the block scans its words' bit counts; then each wave expands its
// non-empty words two at a time, lanes 0-31 one word's bits and lanes 32-63
// the other's, so the stores of a word's nodes are consecutive (one lane per
// set bit, its rank by a popcount of the lower bits).
__global__ void __launch_bounds__(kCompactThreads) k_bits_scatter_list(const uint32_t* __restrict__ m, int64_t nw,
                                                                       int wpl, int nz, const int64_t* offs,
                                                                       int32_t* __restrict__ out, int64_t cap) {
  __shared__ int wave_tot[kCompactThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kBitsChunk;
  const int lane = threadIdx.x & 63, wave0 = threadIdx.x & ~63;
  const int bit = lane & 31;
  int64_t pos = offs[blockIdx.x];
  for (int r = 0; r < kBitsRounds; ++r) {
    const int64_t w0 = base + (int64_t)r * kCompactThreads + wave0;  // this wave's first word
    const int64_t w = w0 + lane;
    const uint32_t bits = w < nw ? m[w] : 0u;
    int total;
    const int excl = block_exclusive_scan(__popc(bits), wave_tot, &total);
    const unsigned long long nzm = __ballot(bits != 0u);
    uint32_t pairs = (uint32_t)nzm | (uint32_t)(nzm >> 32);
    while (pairs) {  // wave-uniform
      const int j = __ffs(pairs) - 1;
      pairs &= pairs - 1u;
      const int src = j + (lane & 32);
      const uint32_t b = (uint32_t)__shfl((int)bits, src, 64);
      const int e = __shfl(excl, src, 64);
      if ((b >> bit) & 1u) {
        const int64_t ws = w0 + src;
        const int64_t line = ws / wpl;
        const int64_t o = pos + e + __popc(b & ((1u << bit) - 1u));
        if (o < cap) out[o] = (int32_t)(line * nz + (ws - line * wpl) * 32 + bit);
      }
    }
    pos += total;
  }
}

// Order-preserving list of the set bits of m; the count stays on the device
// at ws.chunk_counts[chunks] (*count_dev), no host synchronisation.
static int compact_bits(hipStream_t st, const uint32_t* m, int64_t nw, int wpl, int nz, int32_t* out, int64_t cap,
                        Workspace& ws, const int64_t** count_dev) {
  const int64_t chunks = (nw + kBitsChunk - 1) / kBitsChunk;
  SVIMX_REQUIRE(chunks + 1 <= ws.max_chunks, SVIMX_EBADARG, "workspace too small for compaction");
  *count_dev = ws.chunk_counts + chunks;
  if (chunks > 0) {
    hipLaunchKernelGGL(k_bits_count, dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, m, nw, ws.chunk_counts);
    SVIMX_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, ws.chunk_counts, chunks);
  SVIMX_CHECK_LAUNCH();
  if (chunks > 0) {
    hipLaunchKernelGGL(k_bits_scatter_list, dim3((unsigned)chunks), dim3(kCompactThreads), 0, st, m, nw, wpl, nz,
                       ws.chunk_counts, out, cap);
    SVIMX_CHECK_LAUNCH();
  }
  return SVIMX_OK;
}

// Device counters of one reinitialisation (integer slots of the workspace
// header): the work-set size.  A work set larger than the caller's capacity leaves every kernel
// after the compaction idle (the field is not touched) and is reported at
// the end, so the caller can retry with the exact size.
enum { kCtrWork = 0, kCtrGate = 1 };

// The work count the later kernels read; a closed gate leaves it 0, so every
// kernel after this one (geometry, sweeps, rebuild, band compaction) changes
// nothing.
__global__ void k_reinit_counters(long long* ctr, const int64_t* work_total, int gated) {
  if (threadIdx.x == 0) {
    ctr[kCtrWork] = (gated && ctr[kCtrGate] == 0) ? 0 : (long long)*work_total;
  }
}

// The reinitialisation trigger of ErosionSimulation.run (svierosion/evolve.py:
// 232-245) for the step whose results (stage-0 statistics, device dt, volume)
// sit in the workspace: the step is accepted (finite positive dt), needs no
// secant retry (volume >= target), does not reach the target and moves the
// displacement to the threshold.  FP64 in the host's operation order, without
// contraction, so the device opens the gate exactly when the host decides to
// reinitialise.
__global__ void k_reinit_gate(const double* __restrict__ results, long long* ctr, svimx_reinit_gate_t g) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0) return;
  const double dt = results[SVIMX_R_DT], vmax = results[SVIMX_R_VMAX];
  const double vnew = (g.inside_outside + results[SVIMX_R_VOLUME_SUM]) * g.h3;
  bool open = dt > 0.0 && dt < INFINITY;
  if (g.has_target && vnew < g.target) open = false;
  const double disp = g.displacement + dt * vmax;
  const bool reached = g.has_target && vnew <= g.target_stop;
  ctr[kCtrGate] = (open && disp >= g.threshold && !reached) ? 1 : 0;
}

__device__ __forceinline__ int64_t work_count(const long long* ctr, int64_t cap) {
  const int64_t n = (int64_t)ctr[kCtrWork];
  return n > cap ? 0 : n;  // overflow: nothing runs
}

// partials[slot * G + b] = #{i in block b's share : phi[list[i]] < 0}; the list
// length is n, or the device work count when ctr is given.
__global__ void __launch_bounds__(kBandThreads)
    k_count_neg(const float* __restrict__ phi, const int32_t* __restrict__ list, int64_t n, const long long* ctr,
                int64_t cap, double* __restrict__ partials, int slot) {
  __shared__ double red[4];
  if (ctr) n = work_count(ctr, cap);
  double c = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += phi[list[i]] < 0.0f ? 1.0 : 0.0;
  c = block_reduce<0>(c, red);
  if (threadIdx.x == 0) partials[slot * gridDim.x + blockIdx.x] = c;
}

// info word: bits 0-5 crossing flags, bits 6-7 sign code.  Nodes with a
// crossing store dtau and the six distances at geo[8 i].
__global__ void k_reinit_geo(const float* __restrict__ phi, const int32_t* __restrict__ idx, const long long* ctr,
                             int64_t cap, Dims d, float h, float cfl, uint8_t* __restrict__ info,
                             float* __restrict__ geo) {
  const int64_t n = work_count(ctr, cap);
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ReinitGeo g = reinit_geometry<double>(phi, idx[i], d, h, cfl);
  if (g.flags & 63u) {
    float* o = geo + 8 * i;
    o[0] = g.dtau;
    for (int q = 0; q < 6; ++q) o[1 + q] = g.x[q];
  }
  info[i] = (uint8_t)(g.flags & 0xFFu);
}

// Mode 0: out = in + dtau R(in);  mode 1: out = (out + in + dtau R(in)) / 2.
// Each thread takes kSweepNodes work nodes kSweepThreads apart (coalesced) and
// issues the stencil loads of all of them before computing any, so a wave
// keeps kSweepNodes x 13 gathers in flight.  The launch is sized for the
// capacity; blocks past the device work count exit.  Off = uint32_t addresses
// the field by 32-bit byte offsets (grids below 2^30 nodes), int64_t beyond.
constexpr int kSweepNodes = 2;
constexpr int kSweepThreads = 256;
constexpr int kSweepPerBlock = kSweepThreads * kSweepNodes;
template <int Mode, typename Off>
__global__ void __launch_bounds__(kSweepThreads)
    k_reinit_sweep(const float* __restrict__ in, float* out, const int32_t* __restrict__ idx, const long long* ctr,
                   int64_t cap, Off sx4, Off sy4, float ih, float ihh, float dtau0,
                   const uint8_t* __restrict__ info, const float* __restrict__ geo) {
  const uint32_t n = (uint32_t)work_count(ctr, cap);  // < 2^31
  // XCD-contiguous slices of the (sorted) work list keep stencil neighbours in
  // one L2: the blocks needed for n, taken in XCD order
  const uint32_t need = (n + kSweepPerBlock - 1) / kSweepPerBlock;
  const uint32_t g8 = (need + kXcds - 1) / kXcds;  // blocks per XCD
  const uint32_t xcd = blockIdx.x % kXcds, lbx = blockIdx.x / kXcds;
  if (lbx >= g8) return;  // block-uniform
  const uint32_t base = (xcd * g8 + lbx) * kSweepPerBlock + threadIdx.x;
  int32_t node[kSweepNodes];
  uint32_t word[kSweepNodes];
#pragma unroll
  for (int k = 0; k < kSweepNodes; ++k) {
    const uint32_t i = base + k * kSweepThreads;
    node[k] = i < n ? idx[i] : -1;
    word[k] = i < n ? info[i] : 0u;
  }
  Stencil13 S[kSweepNodes];
  float prev[kSweepNodes];
#pragma unroll
  for (int k = 0; k < kSweepNodes; ++k) {
    if (node[k] >= 0) {
      S[k] = load_stencil13_at<Off>(in, (Off)node[k], sx4, sy4);
      prev[k] = Mode == 0 ? 0.0f : out[node[k]];
    }
  }
#pragma unroll
  for (int k = 0; k < kSweepNodes; ++k) {
    if (node[k] < 0) continue;
    const uint32_t i = base + k * kSweepThreads;
    float dtau = dtau0;
    float x[6] = {0, 0, 0, 0, 0, 0};
    if (word[k] & 63u) {
      const float* g = geo + 8 * (size_t)i;
      dtau = g[0];
      for (int q = 0; q < 6; ++q) x[q] = g[1 + q];
    }
    const uint32_t sb = (word[k] >> 6) & 3u;
    const float s = sb == 1u ? 1.0f : (sb == 2u ? -1.0f : 0.0f);
    const float r = reinit_rhs_lean(S[k], s, ih, ihh, word[k], x);
    if (Mode == 0) {
      out[node[k]] = S[k].c + dtau * r;
    } else {
      out[node[k]] = 0.5f * (prev[k] + S[k].c + dtau * r);
    }
  }
}

template <typename Off>
static void launch_sweeps(hipStream_t st, int iterations, float* phi, float* scratch, const int32_t* work_idx,
                          const long long* ctr, int64_t work_cap, const Dims& d, float h, float dtau0,
                          const uint8_t* info_b, const float* geo) {
  // blocks for the largest work set, kXcds times the per-XCD share (idle ones exit)
  const int64_t need = (work_cap + kSweepPerBlock - 1) / kSweepPerBlock;
  const int gs = (int)(((need + kXcds - 1) / kXcds) * kXcds);
  const Off sx4 = (Off)(d.sx() * 4), sy4 = (Off)(d.sy() * 4);
  const float ih = 1.0f / h, ihh = ih * ih;
  for (int it = 0; it < iterations; ++it) {
    hipLaunchKernelGGL((k_reinit_sweep<0, Off>), dim3(gs), dim3(kSweepThreads), 0, st, phi, scratch, work_idx, ctr,
                       work_cap, sx4, sy4, ih, ihh, dtau0, info_b, geo);
    hipLaunchKernelGGL((k_reinit_sweep<1, Off>), dim3(gs), dim3(kSweepThreads), 0, st, scratch, phi, work_idx, ctr,
                       work_cap, sx4, sy4, ih, ihh, dtau0, info_b, geo);
  }
}

// Clamp the work set, copy it into the other stage buffers and reset its
// occupancy.  Partials: slot 2 inside-and-outside-the-new-band count, slots
// 3..5 / 6..8 min / max i, j, k of the new band.
__global__ void __launch_bounds__(kBandThreads)
    k_rebuild_work(float* __restrict__ phi, float* __restrict__ c1, float* __restrict__ c2, float* __restrict__ occ,
                   const int32_t* __restrict__ idx, const long long* ctr, int64_t cap, Dims d, float capv, float thr,
                   double* __restrict__ partials) {
  __shared__ double red[4];
  const int64_t n = work_count(ctr, cap);
  const OccLayout OL = OccLayout::of(d.nx, d.ny, d.nz);
  double outneg = 0.0;
  double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = idx[i];
    const float v = fminf(fmaxf(phi[node], -capv), capv);
    phi[node] = v;
    c1[node] = v;
    c2[node] = v;
    int a, b, c;
    unflat(d, node, a, b, c);
    occ[OL.at(a, b, c)] = v < 0.0f ? 1.0f : 0.0f;
    const bool inb = fabsf(v) < thr;
    if (!inb && v < 0.0f) outneg += 1.0;
    if (inb) {
      lo[0] = fmin(lo[0], (double)a); hi[0] = fmax(hi[0], (double)a);
      lo[1] = fmin(lo[1], (double)b); hi[1] = fmax(hi[1], (double)b);
      lo[2] = fmin(lo[2], (double)c); hi[2] = fmax(hi[2], (double)c);
    }
  }
  double r = block_reduce<0>(outneg, red);
  if (threadIdx.x == 0) partials[2 * gridDim.x + blockIdx.x] = r;
  for (int q = 0; q < 3; ++q) {
    double l = block_reduce<2>(lo[q], red);
    if (threadIdx.x == 0) partials[(3 + q) * gridDim.x + blockIdx.x] = l;
    double u = block_reduce<1>(hi[q], red);
    if (threadIdx.x == 0) partials[(6 + q) * gridDim.x + blockIdx.x] = u;
  }
}

struct WorkBandPred {
  const float* phi;
  const int32_t* idx;
  float thr;
  __device__ bool operator()(int64_t i) const { return fabsf(phi[idx[i]]) < thr; }
};

// Result scratch slots 19..30 (outside the public SVIMX_R_* range): the
// negative counts and the new band's extent, then the work-set and band sizes
// and the gate, so one copy brings all of them to the host.
constexpr int kRsOldNeg = 19, kRsWorkNeg = 20, kRsOutNeg = 21, kRsLo = 22, kRsHi = 25, kRsWork = 28,
              kRsBand = 29, kRsGate = 30, kRsEnd = 31;

__global__ void k_reinit_readout(double* __restrict__ results, const int64_t* __restrict__ work_ctr,
                                 const int64_t* __restrict__ band_count, const long long* __restrict__ ctr,
                                 int gated) {
  if (threadIdx.x == 0) {
    results[kRsWork] = (double)*work_ctr;
    results[kRsBand] = (double)*band_count;
    results[kRsGate] = gated ? (double)ctr[kCtrGate] : 1.0;
  }
}

// Host-visible (pinned) landing area of the readback, one per host thread.
static double* pinned_readout() {
  thread_local double* buf = nullptr;
  if (!buf && hipHostMalloc(reinterpret_cast<void**>(&buf), 16 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    buf = nullptr;
  return buf;
}

}  // namespace svimx

using namespace svimx;

extern "C" {

// Sequence of one reinitialisation, all on `stream` with a single host
// synchronisation at the end: band bits -> dilation -> work list (count on the
// device) -> negatives of the old band and of the work set -> geometry and the
// plain / crossing split -> 2 x iterations sweeps -> rebuild over the work set
// -> new band (order-preserving compaction of the work list) -> one readback of
// the counts and extents.
static int reinit(void* stream, float* phi, float* scratch, float* phi_copy, float* occupancy, int32_t* band,
                  int64_t nb, int64_t inside_outside, int nx, int ny, int nz, double h, double width, int iterations,
                  double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b, int32_t* work_idx,
                  int32_t* work_info, float* geo, int64_t work_cap, void* ws, int64_t ws_bytes,
                  const svimx_reinit_gate_t* gate, int64_t* info) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  const int64_t n = d.size();
  SVIMX_REQUIRE(phi && scratch && phi_copy && occupancy && band && mask_a && mask_b && work_idx && work_info && geo &&
                    info,
                SVIMX_EBADARG, "null reinit buffer");
  SVIMX_REQUIRE(phi != scratch && phi != phi_copy && scratch != phi_copy, SVIMX_EBADARG, "aliased reinit buffers");
  SVIMX_REQUIRE(iterations >= 0 && dilation >= 0 && margin >= 2 && work_cap >= 0, SVIMX_EBADARG,
                "bad reinit parameters (the sweeps' stencils reach 2 nodes: margin >= 2)");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, n, w)) return SVIMX_EBADARG;
  // the work set holds at most every node: larger buffers are used up to n
  // (the compactions and the workspace's chunk counts are sized for n)
  work_cap = std::min<int64_t>(work_cap, n);
  hipStream_t st = as_stream(stream);
  const dim3 tb(256);
  ReduceSpec spec;
  auto one_slot = [&](int slot, int op, int dst) {
    for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
    spec.op[slot] = (int8_t)op;
    spec.dst[slot] = (int8_t)dst;
  };
  // bit masks in the caller's n-byte mask buffers
  const int wpl = (nz + 31) / 32;
  const int64_t nwords = (int64_t)nx * ny * wpl;
  SVIMX_REQUIRE(nwords * 4 <= n && (reinterpret_cast<uintptr_t>(mask_a) & 3) == 0 &&
                    (reinterpret_cast<uintptr_t>(mask_b) & 3) == 0 && dilation <= 31,
                SVIMX_EBADARG, "reinit masks need nz >= 4, 4-byte alignment and dilation <= 31");
  uint32_t* ba = reinterpret_cast<uint32_t*>(mask_a);
  uint32_t* bb = reinterpret_cast<uint32_t*>(mask_b);
  const dim3 gwd(blocks_for(nwords, 256));
  const int gated = gate ? 1 : 0;
  if (gate) {
    hipLaunchKernelGGL(k_reinit_gate, dim3(1), dim3(64), 0, st, w.results, w.counters, *gate);
    SVIMX_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_bits_fill, gwd, tb, 0, st, ba, nwords);
  SVIMX_CHECK_LAUNCH();
  int nblk = reduce_grid(nb);
  hipLaunchKernelGGL(k_count_neg, dim3(nblk), tb, 0, st, phi, band, nb, nullptr, (int64_t)0, w.partials, 0);
  SVIMX_CHECK_LAUNCH();
  one_slot(0, 0, kRsOldNeg);
  int rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  if (nb > 0) {
    hipLaunchKernelGGL(k_bits_scatter, dim3(blocks_for(nb, 256)), tb, 0, st, ba, band, nb, nz, wpl);
    SVIMX_CHECK_LAUNCH();
  }
  // z, x, y passes: a -> b -> a -> b; margin cleared on the last
  hipLaunchKernelGGL(k_bits_dilate_z, gwd, tb, 0, st, ba, bb, nwords, wpl, nz, dilation);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bits_dilate_xy, gwd, tb, 0, st, bb, ba, nx, ny, wpl, nz, 0, dilation, 0);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bits_dilate_xy, gwd, tb, 0, st, ba, bb, nx, ny, wpl, nz, 1, dilation, margin);
  SVIMX_CHECK_LAUNCH();
  const int64_t* work_total = nullptr;
  rc = compact_bits(st, bb, nwords, wpl, nz, work_idx, work_cap, w, &work_total);
  if (rc) return rc;
  long long* ctr = w.counters;
  hipLaunchKernelGGL(k_reinit_counters, dim3(1), dim3(64), 0, st, ctr, work_total, gated);
  SVIMX_CHECK_LAUNCH();
  const float capf = (float)(width * h);
  const float thr = (float)(width * h * (1.0 - 1e-5));
  nblk = reduce_grid(work_cap);
  hipLaunchKernelGGL(k_count_neg, dim3(nblk), tb, 0, st, phi, work_idx, (int64_t)0, ctr, work_cap, w.partials, 1);
  SVIMX_CHECK_LAUNCH();
  one_slot(1, 0, kRsWorkNeg);
  rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  if (work_cap > 0) {
    uint8_t* info_b = reinterpret_cast<uint8_t*>(work_info);
    hipLaunchKernelGGL(k_reinit_geo, dim3(blocks_for(work_cap, 256)), tb, 0, st, phi, work_idx, ctr, work_cap, d,
                       (float)h, (float)cfl, info_b, geo);
    SVIMX_CHECK_LAUNCH();
    const float dtau0 = (float)cfl * fmaxf((float)h, 1e-3f * (float)h);
    if (n * 4 < ((int64_t)1 << 32))
      launch_sweeps<uint32_t>(st, iterations, phi, scratch, work_idx, ctr, work_cap, d, (float)h, dtau0, info_b, geo);
    else
      launch_sweeps<int64_t>(st, iterations, phi, scratch, work_idx, ctr, work_cap, d, (float)h, dtau0, info_b, geo);
    SVIMX_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_rebuild_work, dim3(nblk), tb, 0, st, phi, scratch, phi_copy, occupancy, work_idx, ctr,
                     work_cap, d, capf, thr, w.partials);
  SVIMX_CHECK_LAUNCH();
  for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
  spec.op[2] = 0; spec.dst[2] = kRsOutNeg;
  for (int q = 0; q < 3; ++q) {
    spec.op[3 + q] = 2; spec.dst[3 + q] = (int8_t)(kRsLo + q);
    spec.op[6 + q] = 1; spec.dst[6 + q] = (int8_t)(kRsHi + q);
  }
  rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  const int64_t* band_count = nullptr;
  const int64_t* work_ctr = reinterpret_cast<const int64_t*>(ctr + kCtrWork);
  rc = compact_async(st, WorkBandPred{phi, work_idx, thr}, work_ctr, work_cap, work_ctr, work_cap, band, n, w,
                     &band_count, ListMap{work_idx});
  if (rc) return rc;
  hipLaunchKernelGGL(k_reinit_readout, dim3(1), dim3(64), 0, st, w.results, work_ctr, band_count, ctr, gated);
  SVIMX_CHECK_LAUNCH();
  double* r = pinned_readout();
  SVIMX_REQUIRE(r, SVIMX_EDEVICE, "reinitialisation: no pinned host memory for the readback");
  SVIMX_CHECK_HIP(hipMemcpyAsync(r, w.results + kRsOldNeg, sizeof(double) * (kRsEnd - kRsOldNeg),
                                 hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  const int64_t nw = (int64_t)r[kRsWork - kRsOldNeg], nbn = (int64_t)r[kRsBand - kRsOldNeg];
  if (gate) {
    info[10] = r[kRsGate - kRsOldNeg] != 0.0 ? 1 : 0;
    if (!info[10]) {  // closed: nothing changed; the band is the one passed in
      info[9] = 0;
      info[0] = nb;
      info[1] = inside_outside;
      info[2] = 1;
      for (int a = 0; a < 6; ++a) info[3 + a] = 0;
      return SVIMX_OK;
    }
  }
  info[9] = nw;
  SVIMX_REQUIRE(nw <= work_cap, SVIMX_ECAPACITY, "reinitialisation work set needs %lld entries, capacity %lld",
                (long long)nw, (long long)work_cap);
  info[0] = nbn;
  // inside nodes outside the band: (all negatives before) - (work negatives before) + (new outside-band negatives)
  info[1] = inside_outside + (int64_t)r[0] - (int64_t)r[1] + (int64_t)r[2];
  const int dims[3] = {nx, ny, nz};
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    info[3 + a] = nbn ? (int64_t)r[3 + a] : 0;
    info[6 + a] = nbn ? (int64_t)r[6 + a] : 0;
    if (nbn && (r[3 + a] < margin || r[6 + a] > dims[a] - 1 - margin)) ok = false;
  }
  info[2] = ok ? 1 : 0;
  if (!ok) {
    set_error("narrow band reaches the grid boundary; enlarge the padding");
    return SVIMX_EMARGIN;
  }
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_reinit(void* stream, float* phi, float* scratch, float* phi_copy, float* occupancy, int32_t* band,
                 int64_t nb, int64_t inside_outside, int nx, int ny, int nz, double h, double width, int iterations,
                 double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b, int32_t* work_idx,
                 int32_t* work_info, float* geo, int64_t work_cap, void* ws, int64_t ws_bytes, int64_t* info) {
  return reinit(stream, phi, scratch, phi_copy, occupancy, band, nb, inside_outside, nx, ny, nz, h, width,
                iterations, cfl, dilation, margin, mask_a, mask_b, work_idx, work_info, geo, work_cap, ws, ws_bytes,
                nullptr, info);
}

int svimx_reinit_gated(void* stream, float* phi, float* scratch, float* phi_copy, float* occupancy, int32_t* band,
                       int64_t nb, int64_t inside_outside, int nx, int ny, int nz, double h, double width,
                       int iterations, double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b,
                       int32_t* work_idx, int32_t* work_info, float* geo, int64_t work_cap, void* ws,
                       int64_t ws_bytes, const svimx_reinit_gate_t* gate, int64_t* info) {
  SVIMX_REQUIRE(gate, SVIMX_EBADARG, "svimx_reinit_gated: null gate");
  return reinit(stream, phi, scratch, phi_copy, occupancy, band, nb, inside_outside, nx, ny, nz, h, width,
                iterations, cfl, dilation, margin, mask_a, mask_b, work_idx, work_info, geo, work_cap, ws, ws_bytes,
                gate, info);
}

}  // extern "C"
