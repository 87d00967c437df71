// libsvimx core: error state, workspace, deterministic reductions, band
// rebuild, and the node-list operator entry points.
#include <cstdarg>
#include <cstring>

#include "common.h"
#include "compact.h"

namespace svimx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- workspace
static constexpr int64_t kHeaderBytes = 4096;

int64_t workspace_bytes_for(int64_t n_nodes) {
  int64_t blocks = (n_nodes + kBandThreads - 1) / kBandThreads + 1;
  int64_t chunks = (n_nodes + kCompactChunk - 1) / kCompactChunk + 2;
  int64_t b = kHeaderBytes + blocks * kPartialSlots * (int64_t)sizeof(double) +
              chunks * (int64_t)sizeof(int64_t);
  return (b + 255) & ~(int64_t)255;
}

bool carve_workspace(void* ws, int64_t ws_bytes, int64_t n_nodes, Workspace& out) {
  if (!ws || ws_bytes < workspace_bytes_for(n_nodes)) {
    set_error("workspace of %lld bytes is smaller than the %lld bytes this grid needs",
              (long long)ws_bytes, (long long)workspace_bytes_for(n_nodes));
    return false;
  }
  char* p = static_cast<char*>(ws);
  out.results = reinterpret_cast<double*>(p);
  out.counters = reinterpret_cast<long long*>(p + 2048);
  out.max_partial_blocks = (n_nodes + kBandThreads - 1) / kBandThreads + 1;
  out.partials = reinterpret_cast<double*>(p + kHeaderBytes);
  out.chunk_counts = reinterpret_cast<int64_t*>(p + kHeaderBytes + out.max_partial_blocks * kPartialSlots *
                                                                      (int64_t)sizeof(double));
  out.max_chunks = (n_nodes + kCompactChunk - 1) / kCompactChunk + 2;
  return true;
}

// ---------------------------------------------------------------- reductions
__global__ void __launch_bounds__(kFinalizeThreads) k_finalize(const double* __restrict__ partials, int64_t nblocks,
                                                  double* __restrict__ results, ReduceSpec spec) {
  finalize_body(partials, nblocks, results, spec);
}

int launch_finalize(hipStream_t st, const double* partials, int64_t nblocks, double* results,
                    const ReduceSpec& spec) {
  SVIMX_REQUIRE(nblocks >= 1 && nblocks <= kMaxReduceBlocks, SVIMX_EBADARG, "finalisation of %lld partial blocks",
                (long long)nblocks);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, st, partials, nblocks, results, spec);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

__global__ void __launch_bounds__(1024) k_scan_counts(int64_t* counts, int64_t n) { scan_counts_body(counts, n); }

// ---------------------------------------------------------------- band rebuild
struct BandPred {
  const float* phi;
  float thr;
  __device__ bool operator()(int64_t e) const { return fabsf(phi[e]) < thr; }
};

// Clamp, copy, reset occupancy, and count inside nodes outside the band and
// the band's coordinate extent.  Partials per block: [0] inside-outside count,
// [1..3] min i/j/k, [4..6] max i/j/k of band nodes.
__global__ void __launch_bounds__(kBandThreads) k_band_classify(float* __restrict__ phi, float* c1, float* c2,
                                                                float* __restrict__ occ, Dims d, float cap,
                                                                float thr, double* __restrict__ partials) {
  __shared__ double red[4];
  const int64_t n = d.size();
  double inside_out = 0.0;
  double mn[3] = {1e30, 1e30, 1e30}, mx[3] = {-1e30, -1e30, -1e30};
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    float v = fminf(fmaxf(phi[e], -cap), cap);
    phi[e] = v;
    if (c1) c1[e] = v;
    if (c2) c2[e] = v;
    int i, j, k;
    unflat(d, e, i, j, k);
    occ[OccLayout::of(d.nx, d.ny, d.nz).at(i, j, k)] = v < 0.0f ? 1.0f : 0.0f;
    bool inb = fabsf(v) < thr;
    if (!inb && v < 0.0f) inside_out += 1.0;
    if (inb) {
      mn[0] = fmin(mn[0], (double)i); mx[0] = fmax(mx[0], (double)i);
      mn[1] = fmin(mn[1], (double)j); mx[1] = fmax(mx[1], (double)j);
      mn[2] = fmin(mn[2], (double)k); mx[2] = fmax(mx[2], (double)k);
    }
  }
  double r = block_reduce<0>(inside_out, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
  for (int a = 0; a < 3; ++a) {
    double lo = block_reduce<2>(mn[a], red);
    if (threadIdx.x == 0) partials[(1 + a) * gridDim.x + blockIdx.x] = lo;
    double hi = block_reduce<1>(mx[a], red);
    if (threadIdx.x == 0) partials[(4 + a) * gridDim.x + blockIdx.x] = hi;
  }
}

}  // namespace svimx

using namespace svimx;

extern "C" {

const char* svimx_last_error(void) { return g_err; }

int svimx_version(void) { return 1; }

int64_t svimx_workspace_bytes(int nx, int ny, int nz) {
  return workspace_bytes_for((int64_t)nx * ny * nz);
}

int svimx_band_rebuild(void* stream, float* phi, float* phi_copy1, float* phi_copy2, float* occupancy, int nx,
                       int ny, int nz, double h, double width, int margin, int32_t* band, int64_t band_cap,
                       void* ws, int64_t ws_bytes, int64_t* info) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  const int64_t n = d.size();
  SVIMX_REQUIRE(nx > 0 && ny > 0 && nz > 0 && n < (int64_t)INT32_MAX, SVIMX_EBADARG, "bad grid %dx%dx%d", nx,
                ny, nz);
  SVIMX_REQUIRE(phi && occupancy && band && info, SVIMX_EBADARG, "null buffer");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, n, w)) return SVIMX_EBADARG;
  hipStream_t st = as_stream(stream);
  const double capd = width * h;
  const float cap = (float)capd;
  const float thr = (float)(capd * (1.0 - 1e-5));
  const int nblk = reduce_grid(n);
  hipLaunchKernelGGL(k_band_classify, dim3(nblk), dim3(kBandThreads), 0, st, phi, phi_copy1, phi_copy2,
                     occupancy, d, cap, thr, w.partials);
  SVIMX_CHECK_LAUNCH();
  ReduceSpec spec;
  for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
  // scratch result slots 24..30 (outside the public SVIMX_R_* range)
  spec.op[0] = 0; spec.dst[0] = 24;
  for (int a = 0; a < 3; ++a) {
    spec.op[1 + a] = 2; spec.dst[1 + a] = 25 + a;
    spec.op[4 + a] = 1; spec.dst[4 + a] = 28 + a;
  }
  int rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  int64_t count = 0;
  rc = compact(st, BandPred{phi, thr}, n, band, band_cap, w, &count);
  info[0] = count;
  if (rc) return rc;
  double res[7];
  SVIMX_CHECK_HIP(hipMemcpyAsync(res, w.results + 24, sizeof(res), hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  info[1] = (int64_t)res[0];
  bool ok = true;
  const int dims[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    info[3 + a] = count ? (int64_t)res[1 + a] : 0;
    info[6 + a] = count ? (int64_t)res[4 + a] : 0;
    if (count && (res[1 + a] < margin || res[4 + a] > dims[a] - 1 - margin)) ok = false;
  }
  info[2] = ok ? 1 : 0;
  if (!ok) {
    set_error("narrow band reaches the grid boundary; enlarge the padding");
    return SVIMX_EMARGIN;
  }
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
