// Per-stage band kernels: cell fractions (+ volume, interface flags, N_r tile
// marks), curvature scatter, velocity (+ dt statistics), TVD-RK3 stage update,
// and the node-list operator entry points that share their device code.
#include <mutex>

#include "band_dev.h"

namespace svimx {

// ---------------------------------------------------------------- fractions
__global__ void __launch_bounds__(kBandThreads)
    k_fractions(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
                float* __restrict__ occ, uint8_t* __restrict__ marks, int run_mode, int nty, int ntz, int nzr,
                float cut, Hi3 hi, double* __restrict__ partials) {
  fractions_body(phi, band, nb, d, h, occ, marks, run_mode, nty, ntz, nzr, cut, hi, partials, blockIdx.x, gridDim.x);
}

// 16 bytes per thread: p must be 16-byte aligned (the plan's hipMalloc'd marks)
__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (i + 16 <= n) {
    const uint32_t w = v * 0x01010101u;
    *reinterpret_cast<uint4*>(p + i) = make_uint4(w, w, w, w);
  } else {
    for (int64_t k = i; k < n; ++k) p[k] = v;
  }
}

__global__ void __launch_bounds__(kBandThreads)
    k_curvatures(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
                 float* __restrict__ k12) {
  curvatures_body(phi, band, nb, d, h, k12, blockIdx.x, gridDim.x);
}

template <bool SVI>
__global__ void __launch_bounds__(kBandThreads)
    k_velocity(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
               LawF L, const float* __restrict__ field, const float* __restrict__ k12,
               float cut_lo, float cut_hi, Hi3 hi, int stats,
               float* __restrict__ V, double* __restrict__ partials) {
  velocity_body<SVI>(phi, band, nb, d, h, L, field, k12, cut_lo, cut_hi, hi, stats, V, partials, blockIdx.x,
                     gridDim.x);
}

template <bool SVI, int MODE>
__global__ void __launch_bounds__(kBandThreads)
    k_velocity_rk(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
                  LawF L, const float* __restrict__ field, const float* __restrict__ k12, float cut_lo, float cut_hi,
                  Hi3 hi, int stats, double* __restrict__ partials, RkFuse rk, const double* __restrict__ dtp) {
  if constexpr (MODE >= 2) {
    if (dtp) rk.dt = device_step(dtp);
  }
  velocity_body<SVI, MODE>(phi, band, nb, d, h, L, field, k12, cut_lo, cut_hi, hi, stats, nullptr, partials,
                           blockIdx.x, gridDim.x, rk);
}

// dtp, when set, replaces dt by the device step it points to
__global__ void __launch_bounds__(kBandThreads)
    k_rk_stage(int stage, const float* __restrict__ in, const float* __restrict__ u0, float* __restrict__ out,
               const int32_t* __restrict__ band, int64_t nb, Dims d, float inv_h, float dt,
               const float* __restrict__ V, float* __restrict__ L0, const double* __restrict__ dtp) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (dtp) dt = device_step(dtp);
  if (q < nb) rk_node(stage, in, u0, out, band, q, d, inv_h, dt, V, L0);
}

__global__ void k_stable_dt(double* __restrict__ results, int svi, int curv, double cfl, double h, double feedback,
                            double radius, double parabolic, int clamp, double max_dt) {
  if (threadIdx.x == 0)
    results[SVIMX_R_DT] = stable_dt_of(results, svi != 0, curv != 0, cfl, h, feedback, radius, parabolic,
                                       clamp != 0, max_dt);
}

// ---------------------------------------------------------------- operators
__global__ void k_op_godunov(const float* phi, const int32_t* idx, int64_t n, Dims d, float inv_h, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = godunov_inward(phi, idx[q], d, inv_h);
}
__global__ void k_op_fractions(const float* phi, const int32_t* idx, int64_t n, Dims d, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = cell_fraction(central(phi, idx[q], d));
}
__global__ void k_op_closest(const float* phi, const int32_t* idx, int64_t n, Dims d, float h, Hi3 hi, int extra,
                             float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int64_t node = idx[q];
  int i, j, k;
  unflat(d, node, i, j, k);
  float p[3];
  closest_point(central(phi, node, d), i, j, k, h, hi.v, p);
  newton_refine(phi, d, h, hi.v, p, extra);
  out[3 * q] = p[0];
  out[3 * q + 1] = p[1];
  out[3 * q + 2] = p[2];
}
__global__ void k_op_tricubic(const float* f, Dims d, const float* p, int64_t n, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float pp[3] = {p[3 * q], p[3 * q + 1], p[3 * q + 2]};
  out[q] = tricubic(f, d, pp);
}
__global__ void k_op_tricubic_grad(const float* f, Dims d, const float* p, int64_t n, float* out, float* grad) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float pp[3] = {p[3 * q], p[3 * q + 1], p[3 * q + 2]};
  float g[3];
  out[q] = tricubic_grad(f, d, pp, g);
  grad[3 * q] = g[0];
  grad[3 * q + 1] = g[1];
  grad[3 * q + 2] = g[2];
}
__global__ void k_op_trilinear(const float* f, Dims d, const float* p, int64_t n, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float pp[3] = {p[3 * q], p[3 * q + 1], p[3 * q + 2]};
  out[q] = trilinear(f, d, pp);
}
__global__ void k_op_curv(const float* phi, const int32_t* idx, int64_t n, Dims d, float h, float* k1, float* k2) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) principal_curvatures(phi, idx[q], d, h, k1[q], k2[q]);
}
__global__ void k_op_iface(const float* phi, const int32_t* idx, int64_t n, Dims d, uint8_t* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int64_t node = idx[q];
  float c = phi[node];
  const int64_t st[3] = {d.sx(), d.sy(), 1};
  bool hit = false;
  for (int a = 0; a < 3; ++a) hit = hit || (phi[node + st[a]] * c <= 0.0f) || (phi[node - st[a]] * c <= 0.0f);
  out[q] = hit ? 1 : 0;
}
__global__ void k_op_reinit_geo(const float* phi, const int32_t* idx, int64_t n, Dims d, float h, float cfl,
                                float* geo) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  ReinitGeo g = reinit_geometry<double>(phi, idx[q], d, h, cfl);
  geo[8 * q] = g.dtau;
  for (int s = 0; s < 6; ++s) geo[8 * q + 1 + s] = g.x[s];
  geo[8 * q + 7] = (float)g.flags;
}
__global__ void k_op_ptd(const float* p, const float* t, int64_t n, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float pp[3] = {p[3 * q], p[3 * q + 1], p[3 * q + 2]};
  const float* tt = t + 9 * q;
  out[q] = point_triangle_d2(pp, tt, tt + 3, tt + 6);
}

// Speed law at given values: SVI laws read N from q1 (V, dV/dN); curvature
// laws read (k1, k2) from q1, q2 (V, dV/dk1 + dV/dk2).  The same device
// functions as the velocity pass.
__global__ void k_op_law(LawF L, const float* q1, const float* q2, int64_t n, float* V, float* D) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float v, dv;
  if (L.kind == 0) {
    v = svi_speed(L, q1[q]);
    dv = svi_dspeed(L, q1[q]);
  } else {
    curv_speed(L, q1[q], q2[q], v, dv);
  }
  V[q] = v;
  if (D) D[q] = dv;
}

// Per-point output fields of a surface (svierosion/api.py:171-194): SVI laws
// interpolate N_r tricubically (q1 = N); curvature laws interpolate the dense
// (k1, k2) pairs tricubically (q1 = H = (k1 + k2)/2, q2 = K = k1 k2); both
// then evaluate the law's speed there.  Points are clamped to [1, n - 3.001]
// per axis so every stencil stays on the grid.
__global__ void k_annotate_points(LawF L, const float* __restrict__ field, const float2* __restrict__ k12, Dims d,
                                  Hi3 hi, const float* __restrict__ p, int64_t n, float* q1, float* q2,
                                  float* speed) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  float pp[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) pp[a] = fminf(fmaxf(p[3 * q + a], 1.0f), hi.v[a]);
  if (L.kind == 0) {
    const float N = tricubic(field, d, pp);
    q1[q] = N;
    if (q2) q2[q] = 0.0f;
    speed[q] = svi_speed(L, N);
  } else {
    const float2 k = tricubic2(k12, d, pp);
    float v, dv;
    curv_speed(L, k.x, k.y, v, dv);
    q1[q] = 0.5f * (k.x + k.y);
    if (q2) q2[q] = k.x * k.y;
    speed[q] = v;
  }
}

// The folds of the velocity statistics (partial slot s -> result
// SVIMX_R_VMAX + s): max, min, max, max, sum, sum, sum, max, then the same over
// every evaluated node.
ReduceSpec velocity_stats_spec() {
  ReduceSpec spec;
  const int8_t ops[15] = {1, 2, 1, 1, 0, 0, 0, 1, 2, 1, 1, 0, 0, 0, 1};
  for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
  for (int s = 0; s < 15; ++s) spec.op[s] = ops[s], spec.dst[s] = (int8_t)(SVIMX_R_VMAX + s);
  return spec;
}

// Velocity pass + (with stats) the finalisation of its 15 statistic slots into
// results[SVIMX_R_VMAX ..]: one kernel over the band list, then one fold of the
// block partials.  With statistics the kernel runs even for an empty band (one
// block of identities), so the fold never reads partials of an earlier call.
int velocity_launch(hipStream_t st, const float* phi, const int32_t* band, int64_t nb, Dims d, double h,
                    const svimx_law_t* law, const float* field, const float* k12, double cut_lo, double cut_hi,
                    int stats, float* V, double* partials, double* results) {
  SVIMX_REQUIRE(law && law->kind >= 0 && law->kind <= 2, SVIMX_EBADARG, "bad law");
  SVIMX_REQUIRE(law->kind != 0 || field, SVIMX_EBADARG, "SVI law needs the N_r field");
  SVIMX_REQUIRE(law->kind == 0 || k12, SVIMX_EBADARG, "curvature law needs the (k1, k2) field");
  SVIMX_REQUIRE(nb <= 0 || (phi && band && V), SVIMX_EBADARG, "velocity: NULL field, band or output");
  const int nblk = reduce_grid(nb);
  if (nb > 0 || stats) {
    if (law->kind == 0)
      hipLaunchKernelGGL(k_velocity<true>, dim3(nblk), dim3(kBandThreads), 0, st, phi, band, nb, d, (float)h,
                         law_to_f(*law), field, k12, (float)(cut_lo * h), (float)(cut_hi * h), hi_of(d), stats,
                         V, partials);
    else
      hipLaunchKernelGGL(k_velocity<false>, dim3(nblk), dim3(kBandThreads), 0, st, phi, band, nb, d, (float)h,
                         law_to_f(*law), field, k12, (float)(cut_lo * h), (float)(cut_hi * h), hi_of(d), stats,
                         V, partials);
    SVIMX_CHECK_LAUNCH();
  }
  if (!stats) return SVIMX_OK;
  const ReduceSpec spec = velocity_stats_spec();
  return launch_finalize(st, partials, nblk, results, spec);
}

// Block partials of svimx_velocity_static: a fixed device array, so that entry
// point allocates nothing.  Its size covers the largest reducing grid
// (kMaxReduceBlocks blocks x kPartialSlots doubles = 256 KiB per GPU, held
// until the process ends).  Every call on one GPU shares it: calls that
// overlap on one GPU (different streams or threads) race.  Each call writes
// every partial it folds, so nothing needs clearing between calls.
constexpr int64_t kVelocityScratchDoubles = (int64_t)kMaxReduceBlocks * kPartialSlots;
static __device__ __align__(256) double scratch[kVelocityScratchDoubles];

constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// Device address of `scratch` on the GPU of `stream` (the current device),
// resolved once per device with that device current; host code uses only this
// address, never the array's name.
int velocity_scratch(void* stream, double** out) {
  int dev = -1;
  SVIMX_CHECK_HIP(hipGetDevice(&dev));
  hipDevice_t sdev = dev;
  if (stream) SVIMX_CHECK_HIP(hipStreamGetDevice(as_stream(stream), &sdev));
  SVIMX_REQUIRE(sdev == dev, SVIMX_EBADARG,
                "svimx_velocity_static: the stream belongs to device %d but device %d is current", (int)sdev, dev);
  SVIMX_REQUIRE(dev >= 0 && dev < kMaxDevices, SVIMX_EBADARG, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SVIMX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SVIMX_REQUIRE(p, SVIMX_EDEVICE, "hipGetSymbolAddress returned NULL for the velocity scratch");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SVIMX_OK;
}

}  // namespace svimx

using namespace svimx;

#define GRID1(n) dim3(blocks_for((n), 256)), dim3(256), 0, as_stream(stream)

extern "C" {

int svimx_fractions(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                    double h, float* occupancy, void* nr_plan, double cut_hi, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, d.size(), w)) return SVIMX_EBADARG;
  SVIMX_REQUIRE(nb >= 0 && nb <= d.size(), SVIMX_EBADARG, "bad band size");
  hipStream_t st = as_stream(stream);
  NrPlan* plan = static_cast<NrPlan*>(nr_plan);
  uint8_t* marks = nullptr;
  int nty = 0, ntz = 0, nzr = 0, run_mode = 0;
  if (plan && plan->fft) plan = nullptr;  // the FFT needs no marks
  if (plan) {
    SVIMX_REQUIRE(plan->d.nx == nx && plan->d.ny == ny && plan->d.nz == nz, SVIMX_EBADARG,
                  "N_r plan grid does not match");
    SVIMX_REQUIRE((reinterpret_cast<uintptr_t>(plan->marks) & 15) == 0, SVIMX_EBADARG, "misaligned N_r marks");
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for((plan->nmarks + 15) / 16, 256)), dim3(256), 0, st, plan->marks,
                       plan->nmarks, (uint8_t)0);
    SVIMX_CHECK_LAUNCH();
    marks = plan->marks;
    nty = plan->nty;
    ntz = plan->ntz;
    nzr = plan->nzr;
    run_mode = plan->run_mode;
  }
  int nblk = reduce_grid(nb);
  if (nb > 0) {
    hipLaunchKernelGGL(k_fractions, dim3(nblk), dim3(kBandThreads), 0, st, phi, band, nb, d, (float)h,
                       occupancy, marks, run_mode, nty, ntz, nzr, (float)(cut_hi * h), hi_of(d), w.partials);
    SVIMX_CHECK_LAUNCH();
  }
  ReduceSpec spec;
  for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
  spec.op[0] = 0; spec.dst[0] = SVIMX_R_VOLUME_SUM;
  spec.op[1] = 1; spec.dst[1] = SVIMX_R_HAS_NEG;
  spec.op[2] = 1; spec.dst[2] = SVIMX_R_HAS_POS;
  return launch_finalize(st, w.partials, nblk, w.results, spec);
  SVIMX_TRY_END
}

int svimx_curvatures(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                     double h, float* k12_dense) {
  SVIMX_TRY_BEGIN
  if (nb <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_curvatures, GRID1(nb), phi, band, nb, Dims{nx, ny, nz}, (float)h, k12_dense);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_velocity(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                   double h, const svimx_law_t* law, const float* field, const float* k12, double cut_lo,
                   double cut_hi, int stats, float* V, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, Dims{nx, ny, nz}.size(), w)) return SVIMX_EBADARG;
  return velocity_launch(as_stream(stream), phi, band, nb, Dims{nx, ny, nz}, h, law, field, k12, cut_lo, cut_hi,
                         stats, V, w.partials, w.results);
  SVIMX_TRY_END
}

static int velocity_rk(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                       const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, const svimx_law_t* law,
                       const float* field, const float* k12, double cut_lo, double cut_hi, double dt, float* L0,
                       const float* cp, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, d.size(), w)) return SVIMX_EBADARG;
  SVIMX_REQUIRE(stage >= 0 && stage <= 2, SVIMX_EBADARG, "svimx_velocity_rk: stage must be 0, 1 or 2");
  SVIMX_REQUIRE(law && law->kind >= 0 && law->kind <= 2, SVIMX_EBADARG, "bad law");
  SVIMX_REQUIRE(law->kind != 0 || field, SVIMX_EBADARG, "SVI law needs the N_r field");
  SVIMX_REQUIRE(law->kind == 0 || k12, SVIMX_EBADARG, "curvature law needs the (k1, k2) field");
  SVIMX_REQUIRE(stage != 0 || L0, SVIMX_EBADARG, "stage 0 stores L0");
  SVIMX_REQUIRE(stage == 0 || (u0 && phi_out && phi_out != phi_in && phi_out != u0), SVIMX_EBADARG,
                "stages 1 and 2 need u0 and an output buffer distinct from the inputs");
  hipStream_t st = as_stream(stream);
  const int stats = stage == 0 ? 1 : 0;
  const int nblk = reduce_grid(nb);
  if (nb > 0 || stats) {
    RkFuse rk;
    rk.u0 = u0;
    rk.out = phi_out;
    rk.L0 = L0;
    rk.inv_h = (float)(1.0 / h);
    rk.dt = (float)dt;
    rk.cp = cp;
    const LawF lf = law_to_f(*law);
    const float clo = (float)(cut_lo * h), chi = (float)(cut_hi * h);
    // a negative dt selects the step svimx_stable_dt left in the workspace
    const double* dtp = dt < 0 ? w.results + SVIMX_R_DT : nullptr;
#define SVIMX_VRK(SVI_, MODE_)                                                                                   \
  hipLaunchKernelGGL((k_velocity_rk<SVI_, MODE_>), dim3(nblk), dim3(kBandThreads), 0, st, phi_in, band, nb, d,   \
                     (float)h, lf, field, k12, clo, chi, hi_of(d), stats, w.partials, rk, dtp)
    if (law->kind == 0) {
      if (stage == 0) SVIMX_VRK(true, 1); else if (stage == 1) SVIMX_VRK(true, 2); else SVIMX_VRK(true, 3);
    } else {
      if (stage == 0) SVIMX_VRK(false, 1); else if (stage == 1) SVIMX_VRK(false, 2); else SVIMX_VRK(false, 3);
    }
#undef SVIMX_VRK
    SVIMX_CHECK_LAUNCH();
  }
  if (!stats) return SVIMX_OK;
  const ReduceSpec spec = velocity_stats_spec();
  return launch_finalize(st, w.partials, nblk, w.results, spec);
  SVIMX_TRY_END
}

int svimx_velocity_rk(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                      const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, const svimx_law_t* law,
                      const float* field, const float* k12, double cut_lo, double cut_hi, double dt, float* L0,
                      void* ws, int64_t ws_bytes) {
  return velocity_rk(stream, stage, phi_in, u0, phi_out, band, nb, nx, ny, nz, h, law, field, k12, cut_lo, cut_hi,
                     dt, L0, nullptr, ws, ws_bytes);
}

int svimx_velocity_rk_points(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                             const int32_t* band, int64_t nb, int nx, int ny, int nz, double h,
                             const svimx_law_t* law, const float* field, const float* k12, double cut_lo,
                             double cut_hi, double dt, float* L0, const float* points, void* ws, int64_t ws_bytes) {
  SVIMX_REQUIRE(points || nb <= 0, SVIMX_EBADARG, "svimx_velocity_rk_points: null points");
  return velocity_rk(stream, stage, phi_in, u0, phi_out, band, nb, nx, ny, nz, h, law, field, k12, cut_lo, cut_hi,
                     dt, L0, points, ws, ws_bytes);
}

int svimx_velocity_static(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                          double h, const svimx_law_t* law, const float* field, const float* k12, double cut_lo,
                          double cut_hi, int stats, float* V, double* results) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(!stats || results, SVIMX_EBADARG, "svimx_velocity_static: statistics need the results array");
  SVIMX_REQUIRE(nb >= 0 && nb <= INT32_MAX, SVIMX_EBADARG,
                "svimx_velocity_static: %lld band nodes; the band list holds at most 2^31 - 1", (long long)nb);
  // the partials of reduce_grid(nb) <= kMaxReduceBlocks blocks always fit the static array
  double* partials = nullptr;
  if (stats) {
    const int rc = velocity_scratch(stream, &partials);
    if (rc != SVIMX_OK) return rc;
  }
  return velocity_launch(as_stream(stream), phi, band, nb, Dims{nx, ny, nz}, h, law, field, k12, cut_lo, cut_hi,
                         stats, V, partials, results);
  SVIMX_TRY_END
}

static int rk_stage(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                    const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, double dt,
                    const double* dt_dev, const float* V, float* L0) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(stage >= 0 && stage <= 3, SVIMX_EBADARG, "stage must be 0..3");
  SVIMX_REQUIRE(phi_in != phi_out || stage == 3, SVIMX_EBADARG, "stage output may not alias its input");
  SVIMX_REQUIRE(stage != 3 || L0, SVIMX_EBADARG, "stage 3 replays L0");
  if (nb <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_rk_stage, GRID1(nb), stage, phi_in, u0, phi_out, band, nb, Dims{nx, ny, nz},
                     (float)(1.0 / h), (float)dt, V, L0, dt_dev);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_rk_stage(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                   const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, double dt, const float* V,
                   float* L0) {
  return rk_stage(stream, stage, phi_in, u0, phi_out, band, nb, nx, ny, nz, h, dt, nullptr, V, L0);
}

int svimx_rk_stage_dt(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                      const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, const double* dt_dev,
                      const float* V, float* L0) {
  SVIMX_REQUIRE(dt_dev, SVIMX_EBADARG, "svimx_rk_stage_dt: null dt pointer");
  return rk_stage(stream, stage, phi_in, u0, phi_out, band, nb, nx, ny, nz, h, 0.0, dt_dev, V, L0);
}

int svimx_stable_dt(void* stream, const svimx_law_t* law, double h, double cfl, double feedback_cfl,
                    double parabolic_cfl, double radius, int clamp, double max_dt, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(law && law->kind >= 0 && law->kind <= 2, SVIMX_EBADARG, "bad law");
  SVIMX_REQUIRE(ws && ws_bytes >= (int64_t)(sizeof(double) * SVIMX_R_COUNT), SVIMX_EBADARG,
                "svimx_stable_dt: the workspace must hold the result slots");
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_stable_dt, dim3(1), dim3(64), 0, st, static_cast<double*>(ws), law->kind == 0 ? 1 : 0,
                     law->kind != 0 ? 1 : 0, cfl, h, feedback_cfl, radius, parabolic_cfl, clamp ? 1 : 0, max_dt);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_op_godunov(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny, int nz,
                     double inv_h, float* out) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_godunov, GRID1(n), phi, idx, n, Dims{nx, ny, nz}, (float)inv_h, out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_cell_fractions(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny, int nz,
                            float* out) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_fractions, GRID1(n), phi, idx, n, Dims{nx, ny, nz}, out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_closest_points(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny, int nz,
                            double h, float* out_p) {
  if (n <= 0) return SVIMX_OK;
  Dims d{nx, ny, nz};
  hipLaunchKernelGGL(k_op_closest, GRID1(n), phi, idx, n, d, (float)h, hi_of(d), 0, out_p);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_closest_points_newton(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny,
                                int nz, double h, int newton_steps, float* out_p) {
  SVIMX_REQUIRE(newton_steps >= 1 && newton_steps <= 64, SVIMX_EBADARG, "newton_steps must be in [1, 64]");
  SVIMX_REQUIRE(phi && (idx || n <= 0) && (out_p || n <= 0), SVIMX_EBADARG, "null argument");
  if (n <= 0) return SVIMX_OK;
  Dims d{nx, ny, nz};
  hipLaunchKernelGGL(k_op_closest, GRID1(n), phi, idx, n, d, (float)h, hi_of(d), newton_steps - 1, out_p);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_tricubic(void* stream, const float* field, int nx, int ny, int nz, const float* p, int64_t n,
                      float* out) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_tricubic, GRID1(n), field, Dims{nx, ny, nz}, p, n, out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_tricubic_gradient(void* stream, const float* field, int nx, int ny, int nz, const float* p, int64_t n,
                               float* out, float* grad) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_tricubic_grad, GRID1(n), field, Dims{nx, ny, nz}, p, n, out, grad);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_trilinear(void* stream, const float* field, int nx, int ny, int nz, const float* p, int64_t n,
                       float* out) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_trilinear, GRID1(n), field, Dims{nx, ny, nz}, p, n, out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_principal_curvatures(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny,
                                  int nz, double h, float* k1, float* k2) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_curv, GRID1(n), phi, idx, n, Dims{nx, ny, nz}, (float)h, k1, k2);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_interface_mask(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny, int nz,
                            uint8_t* out) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_iface, GRID1(n), phi, idx, n, Dims{nx, ny, nz}, out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_reinit_geometry(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny,
                             int nz, double h, double cfl, float* geo) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_reinit_geo, GRID1(n), phi, idx, n, Dims{nx, ny, nz}, (float)h, (float)cfl, geo);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_point_triangle_distance(void* stream, const float* p, const float* tri, int64_t n, float* out) {
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_ptd, GRID1(n), p, tri, n, out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_op_law(void* stream, const svimx_law_t* law, const float* q1, const float* q2, int64_t n, float* V,
                 float* D) {
  SVIMX_REQUIRE(law && law->kind >= 0 && law->kind <= 2, SVIMX_EBADARG, "svimx_op_law: bad law");
  SVIMX_REQUIRE(n <= 0 || (q1 && V && (law->kind == 0 || q2)), SVIMX_EBADARG,
                "svimx_op_law: missing input or output (curvature laws need q1 = k1 and q2 = k2)");
  if (n <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_op_law, GRID1(n), law_to_f(*law), q1, q2, n, V, D);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}
int svimx_annotate_points(void* stream, const svimx_law_t* law, const float* field, const float* k12, int nx,
                          int ny, int nz, const float* p, int64_t n, float* q1, float* q2, float* speed) {
  SVIMX_REQUIRE(law && law->kind >= 0 && law->kind <= 2, SVIMX_EBADARG, "svimx_annotate_points: bad law");
  SVIMX_REQUIRE(law->kind != 0 || field, SVIMX_EBADARG, "svimx_annotate_points: SVI laws need the N_r field");
  SVIMX_REQUIRE(law->kind == 0 || k12, SVIMX_EBADARG, "svimx_annotate_points: curvature laws need (k1, k2)");
  SVIMX_REQUIRE(nx >= 5 && ny >= 5 && nz >= 5, SVIMX_EBADARG, "svimx_annotate_points: grid below 5 nodes");
  SVIMX_REQUIRE(n <= 0 || (p && q1 && speed), SVIMX_EBADARG, "svimx_annotate_points: NULL buffer");
  if (n <= 0) return SVIMX_OK;
  Dims d{nx, ny, nz};
  hipLaunchKernelGGL(k_annotate_points, GRID1(n), law_to_f(*law), field, reinterpret_cast<const float2*>(k12), d,
                     hi_of(d), p, n, q1, q2, speed);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

}  // extern "C"
