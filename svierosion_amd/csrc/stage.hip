// Per-stage band kernels: cell fractions (+ volume, interface flags, N_r tile
// marks), curvature scatter, velocity (+ dt statistics), TVD-RK3 stage update,
// and the node-list operator entry points that share their device code.
#include "common.h"
#include "nr_plan.h"

namespace svimx {

// ---------------------------------------------------------------- speed laws
struct LawF {
  int kind, k, s;
  bool has_eps;
  float t, eps, n_cap, alpha, delta, alpha_m1, alpha_m2;
};

static LawF law_to_f(const svimx_law_t& L) {
  LawF f;
  f.kind = L.kind;
  f.k = L.k;
  f.s = L.s;
  f.has_eps = L.eps != 0.0;
  f.t = (float)L.t;
  f.eps = (float)L.eps;
  f.n_cap = (float)L.n_cap;
  f.alpha = (float)L.alpha;
  f.delta = (float)L.delta;
  f.alpha_m1 = (float)(L.alpha - 1.0);
  f.alpha_m2 = (float)(L.alpha - 2.0);
  return f;
}

__device__ __forceinline__ float ipow(float x, int k) {
  float r = 1.0f;
  for (int q = 0; q < k; ++q) r *= x;
  return r;
}

// V = N^k / (t - N)^s (+ eps) with N clamped to [0, 1] and to n_cap when s = 1
// (svierosion/flows.py:89-96).
__device__ __forceinline__ float svi_speed(const LawF& L, float N) {
  N = fminf(fmaxf(N, 0.0f), 1.0f);
  float V;
  if (L.s) {
    N = fminf(N, L.n_cap);
    V = ipow(N, L.k) / (L.t - N);
  } else {
    V = L.k ? ipow(N, L.k) : 1.0f;
  }
  return L.has_eps ? V + L.eps : V;
}

// dV/dN (svierosion/flows.py:98-108).
__device__ __forceinline__ float svi_dspeed(const LawF& L, float N) {
  N = fminf(fmaxf(N, 1e-12f), 1.0f);
  if (L.s) {
    N = fminf(N, L.n_cap);
    float d = L.t - N;
    float V = ipow(N, L.k) / d;
    return V * ((float)L.k / N + 1.0f / d);
  }
  if (L.k == 0) return 0.0f;
  return (float)L.k * ipow(N, L.k - 1);
}

// x^alpha regularised for alpha < 1 and its derivative (svierosion/flows.py:139-158).
__device__ __forceinline__ float curv_power(const LawF& L, float x) {
  if (L.alpha == 1.0f) return x;
  if (L.alpha < 1.0f) return x * powf(x + L.delta, L.alpha_m1);
  return powf(x, L.alpha);
}
__device__ __forceinline__ float curv_dpower(const LawF& L, float x) {
  if (L.alpha == 1.0f) return 1.0f;
  if (L.alpha < 1.0f) return powf(x + L.delta, L.alpha_m2) * (L.alpha * x + L.delta);
  return L.alpha * powf(fmaxf(x, 0.0f), L.alpha_m1);
}

// Speed and diffusivity dV/dk1 + dV/dk2 of the curvature laws (svierosion/flows.py:170-205).
__device__ __forceinline__ void curv_speed(const LawF& L, float k1, float k2, float& V, float& D) {
  if (L.kind == 1) {
    float H = 0.5f * (k1 + k2);
    float Hp = fmaxf(H, 0.0f);
    V = curv_power(L, Hp) + L.eps;
    D = H > 0.0f ? curv_dpower(L, Hp) : 0.0f;
  } else {
    float a = fmaxf(k1, 0.0f), b = fmaxf(k2, 0.0f);
    float Kc = a * b;
    V = curv_power(L, Kc) + L.eps;
    float dF = curv_dpower(L, Kc);
    D = Kc > 0.0f ? dF * (a + b) : dF * 0.0f;
  }
}

// Smoothstep cutoff (svierosion/evolve.py:108-112), a/b are float32(c*h).
__device__ __forceinline__ float cutoff_factor(float x, float a, float b) {
  float t = fminf(fmaxf((b - x) / (b - a), 0.0f), 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

struct Hi3 {
  float v[3];
};
static Hi3 hi_of(const Dims& d) {
  Hi3 h;
  h.v[0] = (float)(d.nx - 3.001);
  h.v[1] = (float)(d.ny - 3.001);
  h.v[2] = (float)(d.nz - 3.001);
  return h;
}

// ---------------------------------------------------------------- fractions
__global__ void __launch_bounds__(kBandThreads)
    k_fractions(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
                float* __restrict__ occ, uint8_t* __restrict__ marks, int run_mode, int nty, int ntz, int nzr,
                float cut, Hi3 hi, double* __restrict__ partials) {
  __shared__ double red[4];
  const OccLayout OL = OccLayout::of(d.nx, d.ny, d.nz);
  double fsum = 0.0, neg = 0.0, pos = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = band[q];
    int i, j, k;
    unflat(d, n, i, j, k);
    Grad g = central(phi, n, d);
    float fr = cell_fraction(g);
    occ[OL.at(i, j, k)] = fr;
    fsum += fr;
    neg = g.c < 0.0f ? 1.0 : neg;
    pos = g.c > 0.0f ? 1.0 : pos;
    if (marks && fabsf(g.c) < cut) {
      float p[3];
      closest_point(g, i, j, k, h, hi.v, p);
      const int bx = (int)floorf(p[0]) - 1, by = (int)floorf(p[1]) - 1, bz = (int)floorf(p[2]) - 1;
      if (run_mode)
        mark_stencil_runs(marks, d.ny, nzr, bx, by, bz);
      else
        mark_stencil_tiles(marks, nty, ntz, bx, by, bz);
    }
  }
  double s = block_reduce<0>(fsum, red);
  double a = block_reduce<1>(neg, red);
  double b = block_reduce<1>(pos, red);
  if (threadIdx.x == 0) {
    double* o = partials + blockIdx.x * kPartialSlots;
    o[0] = s;
    o[1] = a;
    o[2] = b;
  }
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- curvature scatter
__global__ void __launch_bounds__(kBandThreads)
    k_curvatures(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
                 float* __restrict__ k1d, float* __restrict__ k2d) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  const int64_t n = band[q];
  float k1, k2;
  principal_curvatures(phi, n, d, h, k1, k2);
  k1d[n] = k1;
  k2d[n] = k2;
}

// ---------------------------------------------------------------- velocity
// Partial slots: 0 vmax, 1 qmin, 2 qmax, 3 dvmax, 4 capped, 5 vsum, 6 nnear,
// 7 dmax, 8 qmin_all, 9 qmax_all, 10 dvmax_all, 11 capped_all, 12 vsum_all,
// 13 nall, 14 dmax_all.
__global__ void __launch_bounds__(kBandThreads)
    k_velocity(const float* __restrict__ phi, const int32_t* __restrict__ band, int64_t nb, Dims d, float h,
               LawF L, const float* __restrict__ field, const float* __restrict__ k1d,
               const float* __restrict__ k2d, float cut_lo, float cut_hi, Hi3 hi, int stats,
               float* __restrict__ V, double* __restrict__ partials) {
  __shared__ double red[4];
  float vmax = 0.0f;
  float qmin = INFINITY, qmax = -INFINITY, dvmax = -INFINITY, dmax = -INFINITY;
  float qmin_a = INFINITY, qmax_a = -INFINITY, dvmax_a = -INFINITY, dmax_a = -INFINITY;
  double capped = 0, vsum = 0, nnear = 0, capped_a = 0, vsum_a = 0, nall = 0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t n = band[q];
    Grad g = central(phi, n, d);
    const float x = fabsf(g.c);
    const bool near = x < h;
    float v = 0.0f;
    if (L.kind == 0) {
      if (x < cut_hi) {
        int i, j, k;
        unflat(d, n, i, j, k);
        float p[3];
        closest_point(g, i, j, k, h, hi.v, p);
        float N = tricubic(field, d, p);
        float vraw = svi_speed(L, N);
        v = vraw * cutoff_factor(x, cut_lo, cut_hi);
        if (stats) {
          float dv = svi_dspeed(L, N);
          double cp = N > L.n_cap ? 1.0 : 0.0;
          qmin_a = fminf(qmin_a, N); qmax_a = fmaxf(qmax_a, N); dvmax_a = fmaxf(dvmax_a, dv);
          capped_a += cp; vsum_a += vraw; nall += 1;
          if (near) {
            qmin = fminf(qmin, N); qmax = fmaxf(qmax, N); dvmax = fmaxf(dvmax, dv);
            capped += cp; vsum += vraw; nnear += 1;
          }
        }
      }
    } else {
      int i, j, k;
      unflat(d, n, i, j, k);
      float p[3];
      closest_point(g, i, j, k, h, hi.v, p);
      float k1 = trilinear(k1d, d, p), k2 = trilinear(k2d, d, p);
      float vraw, D;
      curv_speed(L, k1, k2, vraw, D);
      v = vraw * cutoff_factor(x, cut_lo, cut_hi);
      if (stats) {
        float H = 0.5f * (k1 + k2);
        qmin_a = fminf(qmin_a, H); qmax_a = fmaxf(qmax_a, H); dmax_a = fmaxf(dmax_a, D);
        vsum_a += vraw; nall += 1;
        if (near) {
          qmin = fminf(qmin, H); qmax = fmaxf(qmax, H); dmax = fmaxf(dmax, D);
          vsum += vraw; nnear += 1;
        }
      }
    }
    V[q] = v;
    vmax = fmaxf(vmax, v);
  }
  if (!stats) return;
  double r[15];
  r[0] = block_reduce<1>(vmax, red);
  r[1] = block_reduce<2>(qmin, red);
  r[2] = block_reduce<1>(qmax, red);
  r[3] = block_reduce<1>(dvmax, red);
  r[4] = block_reduce<0>(capped, red);
  r[5] = block_reduce<0>(vsum, red);
  r[6] = block_reduce<0>(nnear, red);
  r[7] = block_reduce<1>(dmax, red);
  r[8] = block_reduce<2>(qmin_a, red);
  r[9] = block_reduce<1>(qmax_a, red);
  r[10] = block_reduce<1>(dvmax_a, red);
  r[11] = block_reduce<0>(capped_a, red);
  r[12] = block_reduce<0>(vsum_a, red);
  r[13] = block_reduce<0>(nall, red);
  r[14] = block_reduce<1>(dmax_a, red);
  if (threadIdx.x == 0) {
    double* o = partials + blockIdx.x * kPartialSlots;
    for (int s = 0; s < 15; ++s) o[s] = r[s];
  }
}

// ---------------------------------------------------------------- RK stage
__global__ void __launch_bounds__(kBandThreads)
    k_rk_stage(int stage, const float* __restrict__ in, const float* __restrict__ u0, float* __restrict__ out,
               const int32_t* __restrict__ band, int64_t nb, Dims d, float inv_h, float dt,
               const float* __restrict__ V, float* __restrict__ L0) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  const int64_t n = band[q];
  if (stage == 3) {
    out[n] = u0[n] + dt * L0[q];
    return;
  }
  float L = V[q] * godunov_inward(in, n, d, inv_h);
  if (stage == 0) {
    out[n] = u0[n] + dt * L;
    if (L0) L0[q] = L;
  } else if (stage == 1) {
    out[n] = 0.75f * u0[n] + 0.25f * (in[n] + dt * L);
  } else {
    out[n] = u0[n] / 3.0f + (2.0f / 3.0f) * (in[n] + dt * L);
  }
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
__global__ void k_op_closest(const float* phi, const int32_t* idx, int64_t n, Dims d, float h, Hi3 hi, float* out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int64_t node = idx[q];
  int i, j, k;
  unflat(d, node, i, j, k);
  float p[3];
  closest_point(central(phi, node, d), i, j, k, h, hi.v, p);
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
  ReinitGeo g = reinit_geometry(phi, idx[q], d, h, cfl);
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
  if (plan) {
    SVIMX_REQUIRE(plan->d.nx == nx && plan->d.ny == ny && plan->d.nz == nz, SVIMX_EBADARG,
                  "N_r plan grid does not match");
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks_for(plan->nmarks, 256)), dim3(256), 0, st, plan->marks,
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
                     double h, float* k1_dense, float* k2_dense) {
  SVIMX_TRY_BEGIN
  if (nb <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_curvatures, GRID1(nb), phi, band, nb, Dims{nx, ny, nz}, (float)h, k1_dense, k2_dense);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_velocity(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                   double h, const svimx_law_t* law, const float* field, const float* k1, const float* k2,
                   double cut_lo, double cut_hi, int stats, float* V, void* ws, int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, d.size(), w)) return SVIMX_EBADARG;
  SVIMX_REQUIRE(law && law->kind >= 0 && law->kind <= 2, SVIMX_EBADARG, "bad law");
  SVIMX_REQUIRE(law->kind != 0 || field, SVIMX_EBADARG, "SVI law needs the N_r field");
  SVIMX_REQUIRE(law->kind == 0 || (k1 && k2), SVIMX_EBADARG, "curvature law needs k1/k2 fields");
  hipStream_t st = as_stream(stream);
  int nblk = reduce_grid(nb);
  if (nb > 0) {
    hipLaunchKernelGGL(k_velocity, dim3(nblk), dim3(kBandThreads), 0, st, phi, band, nb, d, (float)h,
                       law_to_f(*law), field, k1, k2, (float)(cut_lo * h), (float)(cut_hi * h), hi_of(d), stats,
                       V, w.partials);
    SVIMX_CHECK_LAUNCH();
  }
  if (!stats) return SVIMX_OK;
  ReduceSpec spec;
  const int8_t ops[15] = {1, 2, 1, 1, 0, 0, 0, 1, 2, 1, 1, 0, 0, 0, 1};
  for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
  for (int s = 0; s < 15; ++s) spec.op[s] = ops[s], spec.dst[s] = (int8_t)(SVIMX_R_VMAX + s);
  return launch_finalize(st, w.partials, nblk, w.results, spec);
  SVIMX_TRY_END
}

int svimx_rk_stage(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                   const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, double dt, const float* V,
                   float* L0) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(stage >= 0 && stage <= 3, SVIMX_EBADARG, "stage must be 0..3");
  SVIMX_REQUIRE(phi_in != phi_out || stage == 3, SVIMX_EBADARG, "stage output may not alias its input");
  SVIMX_REQUIRE(stage != 3 || L0, SVIMX_EBADARG, "stage 3 replays L0");
  if (nb <= 0) return SVIMX_OK;
  hipLaunchKernelGGL(k_rk_stage, GRID1(nb), stage, phi_in, u0, phi_out, band, nb, Dims{nx, ny, nz},
                     (float)(1.0 / h), (float)dt, V, L0);
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
  hipLaunchKernelGGL(k_op_closest, GRID1(n), phi, idx, n, d, (float)h, hi_of(d), out_p);
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

}  // extern "C"
