// Device bodies of the per-stage band kernels, shared by the single-job
// kernels (stage.hip) and the job-table kernels of the batch engine
// (engine.hip).  A body processes one "virtual block" lb of a grid of G
// blocks of kBandThreads threads: grid-stride over the band list with stride
// G * kBandThreads, block partials at partials[lb * kPartialSlots].  The
// single-job kernels call it with (blockIdx.x, gridDim.x); the batch engine
// with the job's own (lb, G), so a job's arithmetic and reduction order are
// the same alone and in a batch.
#pragma once

#include "common.h"
#include "nr_plan.h"

namespace svimx {

// ---------------------------------------------------------------- speed laws
struct LawF {
  int kind, k, s;
  bool has_eps;
  float t, eps, n_cap, alpha, delta, alpha_m1, alpha_m2;
};

inline LawF law_to_f(const svimx_law_t& L) {
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
    V = ipow(N, L.k) * rcp(L.t - N);
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
    const float id = rcp(L.t - N);
    float V = ipow(N, L.k) * id;
    return V * ((float)L.k * rcp(N) + id);
  }
  if (L.k == 0) return 0.0f;
  return (float)L.k * ipow(N, L.k - 1);
}

// x^e for x >= 0 as exp2(e log2 x) with the hardware log2 / exp2 (v_log_f32,
// v_exp_f32); relative error ~1e-6 for the exponents and ranges used here,
// x = 0 gives 0 (e > 0) or +inf (e < 0) as powf does.
__device__ __forceinline__ float pow_pos(float x, float e) {
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}

// x^alpha regularised for alpha < 1 and its derivative (svierosion/flows.py:139-158).
__device__ __forceinline__ float curv_power(const LawF& L, float x) {
  if (L.alpha == 1.0f) return x;
  if (L.alpha < 1.0f) return x * pow_pos(x + L.delta, L.alpha_m1);
  return pow_pos(x, L.alpha);
}
__device__ __forceinline__ float curv_dpower(const LawF& L, float x) {
  if (L.alpha == 1.0f) return 1.0f;
  if (L.alpha < 1.0f) return pow_pos(x + L.delta, L.alpha_m2) * (L.alpha * x + L.delta);
  return L.alpha * pow_pos(fmaxf(x, 0.0f), L.alpha_m1);
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
  float t = fminf(fmaxf((b - x) * rcp(b - a), 0.0f), 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

struct Hi3 {
  float v[3];
};
inline Hi3 hi_of(const Dims& d) {
  Hi3 h;
  h.v[0] = (float)(d.nx - 3.001);
  h.v[1] = (float)(d.ny - 3.001);
  h.v[2] = (float)(d.nz - 3.001);
  return h;
}

// ---------------------------------------------------------------- fractions
// Cell fractions into the occupancy layout, N_r stencil marks, and partials
// [0] sum of fractions, [1] any phi < 0, [2] any phi > 0.
__device__ __forceinline__ void fractions_body(const float* __restrict__ phi, const int32_t* __restrict__ band,
                                               int64_t nb, Dims d, float h, float* __restrict__ occ,
                                               uint8_t* __restrict__ marks, int run_mode, int nty, int ntz, int nzr,
                                               float cut, Hi3 hi, double* __restrict__ partials, int64_t lb,
                                               int64_t G) {
  __shared__ double red[4];
  const OccLayout OL = OccLayout::of(d.nx, d.ny, d.nz);
  double fsum = 0.0, neg = 0.0, pos = 0.0;
  for (int64_t q = lb * kBandThreads + threadIdx.x; q < nb; q += G * kBandThreads) {
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
        mark_stencil_runs(marks, d.ny, nzr, run_mode, bx, by, bz);
      else
        mark_stencil_tiles(marks, nty, ntz, bx, by, bz);
    }
  }
  double s = block_reduce<0>(fsum, red);
  double a = block_reduce<1>(neg, red);
  double b = block_reduce<1>(pos, red);
  if (threadIdx.x == 0) {
    double* o = partials + lb * kPartialSlots;
    o[0] = s;
    o[1] = a;
    o[2] = b;
  }
}

// ---------------------------------------------------------------- curvature scatter
__device__ __forceinline__ void curvatures_body(const float* __restrict__ phi, const int32_t* __restrict__ band,
                                                int64_t nb, Dims d, float h, float* __restrict__ k1d,
                                                float* __restrict__ k2d, int64_t lb, int64_t G) {
  for (int64_t q = lb * kBandThreads + threadIdx.x; q < nb; q += G * kBandThreads) {
    const int64_t n = band[q];
    float k1, k2;
    principal_curvatures(phi, n, d, h, k1, k2);
    k1d[n] = k1;
    k2d[n] = k2;
  }
}

// ---------------------------------------------------------------- velocity
// Partial slots: 0 vmax, 1 qmin, 2 qmax, 3 dvmax, 4 capped, 5 vsum, 6 nnear,
// 7 dmax, 8 qmin_all, 9 qmax_all, 10 dvmax_all, 11 capped_all, 12 vsum_all,
// 13 nall, 14 dmax_all.
__device__ __forceinline__ void velocity_body(const float* __restrict__ phi, const int32_t* __restrict__ band,
                                              int64_t nb, Dims d, float h, const LawF& L,
                                              const float* __restrict__ field, const float* __restrict__ k1d,
                                              const float* __restrict__ k2d, float cut_lo, float cut_hi, Hi3 hi,
                                              int stats, float* __restrict__ V, double* __restrict__ partials,
                                              int64_t lb, int64_t G) {
  __shared__ double red[4];
  float vmax = 0.0f;
  float qmin = INFINITY, qmax = -INFINITY, dvmax = -INFINITY, dmax = -INFINITY;
  float qmin_a = INFINITY, qmax_a = -INFINITY, dvmax_a = -INFINITY, dmax_a = -INFINITY;
  double capped = 0, vsum = 0, nnear = 0, capped_a = 0, vsum_a = 0, nall = 0;
  for (int64_t q = lb * kBandThreads + threadIdx.x; q < nb; q += G * kBandThreads) {
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
    double* o = partials + lb * kPartialSlots;
    for (int s = 0; s < 15; ++s) o[s] = r[s];
  }
}

// ---------------------------------------------------------------- RK stage
// Shu-Osher TVD-RK3 stage s of node q (svierosion/evolve.py:172-179); stage 3
// replays stage 0 from the saved L0 = V |grad phi|.
__device__ __forceinline__ void rk_node(int stage, const float* __restrict__ in, const float* __restrict__ u0,
                                        float* __restrict__ out, const int32_t* __restrict__ band, int64_t q, Dims d,
                                        float inv_h, float dt, const float* __restrict__ V, float* __restrict__ L0) {
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
    out[n] = u0[n] * kThird + (2.0f / 3.0f) * (in[n] + dt * L);
  }
}

}  // namespace svimx
