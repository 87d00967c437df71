// Device bodies of the per-stage band kernels, shared by the single-job
// kernels (stage.hip) and the job-table kernels of the batch engine
// (engine.hip).  A body processes one "virtual block" lb of a grid of G
// blocks of kBandThreads threads: grid-stride over the band list with stride
// G * kBandThreads, block partials at partials[slot * G + lb].  The
// single-job kernels call it with (blockIdx.x, gridDim.x); the batch engine
// with the job's own (lb, G), so a job's arithmetic and reduction order are
// the same alone and in a batch.  Only velocity_body gives each XCD a
// contiguous eighth of the list (its tricubic/trilinear gathers share lines
// between neighbouring nodes); for the other bodies that split measured
// slower or within noise.
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
  double fsum = 0.0;
  bool neg = false, pos = false;
  for (int64_t q = lb * kBandThreads + threadIdx.x; q < nb; q += G * kBandThreads) {
    const int64_t n = band[q];
    int i, j, k;
    unflat(d, n, i, j, k);
    Grad g = central(phi, n, d);
    float fr = cell_fraction(g);
    occ[OL.at(i, j, k)] = fr;
    fsum += fr;
    neg = neg || g.c < 0.0f;
    pos = pos || g.c > 0.0f;
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
  // the sign flags are block-wide ORs: ballots, folded into the one LDS exchange of the sum
  __shared__ int flags[4];
  const int wave = threadIdx.x >> 6;
  const unsigned long long bn = __ballot(neg), bp = __ballot(pos);
  if ((threadIdx.x & 63) == 0) flags[wave] = (bn ? 1 : 0) | (bp ? 2 : 0);
  double s = block_reduce<0>(fsum, red);  // its barriers publish flags[]
  if (threadIdx.x == 0) {
    int f = 0;
    for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) f |= flags[w];
    double* o = partials + lb;
    o[0] = s;
    o[G] = (f & 1) ? 1.0 : 0.0;
    o[2 * G] = (f & 2) ? 1.0 : 0.0;
  }
  __syncthreads();  // flags[] is rewritten by the block's next segment
}

// ---------------------------------------------------------------- curvature scatter
__device__ __forceinline__ void curvatures_body(const float* __restrict__ phi, const int32_t* __restrict__ band,
                                                int64_t nb, Dims d, float h, float* __restrict__ k12,
                                                int64_t lb, int64_t G) {
  for (int64_t q = lb * kBandThreads + threadIdx.x; q < nb; q += G * kBandThreads) {
    const int64_t n = band[q];
    float k1, k2;
    principal_curvatures(phi, n, d, h, k1, k2);
    reinterpret_cast<float2*>(k12)[n] = make_float2(k1, k2);
  }
}

// ---------------------------------------------------------------- velocity
// Partial slots: 0 vmax, 1 qmin, 2 qmax, 3 dvmax, 4 capped, 5 vsum, 6 nnear,
// 7 dmax, 8 qmin_all, 9 qmax_all, 10 dvmax_all, 11 capped_all, 12 vsum_all,
// 13 nall, 14 dmax_all.
// SVI selects the law family at compile time (true: SVI power law, false:
// curvature laws), so each variant gets the registers it needs: the SVI
// tricubic holds 64 values in flight, the curvature trilinear 16.
//
// MODE fuses the TVD-RK3 stage that consumes the velocity (svierosion/evolve.py:
// 157-160, :172-179), so V never goes through memory and the gathers of the
// velocity overlap the WENO5 arithmetic of other waves: 0 stores V[q] (the
// plain pass); 1 stores L0[q] = V G (stage 0 before dt is known; the update
// u0 + dt L0 is the stage-3 replay); 2 and 3 write stages 1 and 2 of the
// update into rk.out.  V G, u0 and phi are the same floats as in the separate
// passes (rk_node), combined by the same operations: the fields are
// bit-identical to the unfused kernels.
//
// rk.cp, when set, holds each band entry's closest point (3 floats per entry,
// svimx_closest_points_newton with newton_steps > 1); else the body projects
// the node itself (one step).
struct RkFuse {
  const float* u0 = nullptr;
  float* out = nullptr;
  float* L0 = nullptr;
  float inv_h = 0.0f, dt = 0.0f;
  const float* cp = nullptr;
};

// The stable step of svierosion/evolve.py:162-170 from the finalised velocity
// statistics r[SVIMX_R_VMAX ..], with the max_time clamp of :199-200 when
// `clamp` is set: FP64 in the operation order of the host formula, without
// contraction, so the host's and the device's dt agree bit for bit.
__device__ __forceinline__ double stable_dt_of(const double* __restrict__ r, bool svi, bool curv, double cfl,
                                               double h, double feedback, double radius, double parabolic,
                                               bool clamp, double max_dt) {
#pragma clang fp contract(off)
  const double vmax = r[SVIMX_R_VMAX];
  const bool near = r[SVIMX_R_NNEAR] > 0;
  double dt = vmax > 0 ? (cfl * h) / vmax : INFINITY;
  if (svi && (r[SVIMX_R_NALL] > 0 || near)) {
    const double dv = near ? r[SVIMX_R_DVMAX] : r[SVIMX_R_DVMAX_ALL];
    if (dv > 0) {
      const double x = (feedback * ((4.0 * radius) / 3.0)) / dv;
      dt = x < dt ? x : dt;
    }
  }
  if (curv) {
    const double dm = near ? r[SVIMX_R_DMAX] : r[SVIMX_R_DMAX_ALL];
    if (dm > 0) {
      const double x = ((parabolic * h) * h) / dm;
      dt = x < dt ? x : dt;
    }
  }
  if (clamp) dt = max_dt < dt ? max_dt : dt;
  return dt;
}

// The float step a stage applies when it reads dt from device memory.  A step
// that is not finite and positive is discarded by the caller ("stalled"), so
// its stages apply 0 and keep the fields finite.
__device__ __forceinline__ float device_step(const double* __restrict__ dtp) {
  const double v = *dtp;
  return (v > 0.0 && v < INFINITY) ? (float)v : 0.0f;
}

template <bool SVI, int MODE = 0>
__device__ __forceinline__ void velocity_body(const float* __restrict__ phi, const int32_t* __restrict__ band,
                                              int64_t nb, Dims d, float h, const LawF& L,
                                              const float* __restrict__ field, const float* __restrict__ k12,
                                              float cut_lo, float cut_hi, Hi3 hi,
                                              int stats, float* __restrict__ V, double* __restrict__ partials,
                                              int64_t lb, int64_t G, const RkFuse& rk = RkFuse{}) {
  float vmax = 0.0f;
  float qmin = INFINITY, qmax = -INFINITY, dvmax = -INFINITY, dmax = -INFINITY;
  float qmin_a = INFINITY, qmax_a = -INFINITY, dvmax_a = -INFINITY, dmax_a = -INFINITY;
  double vsum = 0, vsum_a = 0;
  uint32_t capped = 0, nnear = 0, capped_a = 0, nall = 0;
  // Software pipeline over the grid-stride loop: the band index two nodes
  // ahead and the gradient stencil of the next node are loaded while this
  // node's closest point and interpolation run (past the end, the loads
  // repeat an in-bounds node and are not used).
  // With G a multiple of the XCD count, the blocks of XCD x (lb mod 8 = x)
  // stride over the x-th contiguous eighth of the band list, so the N_r lines
  // that neighbouring nodes' tricubic stencils share stay in one XCD's L2.
  int64_t q = lb * kBandThreads + threadIdx.x, stride = G * kBandThreads, qend = nb;
  if (G % kXcds == 0) {
    const int64_t per = (nb + kXcds - 1) / kXcds, x = lb % kXcds;
    q = x * per + (lb / kXcds) * kBandThreads + threadIdx.x;
    stride = (G / kXcds) * kBandThreads;
    qend = min(nb, x * per + per);
  }
  int64_t n = 0, n1 = 0;
  Grad gnext{};
  if (q < qend) {  // band nodes keep their stencils inside the grid
    n = band[q];
    n1 = q + stride < qend ? band[q + stride] : n;
    gnext = central(phi, n, d);
  }
  for (; q < qend; q += stride) {
    const Grad g = gnext;
    gnext = central(phi, n1, d);
    const int64_t n2 = q + 2 * stride < qend ? band[q + 2 * stride] : n1;
    const float x = fabsf(g.c);
    const bool near = x < h;
    float v = 0.0f;
    if constexpr (SVI) {
      if (x < cut_hi) {
        float p[3];
        if (rk.cp) {
          p[0] = rk.cp[3 * q]; p[1] = rk.cp[3 * q + 1]; p[2] = rk.cp[3 * q + 2];
        } else {
          int i, j, k;
          unflat(d, n, i, j, k);
          closest_point(g, i, j, k, h, hi.v, p);
        }
        float N = tricubic(field, d, p);
        float vraw = svi_speed(L, N);
        v = vraw * cutoff_factor(x, cut_lo, cut_hi);
        if (stats) {
          float dv = svi_dspeed(L, N);
          const uint32_t cp = N > L.n_cap ? 1u : 0u;
          qmin_a = fminf(qmin_a, N); qmax_a = fmaxf(qmax_a, N); dvmax_a = fmaxf(dvmax_a, dv);
          capped_a += cp; vsum_a += vraw; nall += 1;
          if (near) {
            qmin = fminf(qmin, N); qmax = fmaxf(qmax, N); dvmax = fmaxf(dvmax, dv);
            capped += cp; vsum += vraw; nnear += 1;
          }
        }
      }
    } else {
      float p[3];
      if (rk.cp) {
        p[0] = rk.cp[3 * q]; p[1] = rk.cp[3 * q + 1]; p[2] = rk.cp[3 * q + 2];
      } else {
        int i, j, k;
        unflat(d, n, i, j, k);
        closest_point(g, i, j, k, h, hi.v, p);
      }
      const float2 kk = trilinear2(reinterpret_cast<const float2*>(k12), d, p);
      const float k1 = kk.x, k2 = kk.y;
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
    if constexpr (MODE == 0) {
      V[q] = v;
    } else {
      const float Lq = v * godunov_inward(phi, n, d, rk.inv_h);
      if constexpr (MODE == 1) {
        rk.L0[q] = Lq;
      } else if constexpr (MODE == 2) {
        rk.out[n] = 0.75f * rk.u0[n] + 0.25f * (phi[n] + rk.dt * Lq);
      } else {
        rk.out[n] = rk.u0[n] * kThird + (2.0f / 3.0f) * (phi[n] + rk.dt * Lq);
      }
    }
    vmax = fmaxf(vmax, v);
    n = n1;
    n1 = n2;
  }
  if (!stats) return;
  // One block reduction of all statistics: max/min in float and counts in
  // uint32 (both exact in any order), the two sums in double with the same
  // xor-shuffle tree and wave order as block_reduce; a single LDS exchange.
  float f[9] = {vmax, qmin, qmax, dvmax, dmax, qmin_a, qmax_a, dvmax_a, dmax_a};
  constexpr bool fmx[9] = {true, false, true, true, true, false, true, true, true};  // max, else min
  uint32_t c[4] = {capped, nnear, capped_a, nall};
  double sm[2] = {vsum, vsum_a};
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const float u = __shfl_xor(f[t], o, 64);
      f[t] = fmx[t] ? fmaxf(f[t], u) : fminf(f[t], u);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) c[t] += __shfl_xor(c[t], o, 64);
#pragma unroll
    for (int t = 0; t < 2; ++t) sm[t] += __shfl_xor(sm[t], o, 64);
  }
  __shared__ float wf[4][9];
  __shared__ uint32_t wc[4][4];
  __shared__ double wd[4][2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int t = 0; t < 9; ++t) wf[wave][t] = f[t];
#pragma unroll
    for (int t = 0; t < 4; ++t) wc[wave][t] = c[t];
    wd[wave][0] = sm[0];
    wd[wave][1] = sm[1];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) {
#pragma unroll
      for (int t = 0; t < 9; ++t) f[t] = fmx[t] ? fmaxf(f[t], wf[w][t]) : fminf(f[t], wf[w][t]);
#pragma unroll
      for (int t = 0; t < 4; ++t) c[t] += wc[w][t];
      sm[0] += wd[w][0];
      sm[1] += wd[w][1];
    }
    double* o = partials + lb;
    o[0] = f[0];
    o[G] = f[1];
    o[2 * G] = f[2];
    o[3 * G] = f[3];
    o[4 * G] = (double)c[0];
    o[5 * G] = sm[0];
    o[6 * G] = (double)c[1];
    o[7 * G] = f[4];
    o[8 * G] = f[5];
    o[9 * G] = f[6];
    o[10 * G] = f[7];
    o[11 * G] = (double)c[2];
    o[12 * G] = sm[1];
    o[13 * G] = (double)c[3];
    o[14 * G] = f[8];
  }
  __syncthreads();  // wf/wc/wd are reused by the block's next segment
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
