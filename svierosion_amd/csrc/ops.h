// Per-node device math shared by every kernel of libsvimx.
//
// Layout conventions: phi and every dense field are C-order (x, y, z) float32
// grids with flat strides (ny*nz, nz, 1); a band is a sorted int32 list of flat
// node indices; points are in index (cell) coordinates.
//
// The formulas restate the reference's stencils (file:line cited per
// function) and keep their operation order, so float32 results agree with the
// reference's single-precision path to rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svimx {

struct Dims {
  int nx, ny, nz;
  __host__ __device__ int64_t sx() const { return (int64_t)ny * nz; }
  __host__ __device__ int64_t sy() const { return nz; }
  __host__ __device__ int64_t size() const { return (int64_t)nx * ny * nz; }
};

// Layout of the occupancy field f: the grid embedded with kOccPad zero cells in
// front of every axis and a z pitch that is a multiple of 4 floats, so the
// N_r summation reads aligned float4 windows and never leaves the array for
// ball offsets up to kOccPad cells.
constexpr int kOccPad = 16;

struct OccLayout {
  int nx, ny, nz;
  int64_t pitch_y;  // ny + 2 kOccPad
  int64_t pitch_z;  // multiple of 4, >= nz + 2 kOccPad + 8
  __host__ __device__ static OccLayout of(int nx, int ny, int nz) {
    OccLayout L;
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    L.pitch_y = ny + 2 * kOccPad;
    L.pitch_z = ((int64_t)nz + 2 * kOccPad + 16 + 3) / 4 * 4;
    return L;
  }
  __host__ __device__ int64_t size() const { return ((int64_t)nx + 2 * kOccPad) * pitch_y * pitch_z; }
  __host__ __device__ int64_t at(int x, int y, int z) const {
    return ((int64_t)(x + kOccPad) * pitch_y + (y + kOccPad)) * pitch_z + (z + kOccPad);
  }
};

__device__ __forceinline__ void unflat(const Dims& d, int64_t n, int& i, int& j, int& k) {
  // node indices fit 31 bits (band lists are int32): 32-bit unsigned divisions,
  // several times cheaper than 64-bit ones on the GPU
  const uint32_t u = (uint32_t)n, nz = (uint32_t)d.nz;
  const uint32_t syz = (uint32_t)d.ny * nz;
  const uint32_t a = u / syz;
  const uint32_t r = u - a * syz;
  const uint32_t b = r / nz;
  i = (int)a;
  j = (int)b;
  k = (int)(r - b * nz);
}

__device__ __forceinline__ float sq(float x) { return x * x; }

// Reciprocal and reciprocal square root by the hardware approximations
// (v_rcp_f32 / v_rsq_f32, within 1 ulp): the stencil formulas below multiply
// by them instead of dividing, which replaces a ~10-instruction IEEE division
// sequence by one instruction.  The results stay within a few ulp of the
// correctly rounded formulas.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

constexpr float kThird = 1.0f / 3.0f, kSixth = 1.0f / 6.0f, kFiveSixths = 5.0f / 6.0f;
constexpr float kSevenSixths = 7.0f / 6.0f, kElevenSixths = 11.0f / 6.0f, kThirteenTwelfths = 13.0f / 12.0f;

// Jiang-Peng WENO5 derivative from five one-sided differences
// (svierosion/kernels.py:83-97); `big` is the largest of v1^2 .. v5^2 (the
// caller forms the maxima of the two derivatives of an axis from one shared
// tree: max is exact in any order).  The normalised weights a_k / sum a with
// a_k = d_k / q_k, q_k = (s_k + eps)^2, are d_k prod_{j!=k} q_j over their
// sum: one reciprocal instead of four (v_rcp_f32 issues at a quarter of the
// rate).  The q_k are scaled by 2^20 first, which leaves the weights as they
// are: q_k >= 1e-20 (eps >= 1e-10), so the products of two stay normal
// floats, and |v| <= 12 in a field clamped to +-6h (s_k < 5e3) keeps them far
// from overflow.
__device__ __forceinline__ float weno5(float v1, float v2, float v3, float v4, float v5, float big) {
  float p1 = v1 * kThird - v2 * kSevenSixths + v3 * kElevenSixths;
  float p2 = -v2 * kSixth + v3 * kFiveSixths + v4 * kThird;
  float p3 = v3 * kThird + v4 * kFiveSixths - v5 * kSixth;
  float s1 = kThirteenTwelfths * sq(v1 - 2.0f * v2 + v3) + 0.25f * sq(v1 - 4.0f * v2 + 3.0f * v3);
  float s2 = kThirteenTwelfths * sq(v2 - 2.0f * v3 + v4) + 0.25f * sq(v2 - v4);
  float s3 = kThirteenTwelfths * sq(v3 - 2.0f * v4 + v5) + 0.25f * sq(3.0f * v3 - 4.0f * v4 + v5);
  float eps = 1e-6f * big + 1e-10f;
  constexpr float kScale = 1048576.0f;  // 2^20
  const float q1 = sq(s1 + eps) * kScale, q2 = sq(s2 + eps) * kScale, q3 = sq(s3 + eps) * kScale;
  const float a1 = 0.1f * (q2 * q3), a2 = 0.6f * (q1 * q3), a3 = 0.3f * (q1 * q2);
  return (a1 * p1 + a2 * p2 + a3 * p3) * rcp(a1 + a2 + a3);
}

// Godunov |grad phi| for phi_t = V |grad phi|, V >= 0, WENO5 one-sided
// derivatives (svierosion/kernels.py:146-169).  The upwind term
// max(min(D-, 0)^2, max(D+, 0)^2) is max(-D-, D+, 0)^2 (one v_max3_f32:
// squaring is monotone on non-negative floats); the square root is the
// hardware v_sqrt_f32 (1 ulp).
__device__ __forceinline__ float godunov_inward(const float* __restrict__ phi, int64_t n,
                                                const Dims& d, float inv_h) {
  const float c = phi[n];
  const int64_t st[3] = {d.sx(), d.sy(), 1};
  float g2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t s = st[a];
    float m3 = phi[n - 3 * s], m2 = phi[n - 2 * s], m1 = phi[n - s];
    float p1 = phi[n + s], p2 = phi[n + 2 * s], p3 = phi[n + 3 * s];
    float d0 = (m2 - m3) * inv_h, d1 = (m1 - m2) * inv_h, d2 = (c - m1) * inv_h;
    float d3 = (p1 - c) * inv_h, d4 = (p2 - p1) * inv_h, d5 = (p3 - p2) * inv_h;
    const float mid = fmaxf(fmaxf(d1 * d1, d2 * d2), fmaxf(d3 * d3, d4 * d4));
    float dm = weno5(d0, d1, d2, d3, d4, fmaxf(mid, d0 * d0));
    float dp = weno5(d5, d4, d3, d2, d1, fmaxf(mid, d5 * d5));
    const float g = fmaxf(fmaxf(-dm, dp), 0.0f);
    g2 = __fadd_rn(g2, __fmul_rn(g, g));
  }
  return __builtin_amdgcn_sqrtf(g2);
}

// Fraction of the unit cube on the side n.xi < -s of a plane; a >= b >= c >= 0
// are the sorted |n_i| (svierosion/kernels.py:104-122, report/report.tex:337-365).
__device__ __forceinline__ float cube_plane_fraction(float s, float a, float b, float c) {
  float M = 0.5f * (a + b + c);
  float y = M - s;
  bool flip = y > M;
  y = flip ? 2.0f * M - y : y;
  y = fmaxf(y, 0.0f);
  const float tiny = 1e-12f;
  float bs = fmaxf(b, tiny), cs = fmaxf(c, tiny);
  const float ib = rcp(bs), ic = rcp(cs), i6a = rcp(6.0f * a);
  float q = 3.0f * y * y - 3.0f * y * c + c * c;
  float H;
  if (y <= c) {
    H = y * y * y * (ib * ic);
  } else if (y <= b) {
    H = q * ib;
  } else if (y <= b + c) {
    float u = fmaxf(y - b, 0.0f);
    H = (q - u * u * u * ic) * ib;
  } else {
    H = 3.0f * (2.0f * y - b - c);
  }
  float ua = fmaxf(y - a, 0.0f);
  float F = H * i6a - ua * ua * ua * (i6a * ib * ic);
  F = flip ? 1.0f - F : F;
  return fminf(fmaxf(F, 0.0f), 1.0f);
}

__device__ __forceinline__ void sort3_abs(float x, float y, float z, float& hi, float& mid, float& lo) {
  x = fabsf(x); y = fabsf(y); z = fabsf(z);
  hi = fmaxf(fmaxf(x, y), z);
  lo = fminf(fminf(x, y), z);
  mid = x + y + z - hi - lo;
}

struct Grad {
  float c, gx, gy, gz;
};

__device__ __forceinline__ Grad central(const float* __restrict__ phi, int64_t n, const Dims& d) {
  Grad g;
  g.c = phi[n];
  g.gx = 0.5f * (phi[n + d.sx()] - phi[n - d.sx()]);
  g.gy = 0.5f * (phi[n + d.sy()] - phi[n - d.sy()]);
  g.gz = 0.5f * (phi[n + 1] - phi[n - 1]);
  return g;
}

// Tangent-plane cell fraction (svierosion/kernels.py:173-181).
__device__ __forceinline__ float cell_fraction(const Grad& g) {
  float gn = fmaxf(sqrtf(g.gx * g.gx + g.gy * g.gy + g.gz * g.gz), 1e-30f);
  const float ign = rcp(gn);
  float a, b, c;
  sort3_abs(g.gx * ign, g.gy * ign, g.gz * ign, a, b, c);
  return cube_plane_fraction(g.c * ign, a, b, c);
}

// First capped Newton projection onto the zero set, clamped to [1, n-3.001]
// (svierosion/kernels.py:185-208).  hi[] = n - 3.001 per axis.
__device__ __forceinline__ void closest_point(const Grad& g, int i, int j, int k, float h,
                                              const float hi[3], float p[3]) {
  float g2 = fmaxf(g.gx * g.gx + g.gy * g.gy + g.gz * g.gz, 1e-30f);
  float ac = fabsf(g.c);
  float length = ac * rsq(g2);
  float cap = 1.25f * ac * rcp(h) + 0.25f;
  float scale = (g.c * rcp(g2)) * fminf(cap * rcp(fmaxf(length, 1e-30f)), 1.0f);
  p[0] = fminf(fmaxf((float)i - scale * g.gx, 1.0f), hi[0]);
  p[1] = fminf(fmaxf((float)j - scale * g.gy, 1.0f), hi[1]);
  p[2] = fminf(fmaxf((float)k - scale * g.gz, 1.0f), hi[2]);
}

// Lagrange cubic weights on nodes -1, 0, 1, 2 (svierosion/kernels.py:134-137).
__device__ __forceinline__ void cubic_w(float t, float w[4]) {
  w[0] = -t * (t - 1.0f) * (t - 2.0f) * kSixth;
  w[1] = (t + 1.0f) * (t - 1.0f) * (t - 2.0f) * 0.5f;
  w[2] = -(t + 1.0f) * t * (t - 2.0f) * 0.5f;
  w[3] = (t + 1.0f) * t * (t - 1.0f) * kSixth;
}

// Tricubic interpolation at index point p (svierosion/kernels.py:212-228).
__device__ __forceinline__ float tricubic(const float* __restrict__ f, const Dims& d, const float p[3]) {
  float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
  float wx[4], wy[4], wz[4];
  cubic_w(p[0] - fx, wx);
  cubic_w(p[1] - fy, wy);
  cubic_w(p[2] - fz, wz);
  int64_t b0 = ((int64_t)fx - 1) * d.sx() + ((int64_t)fy - 1) * d.sy() + ((int64_t)fz - 1);
  float out = 0.0f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float wab = wx[a] * wy[b];
      const float* row = f + b0 + a * d.sx() + b * d.sy();
#pragma unroll
      for (int c = 0; c < 4; ++c) out = out + (wab * wz[c]) * row[c];
    }
  }
  return out;
}

// tricubic() of two interleaved fields at once: per component the same
// operations in the same order, so each equals tricubic() of that field.
__device__ __forceinline__ float2 tricubic2(const float2* __restrict__ f, const Dims& d, const float p[3]) {
  float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
  float wx[4], wy[4], wz[4];
  cubic_w(p[0] - fx, wx);
  cubic_w(p[1] - fy, wy);
  cubic_w(p[2] - fz, wz);
  int64_t b0 = ((int64_t)fx - 1) * d.sx() + ((int64_t)fy - 1) * d.sy() + ((int64_t)fz - 1);
  float2 out = make_float2(0.0f, 0.0f);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      float wab = wx[a] * wy[b];
      const float2* row = f + b0 + a * d.sx() + b * d.sy();
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float2 v = row[c];
        const float w = wab * wz[c];
        out.x = out.x + w * v.x;
        out.y = out.y + w * v.y;
      }
    }
  }
  return out;
}

// Derivatives of the Lagrange cubic weights (svierosion/kernels.py:139-141).
__device__ __forceinline__ void cubic_dw(float t, float w[4]) {
  w[0] = -(3.0f * t * t - 6.0f * t + 2.0f) * kSixth;
  w[1] = (3.0f * t * t - 4.0f * t - 1.0f) * 0.5f;
  w[2] = -(3.0f * t * t - 2.0f * t - 2.0f) * 0.5f;
  w[3] = (3.0f * t * t - 1.0f) * kSixth;
}

// Tricubic value and gradient (per cell) at p (svierosion/kernels.py:232-253).
__device__ __forceinline__ float tricubic_grad(const float* __restrict__ f, const Dims& d, const float p[3],
                                               float g[3]) {
  float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
  float wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
  cubic_w(p[0] - fx, wx); cubic_w(p[1] - fy, wy); cubic_w(p[2] - fz, wz);
  cubic_dw(p[0] - fx, dx); cubic_dw(p[1] - fy, dy); cubic_dw(p[2] - fz, dz);
  int64_t b0 = ((int64_t)fx - 1) * d.sx() + ((int64_t)fy - 1) * d.sy() + ((int64_t)fz - 1);
  float val = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const float* row = f + b0 + a * d.sx() + b * d.sy();
      for (int c = 0; c < 4; ++c) {
        float v = row[c];
        val = val + wx[a] * wy[b] * wz[c] * v;
        gx = gx + dx[a] * wy[b] * wz[c] * v;
        gy = gy + wx[a] * dy[b] * wz[c] * v;
        gz = gz + wx[a] * wy[b] * dz[c] * v;
      }
    }
  g[0] = gx; g[1] = gy; g[2] = gz;
  return val;
}

// Further capped Newton steps from a first projection p (svierosion/levelset.py:
// 139-145): the tricubic value and gradient of f at p, a step val/|g|^2 g of at
// most 1.25 |val|/h + 0.25 cells, then the clamp to where tricubic stencils fit.
__device__ __forceinline__ void newton_refine(const float* __restrict__ f, const Dims& d, float h, const float hi[3],
                                              float p[3], int extra) {
  for (int s = 0; s < extra; ++s) {
    float g[3];
    const float val = tricubic_grad(f, d, p, g);
    const float g2 = fmaxf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2], 1e-30f);
    const float av = fabsf(val);
    const float length = av / sqrtf(g2);
    const float cap = 1.25f * av / h + 0.25f;
    const float scale = (val / g2) * fminf(cap / fmaxf(length, 1e-30f), 1.0f);
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = fminf(fmaxf(p[a] - scale * g[a], 1.0f), hi[a]);
  }
}

// Trilinear interpolation at index point p (svierosion/kernels.py:257-270).
__device__ __forceinline__ float trilinear(const float* __restrict__ f, const Dims& d, const float p[3]) {
  float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
  float t0 = p[0] - fx, t1 = p[1] - fy, t2 = p[2] - fz;
  int64_t b0 = (int64_t)fx * d.sx() + (int64_t)fy * d.sy() + (int64_t)fz;
  float out = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    float wa = a ? t0 : 1.0f - t0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float wb = wa * (b ? t1 : 1.0f - t1);
#pragma unroll
      for (int c = 0; c < 2; ++c)
        out = out + wb * (c ? t2 : 1.0f - t2) * f[b0 + a * d.sx() + b * d.sy() + c];
    }
  }
  return out;
}

// trilinear() of two interleaved fields at once: per component the same
// operations in the same order, so each equals trilinear() of that field.
__device__ __forceinline__ float2 trilinear2(const float2* __restrict__ f, const Dims& d, const float p[3]) {
  float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
  float t0 = p[0] - fx, t1 = p[1] - fy, t2 = p[2] - fz;
  int64_t b0 = (int64_t)fx * d.sx() + (int64_t)fy * d.sy() + (int64_t)fz;
  float o1 = 0.0f, o2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    float wa = a ? t0 : 1.0f - t0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float wb = wa * (b ? t1 : 1.0f - t1);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float w = wb * (c ? t2 : 1.0f - t2);
        const float2 v = f[b0 + a * d.sx() + b * d.sy() + c];
        o1 = o1 + w * v.x;
        o2 = o2 + w * v.y;
      }
    }
  }
  return make_float2(o1, o2);
}

// Principal curvatures k1 >= k2 of the level set through node n from second
// central differences (svierosion/kernels.py:274-299).
__device__ __forceinline__ void principal_curvatures(const float* __restrict__ f, int64_t n,
                                                     const Dims& d, float h, float& k1, float& k2) {
  const int64_t sx = d.sx(), sy = d.sy(), sz = 1;
  float c = f[n];
  float xp = f[n + sx], xm = f[n - sx], yp = f[n + sy], ym = f[n - sy], zp = f[n + sz], zm = f[n - sz];
  const float i2h = rcp(2.0f * h), ihh = rcp(h * h), iq4 = rcp(4.0f * h * h);
  float px = (xp - xm) * i2h, py = (yp - ym) * i2h, pz = (zp - zm) * i2h;
  float pxx = (xp - 2.0f * c + xm) * ihh;
  float pyy = (yp - 2.0f * c + ym) * ihh;
  float pzz = (zp - 2.0f * c + zm) * ihh;
  float pxy = (f[n + sx + sy] - f[n + sx - sy] - f[n - sx + sy] + f[n - sx - sy]) * iq4;
  float pxz = (f[n + sx + sz] - f[n + sx - sz] - f[n - sx + sz] + f[n - sx - sz]) * iq4;
  float pyz = (f[n + sy + sz] - f[n + sy - sz] - f[n - sy + sz] + f[n - sy - sz]) * iq4;
  float g2 = fmaxf(px * px + py * py + pz * pz, 1e-30f);
  float gn = sqrtf(g2);
  float ksum = ((pyy + pzz) * px * px + (pxx + pzz) * py * py + (pxx + pyy) * pz * pz -
                2.0f * (px * py * pxy + px * pz * pxz + py * pz * pyz)) * rcp(g2 * gn);
  float K = (px * px * (pyy * pzz - pyz * pyz) + py * py * (pxx * pzz - pxz * pxz) +
             pz * pz * (pxx * pyy - pxy * pxy) + 2.0f * px * py * (pxz * pyz - pxy * pzz) +
             2.0f * py * pz * (pxy * pxz - pyz * pxx) + 2.0f * px * pz * (pxy * pyz - pxz * pyy)) *
            rcp(g2 * g2);
  float H = 0.5f * ksum;
  float r = sqrtf(fmaxf(H * H - K, 0.0f));
  k1 = H + r;
  k2 = H - r;
}

// ---------------------------------------------------------------- reinitialisation
template <typename Real>
__device__ __forceinline__ Real minmod(Real a, Real b) {
  return (a * b > Real(0)) ? (fabs(a) < fabs(b) ? a : b) : Real(0);
}

// Quadratic-ENO distance from a node (p0) to the zero towards a neighbour (p1)
// (Min 2010; svierosion/kernels.py:312-331), in the arithmetic of Real.
template <typename Real>
__device__ __forceinline__ Real interface_distance(Real p0, Real p1, Real dd, Real h) {
  const Real den = p0 - p1;
  const Real lin = h * p0 / (den != Real(0) ? den : Real(1e-30));
  const Real c2 = Real(0.5) * dd / (h * h);
  const Real c1 = (p1 - p0) / h;
  const Real c0 = Real(0.5) * (p0 + p1) - c2 * h * h / Real(4);
  const Real disc = c1 * c1 - Real(4) * c2 * c0;
  const Real sqd = sqrt(fmax(disc, Real(0)));
  const Real q = Real(-0.5) * (c1 + (c1 >= Real(0) ? sqd : -sqd));
  const Real xi = c0 / (q != Real(0) ? q : Real(1e-30));
  const Real quad = h / Real(2) + xi;
  const bool ok = (disc >= Real(0)) && (fabs(dd) > Real(1e-10) * h * h) && (quad >= Real(0)) && (quad <= h);
  return fmin(fmax(ok ? quad : lin, Real(0)), h);
}

// Reinit geometry of one node: flags bit (2a) = crossing towards +axis a,
// bit (2a+1) = towards -axis a; bits 6-7 encode sgn (1 -> +1, 2 -> -1).
// x[2a] / x[2a+1] = subcell distances (svierosion/kernels.py:335-359).
struct ReinitGeo {
  float dtau;
  float x[6];
  uint32_t flags;
};

// The crossings, subcell distances and local dtau are formed in the
// arithmetic of Real from the float32 field and rounded to float32 once.
// The reinitialisation uses Real = double: next to the zero set the
// quadratic root is ill-conditioned (a node within ~1e-5 h of zero), and
// float32 roots there moved the reinitialised field by up to 3.5e-4 h from
// the same algorithm in FP64, against 8e-5 h with FP64 roots.
template <typename Real>
__device__ __forceinline__ ReinitGeo reinit_geometry(const float* __restrict__ f, int64_t n, const Dims& d,
                                                     float hf, float cflf) {
  ReinitGeo g;
  const int64_t st[3] = {d.sx(), d.sy(), 1};
  const Real h = (Real)hf, cfl = (Real)cflf;
  const Real c = (Real)f[n];
  Real hmin = h;
  uint32_t flags = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t s = st[a];
    const Real m2 = (Real)f[n - 2 * s], m1 = (Real)f[n - s], p1 = (Real)f[n + s], p2 = (Real)f[n + 2 * s];
    const Real d_m1 = c - Real(2) * m1 + m2;
    const Real d_0 = p1 - Real(2) * c + m1;
    const Real d_p1 = p2 - Real(2) * p1 + c;
    const bool cp = (c * p1) < Real(0);
    const bool cm = (c * m1) < Real(0);
    // a distance is read only where its crossing flag is set (most work
    // nodes have none): the root is solved only there
    const Real xp = cp ? interface_distance<Real>(c, p1, minmod<Real>(d_0, d_p1), h) : h;
    const Real xm = cm ? interface_distance<Real>(c, m1, minmod<Real>(d_0, d_m1), h) : h;
    if (cp) hmin = fmin(hmin, xp);
    if (cm) hmin = fmin(hmin, xm);
    flags |= (cp ? 1u : 0u) << (2 * a);
    flags |= (cm ? 1u : 0u) << (2 * a + 1);
    g.x[2 * a] = (float)fmax(xp, Real(1e-6) * h);
    g.x[2 * a + 1] = (float)fmax(xm, Real(1e-6) * h);
  }
  flags |= (c > Real(0) ? 1u : (c < Real(0) ? 2u : 0u)) << 6;
  g.flags = flags;
  g.dtau = (float)(cfl * fmax(hmin, Real(1e-3) * h));
  return g;
}

// -sgn(phi0)(|grad phi| - 1) with ENO2 differences and the subcell fix
// (svierosion/kernels.py:363-379).
// The 13 values the right-hand side reads: the node and +-1, +-2 along each
// axis (axis 0 = x, 2 = z), loaded apart from the arithmetic so a caller can
// put the loads of several nodes in flight before computing any of them.
struct Stencil13 {
  float c, m2[3], m1[3], p1[3], p2[3];
};

// Loads through byte offsets of type Off from the field's base: with uint32_t
// offsets (grids below 2^30 nodes) every load is one scalar-base + 32-bit
// vector-offset instruction and the z neighbours share the node's offset
// through the instruction's immediate field.
template <typename Off>
__device__ __forceinline__ Stencil13 load_stencil13_at(const float* __restrict__ f, Off node, Off sx4, Off sy4) {
  const char* B = reinterpret_cast<const char*>(f);
  const Off o = node * (Off)4;
  const float* p = reinterpret_cast<const float*>(B + o);
  Stencil13 S;
  S.c = p[0];
  S.m2[2] = p[-2];
  S.m1[2] = p[-1];
  S.p1[2] = p[1];
  S.p2[2] = p[2];
  S.m2[1] = *reinterpret_cast<const float*>(B + (o - 2 * sy4));
  S.m1[1] = *reinterpret_cast<const float*>(B + (o - sy4));
  S.p1[1] = *reinterpret_cast<const float*>(B + (o + sy4));
  S.p2[1] = *reinterpret_cast<const float*>(B + (o + 2 * sy4));
  S.m2[0] = *reinterpret_cast<const float*>(B + (o - 2 * sx4));
  S.m1[0] = *reinterpret_cast<const float*>(B + (o - sx4));
  S.p1[0] = *reinterpret_cast<const float*>(B + (o + sx4));
  S.p2[0] = *reinterpret_cast<const float*>(B + (o + 2 * sx4));
  return S;
}

// minmod(a, b) as the median of (a, b, 0), one v_med3_f32: the smaller
// magnitude when a and b share a sign, 0 otherwise.  It equals the
// sign-product form (a b > 0 ? ...) except where a b underflows to 0 for two
// nonzero values of one sign (|a b| < 1e-38: second differences of a
// distance field are never that small unless one is exactly 0).
__device__ __forceinline__ float minmod_med3(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.0f); }

// Right-hand side at one node.  s = sgn(phi0) in {+1, -1, 0} (info bits 6-7),
// ih = 1/h, ihh = 1/h^2.  The Godunov upwind choice of the reference,
//   outside: max(max(Dm,0)^2, min(Dp,0)^2)   inside: max(min(Dm,0)^2, max(Dp,0)^2),
// is max(s Dm, -s Dp, 0)^2 (squaring is monotone on non-negative floats, so
// the maximum may be taken before it); s is folded into the factors of each
// one-sided difference (a sign flip of one factor flips the rounded result
// exactly), so each axis costs one v_max3_f32 after the differences.
// sqrt is the hardware v_sqrt_f32 (1 ulp).
__device__ __forceinline__ float reinit_rhs_lean(const Stencil13& S, float s, float ih, float ihh, uint32_t flags,
                                                 const float x[6]) {
  const float c = S.c;
  const float sih = s * ih, sihh = s * ihh, sc = s * c;
  float g2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float m2 = S.m2[a], m1 = S.m1[a], p1 = S.p1[a], p2 = S.p2[a];
    const float d0 = p1 - 2.0f * c + m1;
    const float mm_p = minmod_med3(d0, p2 - 2.0f * p1 + c);
    const float mm_m = minmod_med3(d0, c - 2.0f * m1 + m2);
    const bool cp = (flags >> (2 * a)) & 1u;
    const bool cm = (flags >> (2 * a + 1)) & 1u;
    const float xp = x[2 * a], xm = x[2 * a + 1];
    // s * Dp and s * Dm
    const float sDp = cp ? (-sc * rcp(xp) - 0.5f * xp * mm_p * sihh) : ((p1 - c - 0.5f * mm_p) * sih);
    const float sDm = cm ? (sc * rcp(xm) + 0.5f * xm * mm_m * sihh) : ((c - m1 + 0.5f * mm_m) * sih);
    const float g = fmaxf(fmaxf(sDm, -sDp), 0.0f);  // one v_max3_f32
    g2 = __fadd_rn(g2, __fmul_rn(g, g));
  }
  return -s * (__builtin_amdgcn_sqrtf(g2) - 1.0f);
}

// ---------------------------------------------------------------- point-triangle distance
// Squared distance from p to triangle (a, b, c), Ericson 5.1.5 region tests with
// the reference's priority order (svierosion/kernels.py:382-427).  Inputs are
// in a coordinate frame local to the triangle so float32 keeps ~1e-7 cells.
__device__ __forceinline__ float point_triangle_d2(const float p[3], const float a[3],
                                                   const float b[3], const float c[3]) {
  float ab[3], ac[3], ap[3], bp[3], cp[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    ab[q] = b[q] - a[q]; ac[q] = c[q] - a[q];
    ap[q] = p[q] - a[q]; bp[q] = p[q] - b[q]; cp[q] = p[q] - c[q];
  }
  float d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
  float d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
  float d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
  float d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
  float d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
  float d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
  float va = d3 * d6 - d5 * d4;
  float vb = d5 * d2 - d1 * d6;
  float vc = d1 * d4 - d3 * d2;
  const float tiny = 1e-30f;
  float den = va + vb + vc;
  den = fabsf(den) > tiny ? den : tiny;
  float v, w;
  // highest priority first: the first matching region wins
  if (d1 <= 0.0f && d2 <= 0.0f) {                       // vertex a
    v = 0.0f; w = 0.0f;
  } else if (d3 >= 0.0f && d4 <= d3) {                  // vertex b
    v = 1.0f; w = 0.0f;
  } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {  // edge ab
    v = d1 / fmaxf(d1 - d3, tiny); w = 0.0f;
  } else if (d6 >= 0.0f && d5 <= d6) {                  // vertex c
    v = 0.0f; w = 1.0f;
  } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {  // edge ac
    v = 0.0f; w = d2 / fmaxf(d2 - d6, tiny);
  } else {
    float e1 = d4 - d3, e2 = d5 - d6;
    if (va <= 0.0f && e1 >= 0.0f && e2 >= 0.0f) {        // edge bc
      float t = e1 / fmaxf(e1 + e2, tiny);
      v = 1.0f - t; w = t;
    } else {                                             // interior
      v = vb / den; w = vc / den;
    }
  }
  float r = 0.0f;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float qq = a[q] + v * ab[q] + w * ac[q];
    r += sq(p[q] - qq);
  }
  return r;
}

}  // namespace svimx
