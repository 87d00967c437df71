// CPU restatement of the SVI erosion step in C++17 with OpenMP, float32 like
// the reference's CPU path (svierosion/evolve.py:108-250 with
// svierosion/kernels.py, svierosion/levelset.py): TEST INFRASTRUCTURE, the
// timed CPU baseline of bench.py and a second restatement that the tests check
// against the NumPy oracle.  Nothing in svierosion_amd links it.
//
// One step: band (|phi| < width h (1 - 1e-5)) -> for each TVD-RK3 stage:
// tangent-plane cell fractions on the band and the dense occupancy, closest
// points and tricubic N_r at them (N_r summed directly over the ball weights at
// the nodes the tricubic stencils read, float32 accumulation), SVI speed,
// smoothstep cutoff, WENO5-Godunov |grad phi| -> stage update; FP64 volume;
// the displacement trigger; reinitialisation (Min 2010: band dilated by 3,
// quadratic-ENO subcell crossings, TVD-RK2 pseudo time) and band rebuild.
// The ball weights and n_cap come from the caller (the NumPy oracle computes
// them).  Parallel loops use OpenMP; every reduction is deterministic per
// thread count (static schedule).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Grid {
  int nx, ny, nz;
  int64_t sx, sy, n;
};

inline float sq(float x) { return x * x; }

// ---------------------------------------------------------------- per-node formulas
// svierosion/kernels.py:83-97
inline float weno5(float v1, float v2, float v3, float v4, float v5) {
  const float p1 = v1 / 3.0f - 7.0f * v2 / 6.0f + 11.0f * v3 / 6.0f;
  const float p2 = -v2 / 6.0f + 5.0f * v3 / 6.0f + v4 / 3.0f;
  const float p3 = v3 / 3.0f + 5.0f * v4 / 6.0f - v5 / 6.0f;
  const float s1 = 13.0f / 12.0f * sq(v1 - 2.0f * v2 + v3) + 0.25f * sq(v1 - 4.0f * v2 + 3.0f * v3);
  const float s2 = 13.0f / 12.0f * sq(v2 - 2.0f * v3 + v4) + 0.25f * sq(v2 - v4);
  const float s3 = 13.0f / 12.0f * sq(v3 - 2.0f * v4 + v5) + 0.25f * sq(3.0f * v3 - 4.0f * v4 + v5);
  const float big = std::max(std::max(std::max(v1 * v1, v2 * v2), std::max(v3 * v3, v4 * v4)), v5 * v5);
  const float eps = 1e-6f * big + 1e-10f;
  const float a1 = 0.1f / sq(s1 + eps), a2 = 0.6f / sq(s2 + eps), a3 = 0.3f / sq(s3 + eps);
  return (a1 * p1 + a2 * p2 + a3 * p3) / (a1 + a2 + a3);
}

// svierosion/kernels.py:146-169
inline float godunov(const float* f, int64_t n, const Grid& g, float ih) {
  const float c = f[n];
  const int64_t st[3] = {g.sx, g.sy, 1};
  float acc = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const int64_t s = st[a];
    const float d0 = (f[n - 2 * s] - f[n - 3 * s]) * ih, d1 = (f[n - s] - f[n - 2 * s]) * ih;
    const float d2 = (c - f[n - s]) * ih, d3 = (f[n + s] - c) * ih;
    const float d4 = (f[n + 2 * s] - f[n + s]) * ih, d5 = (f[n + 3 * s] - f[n + 2 * s]) * ih;
    const float back = weno5(d0, d1, d2, d3, d4), fwd = weno5(d5, d4, d3, d2, d1);
    acc += std::max(sq(std::min(back, 0.0f)), sq(std::max(fwd, 0.0f)));
  }
  return std::sqrt(acc);
}

// svierosion/kernels.py:104-131, :173-181
inline float cube_plane_fraction(float s, float a, float b, float c) {
  const float half = 0.5f * (a + b + c);
  float y = half - s;
  const bool flip = y > half;
  y = std::max(flip ? 2.0f * half - y : y, 0.0f);
  const float bs = std::max(b, 1e-12f), cs = std::max(c, 1e-12f);
  const float q = 3.0f * y * y - 3.0f * y * c + c * c;
  float G;
  if (y <= c) G = y * y * y / (bs * cs);
  else if (y <= b) G = q / bs;
  else if (y <= b + c) G = (q - std::pow(std::max(y - b, 0.0f), 3.0f) / cs) / bs;
  else G = 3.0f * (2.0f * y - b - c);
  const float oa = std::max(y - a, 0.0f);
  float F = G / (6.0f * a) - oa * oa * oa / (6.0f * a * bs * cs);
  F = flip ? 1.0f - F : F;
  return std::min(std::max(F, 0.0f), 1.0f);
}

inline float cell_fraction(const float* f, int64_t n, const Grid& g) {
  const float gx = 0.5f * (f[n + g.sx] - f[n - g.sx]), gy = 0.5f * (f[n + g.sy] - f[n - g.sy]);
  const float gz = 0.5f * (f[n + 1] - f[n - 1]);
  const float gn = std::max(std::sqrt(gx * gx + gy * gy + gz * gz), 1e-30f);
  float x = std::fabs(gx / gn), y = std::fabs(gy / gn), z = std::fabs(gz / gn);
  const float hi = std::max(std::max(x, y), z), lo = std::min(std::min(x, y), z);
  return cube_plane_fraction(f[n] / gn, hi, x + y + z - hi - lo, lo);
}

// svierosion/kernels.py:185-208 (one capped Newton step, clamped to [1, n - 3.001])
inline void closest_point(const float* f, int64_t n, const Grid& g, float h, float p[3]) {
  const float c = f[n];
  const float gx = 0.5f * (f[n + g.sx] - f[n - g.sx]), gy = 0.5f * (f[n + g.sy] - f[n - g.sy]);
  const float gz = 0.5f * (f[n + 1] - f[n - 1]);
  const float g2 = std::max(gx * gx + gy * gy + gz * gz, 1e-30f);
  const float len = std::fabs(c) / std::sqrt(g2);
  const float cap = 1.25f * std::fabs(c) / h + 0.25f;
  const float scale = (c / g2) * std::min(cap / std::max(len, 1e-30f), 1.0f);
  const int64_t i = n / g.sx, j = (n / g.sy) % g.ny, k = n % g.nz;
  const float hi[3] = {(float)(g.nx - 3.001), (float)(g.ny - 3.001), (float)(g.nz - 3.001)};
  const float q[3] = {(float)i - scale * gx, (float)j - scale * gy, (float)k - scale * gz};
  for (int a = 0; a < 3; ++a) p[a] = std::min(std::max(q[a], 1.0f), hi[a]);
}

// svierosion/kernels.py:134-137, :212-228
inline void cubic_w(float t, float w[4]) {
  w[0] = -t * (t - 1.0f) * (t - 2.0f) / 6.0f;
  w[1] = (t + 1.0f) * (t - 1.0f) * (t - 2.0f) / 2.0f;
  w[2] = -(t + 1.0f) * t * (t - 2.0f) / 2.0f;
  w[3] = (t + 1.0f) * t * (t - 1.0f) / 6.0f;
}
inline int64_t stencil_base(const float p[3], const Grid& g) {
  return ((int64_t)std::floor(p[0]) - 1) * g.sx + ((int64_t)std::floor(p[1]) - 1) * g.sy +
         ((int64_t)std::floor(p[2]) - 1);
}
inline float tricubic(const float* f, const Grid& g, const float p[3]) {
  float wx[4], wy[4], wz[4];
  cubic_w(p[0] - std::floor(p[0]), wx);
  cubic_w(p[1] - std::floor(p[1]), wy);
  cubic_w(p[2] - std::floor(p[2]), wz);
  const int64_t b0 = stencil_base(p, g);
  float out = 0.0f;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const float wab = wx[a] * wy[b];
      const float* row = f + b0 + a * g.sx + b * g.sy;
      for (int c = 0; c < 4; ++c) out += (wab * wz[c]) * row[c];
    }
  return out;
}

struct Law {  // svierosion/flows.py:29-122
  int k, s;
  float t, eps, n_cap;
  float speed(float N) const {
    N = std::min(std::max(N, 0.0f), 1.0f);
    float V;
    if (s) {
      N = std::min(N, n_cap);
      V = std::pow(N, (float)k) / (t - N);
    } else {
      V = k ? std::pow(N, (float)k) : 1.0f;
    }
    return eps != 0.0f ? V + eps : V;
  }
  float dspeed(float N) const {
    N = std::min(std::max(N, 1e-12f), 1.0f);
    if (s) {
      N = std::min(N, n_cap);
      const float d = t - N, V = std::pow(N, (float)k) / d;
      return V * ((float)k / N + 1.0f / d);
    }
    return k ? (float)k * std::pow(N, (float)(k - 1)) : 0.0f;
  }
};

// svierosion/kernels.py:100-101, :312-379 (Min 2010)
inline float minmod(float a, float b) { return a * b > 0.0f ? (std::fabs(a) < std::fabs(b) ? a : b) : 0.0f; }
inline float interface_distance(float p0, float p1, float dd, float h) {
  const float den = p0 - p1;
  const float lin = h * p0 / (den != 0.0f ? den : 1e-30f);
  const float c2 = 0.5f * dd / (h * h), c1 = (p1 - p0) / h, c0 = 0.5f * (p0 + p1) - c2 * h * h / 4.0f;
  const float disc = c1 * c1 - 4.0f * c2 * c0;
  const float sd = std::sqrt(std::max(disc, 0.0f));
  const float q = -0.5f * (c1 + (c1 >= 0.0f ? sd : -sd));
  const float xi = c0 / (q != 0.0f ? q : 1e-30f);
  const float quad = h / 2.0f + xi;
  const bool ok = disc >= 0.0f && std::fabs(dd) > 1e-10f * h * h && quad >= 0.0f && quad <= h;
  return std::min(std::max(ok ? quad : lin, 0.0f), h);
}
struct Geo {
  float sgn, dtau, x[6];
  uint8_t cross;  // bit 2a: +axis a, bit 2a+1: -axis a
};
inline Geo reinit_geometry(const float* f, int64_t n, const Grid& g, float h, float cfl) {
  Geo o;
  const float c = f[n];
  o.sgn = c > 0.0f ? 1.0f : (c < 0.0f ? -1.0f : 0.0f);
  float hmin = h;
  o.cross = 0;
  const int64_t st[3] = {g.sx, g.sy, 1};
  for (int a = 0; a < 3; ++a) {
    const int64_t s = st[a];
    const float m2 = f[n - 2 * s], m1 = f[n - s], p1 = f[n + s], p2 = f[n + 2 * s];
    const float dm1 = c - 2.0f * m1 + m2, d0 = p1 - 2.0f * c + m1, dp1 = p2 - 2.0f * p1 + c;
    const bool cp = c * p1 < 0.0f, cm = c * m1 < 0.0f;
    const float xp = interface_distance(c, p1, minmod(d0, dp1), h);
    const float xm = interface_distance(c, m1, minmod(d0, dm1), h);
    if (cp) hmin = std::min(hmin, xp);
    if (cm) hmin = std::min(hmin, xm);
    o.cross |= (uint8_t)((cp ? 1 : 0) << (2 * a)) | (uint8_t)((cm ? 1 : 0) << (2 * a + 1));
    o.x[2 * a] = std::max(xp, 1e-6f * h);
    o.x[2 * a + 1] = std::max(xm, 1e-6f * h);
  }
  o.dtau = cfl * std::max(hmin, 1e-3f * h);
  return o;
}
inline float reinit_rhs(const float* f, int64_t n, const Grid& g, float h, const Geo& q) {
  const float c = f[n];
  const bool outside = q.sgn > 0.0f;
  const int64_t st[3] = {g.sx, g.sy, 1};
  float acc = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const int64_t s = st[a];
    const float m2 = f[n - 2 * s], m1 = f[n - s], p1 = f[n + s], p2 = f[n + 2 * s];
    const float mmp = minmod(p1 - 2.0f * c + m1, p2 - 2.0f * p1 + c);
    const float mmm = minmod(p1 - 2.0f * c + m1, c - 2.0f * m1 + m2);
    const float xp = q.x[2 * a], xm = q.x[2 * a + 1];
    const float Dp = (q.cross >> (2 * a)) & 1 ? -c / xp - 0.5f * xp * mmp / (h * h) : (p1 - c - 0.5f * mmp) / h;
    const float Dm = (q.cross >> (2 * a + 1)) & 1 ? c / xm + 0.5f * xm * mmm / (h * h) : (c - m1 + 0.5f * mmm) / h;
    const float pos = std::max(sq(std::max(Dm, 0.0f)), sq(std::min(Dp, 0.0f)));
    const float neg = std::max(sq(std::min(Dm, 0.0f)), sq(std::max(Dp, 0.0f)));
    acc += outside ? pos : neg;
  }
  return -q.sgn * (std::sqrt(acc) - 1.0f);
}

// ---------------------------------------------------------------- the simulation
struct Sim {
  Grid g;
  float h, width, cut_lo, cut_hi;
  double radius, h3_over_bv;
  Law law;
  std::vector<float> phi, occ, nr, u0, L, V, scratch;  // occ: zero-padded by pm cells on every side
  int pm = 0;
  int64_t psx = 0, psy = 0;
  int64_t pidx(int64_t n) const {
    const int64_t i = n / g.sx, j = (n / g.sy) % g.ny, k = n % g.nz;
    return (i + pm) * psx + (j + pm) * psy + (k + pm);
  }
  std::vector<uint8_t> mark;
  std::vector<int64_t> band;
  std::vector<int64_t> tap_off;  // nonzero ball weights: flat offsets and float32 weights
  std::vector<float> tap_w;
  double cfl, feedback_cfl, reinit_distance, displacement = 0.0;
  int reinit_iterations;

  // in node order; per-thread pieces concatenated in thread order
  template <class Pred>
  std::vector<int64_t> select(Pred pred) const {
    const int nt = omp_get_max_threads();
    std::vector<std::vector<int64_t>> part(nt);
#pragma omp parallel
    {
      const int t = omp_get_thread_num();
      const int64_t lo = g.n * t / nt, hi = g.n * (t + 1) / nt;
      for (int64_t n = lo; n < hi; ++n)
        if (pred(n)) part[t].push_back(n);
    }
    std::vector<int64_t> out;
    for (auto& p : part) out.insert(out.end(), p.begin(), p.end());
    return out;
  }

  void rebuild_band() {  // svierosion/levelset.py:70-86
    const float cap = width * h, thr = (float)(width * h * (1.0 - 1e-5));
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < g.n; ++n) phi[n] = std::min(std::max(phi[n], -cap), cap);
    band = select([&](int64_t n) { return std::fabs(phi[n]) < thr; });
  }

  double volume() {  // svierosion/levelset.py:123-126
    double frac = 0.0;
    int64_t inside = 0;
#pragma omp parallel for schedule(static) reduction(+ : frac)
    for (int64_t q = 0; q < (int64_t)band.size(); ++q) {
      const int64_t n = band[q];
      frac += (double)cell_fraction(phi.data(), n, g);
      if (phi[n] < 0.0f) frac -= 1.0;  // replaces the inside count of the band node
    }
#pragma omp parallel for schedule(static) reduction(+ : inside)
    for (int64_t n = 0; n < g.n; ++n) inside += phi[n] < 0.0f ? 1 : 0;
    return ((double)inside + frac) * (double)h * (double)h * (double)h;
  }

  // L = V |grad phi| on the band; stats: vmax and dvmax over |phi| < h
  void rhs(bool stats, float& vmax, float& dvmax) {
    const float* f = phi.data();
    // dense occupancy: 1 inside, band fractions (svierosion/levelset.py:117-121)
#pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < g.n; ++n) occ[pidx(n)] = f[n] < 0.0f ? 1.0f : 0.0f;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)band.size(); ++q) occ[pidx(band[q])] = cell_fraction(f, band[q], g);
    // N_r at the tricubic stencils of the closest points (svierosion/svi.py:84-93)
    std::fill(mark.begin(), mark.end(), 0);
    const int64_t nb = (int64_t)band.size();
    std::vector<float> P(3 * nb);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb; ++q) {
      const int64_t n = band[q];
      if (!(std::fabs(f[n]) < cut_hi)) continue;
      closest_point(f, n, g, h, &P[3 * q]);
      const int64_t b0 = stencil_base(&P[3 * q], g);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          for (int c = 0; c < 4; ++c) mark[b0 + a * g.sx + b * g.sy + c] = 1;
    }
    const int64_t ntap = (int64_t)tap_off.size();
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t n = 0; n < g.n; ++n) {
      if (!mark[n]) continue;
      const float* o = occ.data() + pidx(n);
      float s = 0.0f;
      for (int64_t t = 0; t < ntap; ++t) s += tap_w[t] * o[tap_off[t]];
      nr[n] = (float)(1.0 - h3_over_bv * (double)s);
    }
    float vm = 0.0f, dvm = 0.0f;
#pragma omp parallel for schedule(static) reduction(max : vm, dvm)
    for (int64_t q = 0; q < nb; ++q) {
      const int64_t n = band[q];
      const float x = std::fabs(f[n]);
      float v = 0.0f;
      if (x < cut_hi) {
        const float N = tricubic(nr.data(), g, &P[3 * q]);
        const float tt = std::min(std::max((cut_hi - x) / (cut_hi - cut_lo), 0.0f), 1.0f);
        v = law.speed(N) * (tt * tt * (3.0f - 2.0f * tt));
        if (stats && x < h) dvm = std::max(dvm, law.dspeed(N));
      }
      V[q] = v;
      vm = std::max(vm, v);
      L[q] = v * godunov(f, n, g, 1.0f / h);
    }
    vmax = vm;
    dvmax = dvm;
  }

  void reinitialize() {  // svierosion/levelset.py:154-183
    const int dil = 3, margin = 4;
    std::vector<uint8_t> m1(g.n, 0), m2(g.n, 0);
    for (int64_t n : band) m1[n] = 1;
    const int64_t st[3] = {g.sx, g.sy, 1};
    const int dims[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {  // separable (2 dil + 1)^3 dilation
#pragma omp parallel for schedule(static)
      for (int64_t n = 0; n < g.n; ++n) {
        const int64_t c = (n / st[a]) % dims[a];
        uint8_t v = 0;
        for (int d = -dil; d <= dil; ++d)
          if (c + d >= 0 && c + d < dims[a]) v |= m1[n + d * st[a]];
        m2[n] = v;
      }
      std::swap(m1, m2);
    }
    const auto work = select([&](int64_t n) {
      const int64_t i = n / g.sx, j = (n / g.sy) % g.ny, k = n % g.nz;
      return m1[n] && i >= margin && i < g.nx - margin && j >= margin && j < g.ny - margin && k >= margin &&
             k < g.nz - margin;
    });
    const int64_t nw = (int64_t)work.size();
    std::vector<Geo> geo(nw);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nw; ++q) geo[q] = reinit_geometry(phi.data(), work[q], g, h, 0.45f);
    scratch = phi;
    for (int it = 0; it < reinit_iterations; ++it) {
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < nw; ++q) {
        const int64_t n = work[q];
        scratch[n] = phi[n] + geo[q].dtau * reinit_rhs(phi.data(), n, g, h, geo[q]);
      }
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < nw; ++q) {
        const int64_t n = work[q];
        phi[n] = 0.5f * (phi[n] + scratch[n] + geo[q].dtau * reinit_rhs(scratch.data(), n, g, h, geo[q]));
      }
    }
    rebuild_band();
  }

  // one accepted TVD-RK3 step (svierosion/evolve.py:172-179, :203-242); returns dt
  double step(double& vol, int& did) {
    const int64_t nb = (int64_t)band.size();
    for (int64_t q = 0; q < nb; ++q) u0[q] = phi[band[q]];
    float vmax, dvmax;
    rhs(true, vmax, dvmax);
    double dt = vmax > 0 ? cfl * h / vmax : INFINITY;
    if (dvmax > 0) dt = std::min(dt, feedback_cfl * (4.0 * radius / 3.0) / dvmax);
    const float d = (float)dt;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb; ++q) phi[band[q]] = u0[q] + d * L[q];
    float a, b;
    rhs(false, a, b);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb; ++q) phi[band[q]] = 0.75f * u0[q] + 0.25f * (phi[band[q]] + d * L[q]);
    rhs(false, a, b);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb; ++q) phi[band[q]] = u0[q] / 3.0f + (2.0f / 3.0f) * (phi[band[q]] + d * L[q]);
    displacement += dt * vmax;
    did = 0;
    if (displacement >= reinit_distance * h) {
      reinitialize();
      u0.resize(band.size());
      L.resize(band.size());
      V.resize(band.size());
      displacement = 0.0;
      did = 1;
    }
    vol = volume();
    return dt;
  }
};

}  // namespace

extern "C" {

// Runs `nsteps` accepted steps of the SVI law on the dense float32 field `phi`
// (nx*ny*nz, C order, updated in place).  w_ball: the (2m+1)^3 ball cell weights
// (FP64, float32 in the sums); records[4 i .. 4 i + 3] = dt, volume, vmax (0),
// reinitialised of step i.  Returns the OpenMP thread count used, or a
// negative value for bad arguments.
int svicpu_run(float* phi, int nx, int ny, int nz, double h, double width, double radius, int k, int s, double t,
               double eps, double n_cap, const double* w_ball, int m, double cfl, double feedback_cfl,
               double reinit_distance, int reinit_iterations, double cut_lo, double cut_hi, int nsteps,
               double* records) {
  if (!phi || !w_ball || !records || nx < 16 || ny < 16 || nz < 16 || m < 1 || nsteps < 0) return -1;
  Sim S;
  S.g = Grid{nx, ny, nz, (int64_t)ny * nz, (int64_t)nz, (int64_t)nx * ny * nz};
  S.h = (float)h;
  S.width = (float)width;
  S.cut_lo = (float)(cut_lo * h);
  S.cut_hi = (float)(cut_hi * h);
  S.radius = radius;
  S.h3_over_bv = h * h * h / (4.0 / 3.0 * M_PI * radius * radius * radius);
  S.law = Law{k, s, (float)t, (float)eps, (float)n_cap};
  S.cfl = cfl;
  S.feedback_cfl = feedback_cfl;
  S.reinit_distance = reinit_distance;
  S.reinit_iterations = reinit_iterations;
  S.phi.assign(phi, phi + S.g.n);
  S.pm = m;
  S.psy = nz + 2 * m;
  S.psx = (int64_t)(ny + 2 * m) * S.psy;
  S.occ.assign((size_t)(nx + 2 * m) * S.psx, 0.0f);
  S.nr.assign(S.g.n, 0.0f);
  S.mark.assign(S.g.n, 0);
  const int K = 2 * m + 1;
  for (int a = 0; a < K; ++a)
    for (int b = 0; b < K; ++b)
      for (int c = 0; c < K; ++c) {
        const float w = (float)w_ball[((int64_t)a * K + b) * K + c];
        if (w != 0.0f) {
          S.tap_off.push_back((a - m) * S.psx + (b - m) * S.psy + (c - m));
          S.tap_w.push_back(w);
        }
      }
  S.rebuild_band();
  S.u0.resize(S.band.size());
  S.L.resize(S.band.size());
  S.V.resize(S.band.size());
  for (int i = 0; i < nsteps; ++i) {
    double vol;
    int did;
    const double dt = S.step(vol, did);
    records[4 * i] = dt;
    records[4 * i + 1] = vol;
    records[4 * i + 2] = 0.0;
    records[4 * i + 3] = did;
  }
  std::memcpy(phi, S.phi.data(), sizeof(float) * S.g.n);
  return omp_get_max_threads();
}

// Reinitialisation alone (band of |phi| < width h, then svierosion/levelset.py:
// 154-183 with `iterations`) of the field `phi`, in place.  Returns the new band size.
int64_t svicpu_reinit(float* phi, int nx, int ny, int nz, double h, double width, int iterations) {
  if (!phi || nx < 16 || ny < 16 || nz < 16 || iterations < 0) return -1;
  Sim S;
  S.g = Grid{nx, ny, nz, (int64_t)ny * nz, (int64_t)nz, (int64_t)nx * ny * nz};
  S.h = (float)h;
  S.width = (float)width;
  S.reinit_iterations = iterations;
  S.phi.assign(phi, phi + S.g.n);
  S.rebuild_band();
  S.reinitialize();
  std::memcpy(phi, S.phi.data(), sizeof(float) * S.g.n);
  return (int64_t)S.band.size();
}

}  // extern "C"
