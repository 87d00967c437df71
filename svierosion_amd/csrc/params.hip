// Host-side set-up of the API layer: grid sizing and speed-law parameters.
//
// Grid sizing (svierosion/grid.py:18-56): the grid holds the box [lo, hi]
// centred, with at least `padding` on every side, each dimension rounded up
// to a 7-smooth size.  The node positions define every compared field, so the
// FP64 operations are those of the reference, in its order.
//
// Speed laws (svierosion/flows.py:29-213): parameter checks, and for the
// geometric SVI form (s = 1) the cap n_cap with V(n_cap) = ratio * V(1/2),
// the largest double below the root of that equation (svierosion/flows.py:71-80
// defines it as the limit of a bisection on [1/2, t]).  Here the root is found
// by a bracketed Newton iteration, and the bracket is then closed to two
// adjacent doubles.
//
// No GPU is touched: these also serve the host-only tests.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

#include "common.h"

using namespace svimx;

namespace {

// Smallest 2^a 3^b 5^c 7^d >= n.  Walks the exponent lattice: for every
// prefix product below n, the next factor is tried, and each chain stops at
// its first product >= n (a minimal smooth m >= n has every proper prefix
// product below n, so it is among the chain ends).
int64_t smooth7_at_least(int64_t n) {
  if (n <= 1) return 1;
  int64_t best = std::numeric_limits<int64_t>::max();
  for (int64_t p2 = 1;; p2 *= 2) {
    for (int64_t p3 = p2;; p3 *= 3) {
      for (int64_t p5 = p3;; p5 *= 5) {
        int64_t p7 = p5;
        while (p7 < n) p7 *= 7;
        best = std::min(best, p7);
        if (p5 >= n) break;
      }
      if (p3 >= n) break;
    }
    if (p2 >= n) break;
  }
  return best;
}

// Raw SVI speed n^k / (t - n)^s + eps in FP64 (the host-side law).
double svi_raw(const svimx_law_t& L, double n) {
  double v = std::pow(n, (double)L.k);
  if (L.s) v /= (L.t - n);
  return v + L.eps;
}

// n_cap for s = 1: the largest double x in [1/2, t) with raw(x) < goal.
// raw is increasing on (0, t), so the bracket [lo, hi] with raw(lo) < goal <=
// raw(hi) stays valid; Newton steps that leave it are replaced by midpoints.
double solve_cap(const svimx_law_t& L, double goal) {
  double lo = 0.5, hi = L.t;
  if (!(svi_raw(L, lo) < goal)) return lo;
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    const double f = svi_raw(L, x) - goal;
    if (f < 0.0)
      lo = x;
    else
      hi = x;
    if (std::nextafter(lo, hi) >= hi) break;
    // d/dx [x^k / (t - x)] = (x^k / (t - x)) (k / x + 1 / (t - x))
    const double core = svi_raw(L, x) - L.eps;
    const double slope = core * ((double)L.k / x + 1.0 / (L.t - x));
    double nx = (slope > 0.0 && std::isfinite(slope)) ? x - f / slope : 0.5 * (lo + hi);
    if (!(nx > lo && nx < hi)) nx = 0.5 * (lo + hi);
    // once Newton has converged to within a few ulps, close the bracket by halving
    if (hi - lo < 64.0 * std::numeric_limits<double>::epsilon() * hi) nx = 0.5 * (lo + hi);
    if (nx == x) nx = 0.5 * (lo + hi);
    x = nx;
  }
  return lo;
}

}  // namespace

extern "C" {

int svimx_smooth_size(int64_t n, int64_t* out) {
  SVIMX_REQUIRE(out, SVIMX_EBADARG, "svimx_smooth_size: output pointer is NULL");
  SVIMX_REQUIRE(n < ((int64_t)1 << 56), SVIMX_EBADARG, "svimx_smooth_size: %lld is too large", (long long)n);
  *out = smooth7_at_least(n);
  return SVIMX_OK;
}

int svimx_grid_around(const double* lo, const double* hi, double spacing, double padding, int fast,
                      int64_t* shape, double* origin) {
  SVIMX_REQUIRE(lo && hi && shape && origin, SVIMX_EBADARG, "svimx_grid_around: NULL argument");
  SVIMX_REQUIRE(spacing > 0.0 && std::isfinite(spacing), SVIMX_EBADARG,
                "grid spacing must be a positive finite number (got %g)", spacing);
  SVIMX_REQUIRE(padding >= 0.0 && std::isfinite(padding), SVIMX_EBADARG, "grid padding must be >= 0 (got %g)",
                padding);
  for (int a = 0; a < 3; ++a) {
    SVIMX_REQUIRE(std::isfinite(lo[a]) && std::isfinite(hi[a]) && hi[a] >= lo[a], SVIMX_EBADARG,
                  "grid box [%g, %g] on axis %d is empty or not finite", lo[a], hi[a], a);
    const double cells = std::ceil((hi[a] - lo[a] + 2.0 * padding) / spacing);
    SVIMX_REQUIRE(cells < 1e9, SVIMX_EBADARG, "grid axis %d would need %g nodes", a, cells);
    const int64_t need = (int64_t)cells + 1;
    shape[a] = fast ? smooth7_at_least(need) : need;
  }
  for (int a = 0; a < 3; ++a) {
    const double centre = 0.5 * (lo[a] + hi[a]);
    origin[a] = centre - 0.5 * (double)(shape[a] - 1) * spacing;
  }
  return SVIMX_OK;
}

int svimx_law_init(int kind, double radius, double k, double s, double t, double eps, double max_speed_ratio,
                   double alpha, double delta, svimx_law_t* out) {
  SVIMX_REQUIRE(out, SVIMX_EBADARG, "svimx_law_init: output pointer is NULL");
  SVIMX_REQUIRE(kind >= 0 && kind <= 2, SVIMX_EBADARG, "law kind %d is not 0 (SVI), 1 (mean) or 2 (Gauss)", kind);
  SVIMX_REQUIRE(eps >= 0.0, SVIMX_EBADARG, "the additive speed eps must be non-negative (got %g)", eps);
  svimx_law_t L{};
  L.kind = kind;
  L.eps = eps;
  L.n_cap = 1.0;
  L.alpha = 1.0;
  if (kind == 0) {
    SVIMX_REQUIRE(radius > 0.0, SVIMX_EBADARG, "the SVI ball radius must be positive (got %g)", radius);
    SVIMX_REQUIRE(k >= 0.0 && k == std::floor(k) && k < 64.0, SVIMX_EBADARG,
                  "the SVI exponent k must be a non-negative integer (got %g)", k);
    SVIMX_REQUIRE(s == 0.0 || s == 1.0, SVIMX_EBADARG, "the SVI switch s must be 0 or 1 (got %g)", s);
    L.k = (int32_t)k;
    L.s = (int32_t)s;
    if (L.s == 1) {
      SVIMX_REQUIRE(t > 0.5 && t < 1.0, SVIMX_EBADARG,
                    "with s = 1 the singular value t must lie strictly between 1/2 and 1, above the flat-surface "
                    "value N = 1/2 (got t = %g)", t);
      L.t = t;
      L.n_cap = solve_cap(L, max_speed_ratio * svi_raw(L, 0.5));
    } else {
      L.t = std::isfinite(t) ? t : 0.0;
    }
  } else {
    SVIMX_REQUIRE(alpha > 0.0, SVIMX_EBADARG, "the curvature exponent alpha must be positive (got %g)", alpha);
    SVIMX_REQUIRE(delta >= 0.0, SVIMX_EBADARG, "the regularisation delta must be non-negative (got %g)", delta);
    SVIMX_REQUIRE(alpha >= 1.0 || delta > 0.0, SVIMX_EBADARG,
                  "alpha = %g < 1 requires a regularisation delta > 0: the unregularised power has an unbounded "
                  "derivative at zero curvature, and the parabolic time step limit would vanish", alpha);
    L.alpha = alpha;
    L.delta = delta;
  }
  *out = L;
  return SVIMX_OK;
}

int svimx_law_raw(const svimx_law_t* law, double n, double* out) {
  SVIMX_REQUIRE(law && out && law->kind == 0, SVIMX_EBADARG, "svimx_law_raw needs an SVI law");
  *out = svi_raw(*law, n);
  return SVIMX_OK;
}

}  // extern "C"
