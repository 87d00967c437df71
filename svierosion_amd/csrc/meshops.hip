// Host-side mesh measures of the API layer (TriMesh.signed_volume /
// edge_topology, svierosion/meshio.py:80-105): plain C++ loops over the face
// array instead of NumPy temporaries.  They run once per job and per output
// mesh (1.3M triangles: ~10 ms each instead of 0.13 s), need no GPU and so
// also serve the host-only tests.
#include <algorithm>
#include <vector>

#include "common.h"

using namespace svimx;

extern "C" {

int svimx_mesh_signed_volume(const double* vertices, int64_t nv, const int64_t* faces, int64_t nf, double* volume) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(volume && nv >= 0 && nf >= 0 && (vertices || nv == 0) && (faces || nf == 0), SVIMX_EBADARG,
                "bad mesh arguments");
  // about the vertex centroid, as the reference does, so large offsets cancel
  double c[3] = {0.0, 0.0, 0.0};
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) c[a] += vertices[3 * i + a];
  if (nv > 0)
    for (int a = 0; a < 3; ++a) c[a] /= (double)nv;
  double sum = 0.0;
  for (int64_t t = 0; t < nf; ++t) {
    double p[3][3];
    for (int q = 0; q < 3; ++q) {
      const int64_t v = faces[3 * t + q];
      SVIMX_REQUIRE(v >= 0 && v < nv, SVIMX_EBADARG, "face %lld names vertex %lld of %lld", (long long)t,
                    (long long)v, (long long)nv);
      for (int a = 0; a < 3; ++a) p[q][a] = vertices[3 * v + a] - c[a];
    }
    // p0 . (p1 x p2)
    sum += p[0][0] * (p[1][1] * p[2][2] - p[1][2] * p[2][1]) + p[0][1] * (p[1][2] * p[2][0] - p[1][0] * p[2][2]) +
           p[0][2] * (p[1][0] * p[2][1] - p[1][1] * p[2][0]);
  }
  *volume = sum / 6.0;
  return SVIMX_OK;
  SVIMX_TRY_END
}

// out[0] edges, out[1] boundary edges (one face), out[2] non-manifold edges
// (more than two faces), out[3] misoriented edges (two faces traversing the
// edge in the same direction).  Edges are grouped by their lower vertex (a
// counting sort), then by the upper one within each small group.
int svimx_mesh_edge_topology(const int64_t* faces, int64_t nf, int64_t nv, int64_t* out) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(out && nf >= 0 && nv >= 0 && (faces || nf == 0), SVIMX_EBADARG, "bad mesh arguments");
  for (int k = 0; k < 4; ++k) out[k] = 0;
  const int64_t ne = 3 * nf;
  std::vector<int64_t> start((size_t)nv + 1, 0);
  for (int64_t i = 0; i < ne; ++i) {
    const int64_t a = faces[(i / 3) * 3 + i % 3], b = faces[(i / 3) * 3 + (i % 3 + 1) % 3];
    SVIMX_REQUIRE(a >= 0 && a < nv && b >= 0 && b < nv, SVIMX_EBADARG, "face %lld names a vertex outside [0, %lld)",
                  (long long)(i / 3), (long long)nv);
    ++start[(size_t)std::min(a, b) + 1];
  }
  for (int64_t v = 0; v < nv; ++v) start[(size_t)v + 1] += start[(size_t)v];
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  std::vector<int64_t> upper((size_t)ne);
  std::vector<int8_t> sign((size_t)ne);
  for (int64_t i = 0; i < ne; ++i) {
    const int64_t a = faces[(i / 3) * 3 + i % 3], b = faces[(i / 3) * 3 + (i % 3 + 1) % 3];
    const int64_t lo = std::min(a, b);
    const int64_t slot = fill[(size_t)lo]++;
    upper[(size_t)slot] = std::max(a, b);
    sign[(size_t)slot] = a < b ? 1 : -1;
  }
  std::vector<std::pair<int64_t, int8_t>> group;
  for (int64_t v = 0; v < nv; ++v) {
    const int64_t s0 = start[(size_t)v], s1 = start[(size_t)v + 1];
    if (s0 == s1) continue;
    group.clear();
    for (int64_t s = s0; s < s1; ++s) group.emplace_back(upper[(size_t)s], sign[(size_t)s]);
    std::sort(group.begin(), group.end(),
              [](const std::pair<int64_t, int8_t>& x, const std::pair<int64_t, int8_t>& y) { return x.first < y.first; });
    for (size_t i = 0; i < group.size();) {
      size_t j = i;
      int64_t balance = 0;
      while (j < group.size() && group[j].first == group[i].first) balance += group[j++].second;
      const int64_t count = (int64_t)(j - i);
      ++out[0];
      if (count == 1) ++out[1];
      if (count > 2) ++out[2];
      if (count == 2 && balance != 0) ++out[3];
      i = j;
    }
  }
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"

// ---------------------------------------------------------------- SDF host preparation
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace {
struct PointKey {
  uint64_t b[3];
  bool operator==(const PointKey& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
};
struct PointKeyHash {
  size_t operator()(const PointKey& k) const {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (int a = 0; a < 3; ++a) {
      h ^= k.b[a] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
      h *= 0xBF58476D1CE4E5B9ull;
    }
    return (size_t)(h ^ (h >> 31));
  }
};
}  // namespace

extern "C" {

// The per-triangle arrays of the SDF build for a mesh none of whose
// triangles needs splitting (every edge <= max_edge cells); see
// include/svimx.h.  With a long edge present it writes nothing but *nsplit.
int svimx_sdf_prepare(const double* vertices, int64_t nv, const int64_t* faces, int64_t nf, const double* origin,
                      double h, const int32_t* shape, double exact_band, double max_edge, double* tri_idx,
                      int32_t* org, float* local, int32_t* box, int32_t* vid, int32_t* star, int32_t* deg,
                      int32_t* first, int64_t* nvid, int64_t* nsplit) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(vertices && faces && origin && shape && tri_idx && org && local && box && vid && star && deg && first &&
                    nvid && nsplit && nv >= 0 && nf >= 0 && h > 0,
                SVIMX_EBADARG, "bad SDF preparation arguments");
  SVIMX_REQUIRE(nv < (int64_t)INT32_MAX && 3 * nf < (int64_t)INT32_MAX, SVIMX_EBADARG, "mesh too large");
  *nsplit = 0;
  *nvid = 0;
  // index coordinates and the longest edge of every face
  int64_t split = 0;
  for (int64_t t = 0; t < nf; ++t) {
    double* p = tri_idx + 9 * t;
    for (int q = 0; q < 3; ++q) {
      const int64_t v = faces[3 * t + q];
      SVIMX_REQUIRE(v >= 0 && v < nv, SVIMX_EBADARG, "face %lld names vertex %lld of %lld", (long long)t,
                    (long long)v, (long long)nv);
      for (int a = 0; a < 3; ++a) p[3 * q + a] = (vertices[3 * v + a] - origin[a]) / h;
    }
    double emax = 0.0;
    for (int q = 0; q < 3; ++q) {
      const double* u = p + 3 * q;
      const double* w = p + 3 * ((q + 1) % 3);
      const double d = std::sqrt((w[0] - u[0]) * (w[0] - u[0]) + (w[1] - u[1]) * (w[1] - u[1]) +
                                 (w[2] - u[2]) * (w[2] - u[2]));
      emax = std::max(emax, d);
    }
    if (!(emax <= max_edge)) ++split;
  }
  if (split) {
    *nsplit = split;
    return SVIMX_OK;
  }
  // local frames and node boxes
  for (int64_t t = 0; t < nf; ++t) {
    const double* p = tri_idx + 9 * t;
    int64_t o[3];
    for (int a = 0; a < 3; ++a) {
      o[a] = (int64_t)std::floor(p[a]);
      org[3 * t + a] = (int32_t)o[a];
    }
    for (int q = 0; q < 3; ++q)
      for (int a = 0; a < 3; ++a) local[9 * t + 3 * q + a] = (float)(p[3 * q + a] - (double)o[a]);
    for (int a = 0; a < 3; ++a) {
      const double mn = std::min(p[a], std::min(p[3 + a], p[6 + a]));
      const double mx = std::max(p[a], std::max(p[3 + a], p[6 + a]));
      const int64_t lo = std::min<int64_t>(std::max<int64_t>((int64_t)std::ceil(mn - exact_band), 0), shape[a] - 1);
      const int64_t hi = std::min<int64_t>(std::max<int64_t>((int64_t)std::floor(mx + exact_band), 0), shape[a] - 1);
      box[6 * t + a] = (int32_t)lo;
      box[6 * t + 3 + a] = (int32_t)hi;
    }
  }
  // ids shared by exactly coincident vertices (bitwise equal coordinates), in
  // first-appearance order: they only group the corners of a vertex star
  std::unordered_map<PointKey, int32_t, PointKeyHash> ids;
  ids.reserve((size_t)nv * 2);
  std::vector<int32_t> pid((size_t)nv);
  for (int64_t v = 0; v < nv; ++v) {
    PointKey k;
    std::memcpy(k.b, vertices + 3 * v, sizeof(k.b));
    auto it = ids.emplace(k, (int32_t)ids.size()).first;
    pid[(size_t)v] = it->second;
  }
  const int64_t nid = (int64_t)ids.size();
  for (int64_t c = 0; c < 3 * nf; ++c) vid[c] = pid[(size_t)faces[c]];
  // star CSR: the corners of every id in corner order (a counting sort)
  for (int64_t v = 0; v < nid; ++v) deg[v] = 0;
  for (int64_t c = 0; c < 3 * nf; ++c) ++deg[vid[c]];
  int64_t run = 0;
  for (int64_t v = 0; v < nid; ++v) {
    first[v] = (int32_t)run;
    run += deg[v];
  }
  std::vector<int32_t> fill(first, first + nid);
  for (int64_t c = 0; c < 3 * nf; ++c) star[fill[(size_t)vid[c]]++] = (int32_t)(c / 3);
  *nvid = nid;
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
