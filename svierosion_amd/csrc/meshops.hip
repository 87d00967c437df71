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
