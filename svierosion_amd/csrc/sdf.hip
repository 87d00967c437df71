// Signed distance of a triangle mesh on the grid (svierosion/sdf.py:365-415).
//
// 1. Exact near field, triangle-bucketed: the grid is cut into 8^3-node
//    bricks, every (split) triangle is listed in each brick its box (dilated
//    by exact_band cells) overlaps, and one workgroup per brick takes the
//    brick's triangles, each thread a triangle visiting the nodes of its box
//    inside the brick, min-reducing a packed key (float bits of the distance
//    << 32 | triangle id) in the brick's 512 LDS keys; the brick then writes
//    its keys with plain stores.  The minimum is order independent, so the
//    field is deterministic and equals a global atomic reduction over the
//    same (node, triangle) pairs (svierosion/sdf.py:100-145).
// 2. Ring propagation: rings of nodes around the known set test the closest
//    triangles of their 26 neighbours, Jacobi per ring (svierosion/sdf.py:163-202).
// 3. Vertex-star refinement of nodes with exact_band < d <= band + 1 (:211-265).
// 4. Sign by scan-line ray casting along grid lines with an FP64
//    Simulation-of-Simplicity point-in-triangle test, winding or parity count,
//    majority over the requested axes (:269-361).
//
// Distances are computed in a frame local to each triangle (vertices stored
// relative to an integer node), so float32 keeps ~1e-7 cells instead of the
// ~1e-5 cells of float32 global index coordinates at index ~400.
#include <vector>

#include "common.h"
#include "compact.h"

namespace svimx {

struct TriSoup {
  const float* local;   // 9 per triangle, index units, relative to org
  const int32_t* org;   // 3 per triangle
};

__device__ __forceinline__ float tri_distance(const TriSoup& T, int64_t t, int i, int j, int k) {
  const float* v = T.local + 9 * t;
  const int32_t* o = T.org + 3 * t;
  float p[3] = {(float)(i - o[0]), (float)(j - o[1]), (float)(k - o[2])};
  return sqrtf(point_triangle_d2(p, v, v + 3, v + 6));
}

__device__ __forceinline__ uint64_t pack_key(float d, int64_t t) {
  return ((uint64_t)(uint32_t)__float_as_int(d) << 32) | (uint64_t)(uint32_t)t;
}

__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u8s(uint8_t* p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- near field
constexpr int kSdfBrick = 8;  // nodes per brick edge (512 keys of 8 B in LDS)

struct BrickGrid {
  int bx, by, bz;  // bricks per axis
  __host__ __device__ int64_t count() const { return (int64_t)bx * by * bz; }
};

// Bricks overlapped by each triangle's box: counts[brick] += 1.
__global__ void k_brick_count(const int32_t* __restrict__ box, int64_t ntri, BrickGrid B,
                              unsigned long long* __restrict__ counts) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  const int32_t* b = box + 6 * t;
  for (int x = b[0] / kSdfBrick; x <= b[3] / kSdfBrick; ++x)
    for (int y = b[1] / kSdfBrick; y <= b[4] / kSdfBrick; ++y)
      for (int z = b[2] / kSdfBrick; z <= b[5] / kSdfBrick; ++z)
        atomicAdd(&counts[((int64_t)x * B.by + y) * B.bz + z], 1ull);
}

// The triangle lists of the bricks, at the scanned offsets (cursor = a copy
// of the offsets, advanced atomically; the order within a brick is free).
__global__ void k_brick_fill(const int32_t* __restrict__ box, int64_t ntri, BrickGrid B,
                             unsigned long long* __restrict__ cursor, int32_t* __restrict__ list) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  const int32_t* b = box + 6 * t;
  for (int x = b[0] / kSdfBrick; x <= b[3] / kSdfBrick; ++x)
    for (int y = b[1] / kSdfBrick; y <= b[4] / kSdfBrick; ++y)
      for (int z = b[2] / kSdfBrick; z <= b[5] / kSdfBrick; ++z)
        list[atomicAdd(&cursor[((int64_t)x * B.by + y) * B.bz + z], 1ull)] = (int32_t)t;
}

// One workgroup per brick: its triangles' (node, triangle) pairs inside the
// brick, min-reduced in LDS, then one store per node of the brick.
__global__ void __launch_bounds__(256) k_near_bricks(TriSoup T, const int32_t* __restrict__ box,
                                                     const int64_t* __restrict__ offs,
                                                     const int32_t* __restrict__ list, BrickGrid B, Dims d,
                                                     float eb, unsigned long long* __restrict__ keys) {
  constexpr int N3 = kSdfBrick * kSdfBrick * kSdfBrick;
  __shared__ unsigned long long lk[N3];
  const int64_t brick = blockIdx.x;
  const int64_t b0 = offs[brick], b1 = offs[brick + 1];
  if (b0 == b1) return;  // block-uniform: no triangle reaches this brick
  const int bz = (int)(brick % B.bz), by = (int)((brick / B.bz) % B.by), bx = (int)(brick / ((int64_t)B.bz * B.by));
  const int x0 = bx * kSdfBrick, y0 = by * kSdfBrick, z0 = bz * kSdfBrick;
  for (int e = threadIdx.x; e < N3; e += blockDim.x) lk[e] = 0x7FFFFFFFFFFFFFFFull;
  __syncthreads();
  for (int64_t q = b0 + threadIdx.x; q < b1; q += blockDim.x) {
    const int32_t t = list[q];
    const int32_t* bb = box + 6 * (int64_t)t;
    const int ilo = max(bb[0], x0), ihi = min(bb[3], x0 + kSdfBrick - 1);
    const int jlo = max(bb[1], y0), jhi = min(bb[4], y0 + kSdfBrick - 1);
    const int klo = max(bb[2], z0), khi = min(bb[5], z0 + kSdfBrick - 1);
    for (int i = ilo; i <= ihi; ++i)
      for (int j = jlo; j <= jhi; ++j)
        for (int k = klo; k <= khi; ++k) {
          const float dist = tri_distance(T, t, i, j, k);
          if (dist <= eb)
            atomicMin(&lk[((i - x0) * kSdfBrick + (j - y0)) * kSdfBrick + (k - z0)],
                      (unsigned long long)pack_key(dist, t));
        }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < N3; e += blockDim.x) {
    const int i = x0 + e / (kSdfBrick * kSdfBrick), j = y0 + (e / kSdfBrick) % kSdfBrick, k = z0 + e % kSdfBrick;
    if (i < d.nx && j < d.ny && k < d.nz) keys[((int64_t)i * d.ny + j) * d.nz + k] = lk[e];
  }
}

__global__ void k_unpack(const uint64_t* __restrict__ keys, int64_t n, float* __restrict__ dist,
                         int32_t* __restrict__ tid, uint8_t* __restrict__ known) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  uint64_t kk = keys[e];
  bool found = kk != 0x7FFFFFFFFFFFFFFFull;
  dist[e] = found ? __int_as_float((int)(uint32_t)(kk >> 32)) : INFINITY;
  tid[e] = found ? (int32_t)(uint32_t)(kk & 0xFFFFFFFFull) : -1;
  known[e] = found ? 1 : 0;
}

// Ring = 26-neighbourhood dilation of `known` minus `known`, never on the outer layer.
__global__ void k_ring_mask(const uint8_t* __restrict__ known, uint8_t* __restrict__ ring, Dims d) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.size()) return;
  int i, j, k;
  unflat(d, e, i, j, k);
  uint8_t r = 0;
  if (!known[e] && i > 0 && j > 0 && k > 0 && i < d.nx - 1 && j < d.ny - 1 && k < d.nz - 1) {
    for (int a = -1; a <= 1 && !r; ++a)
      for (int b = -1; b <= 1 && !r; ++b)
        for (int c = -1; c <= 1 && !r; ++c) r = known[e + a * d.sx() + b * d.sy() + c];
  }
  ring[e] = r;
}

struct U8Pred {
  const uint8_t* m;
  __device__ bool operator()(int64_t e) const { return m[e] != 0; }
};

// Candidates from the 26 neighbours' closest triangles; results to (nd, nt) by list position.
__global__ void k_propagate(TriSoup T, const int32_t* __restrict__ upd, int64_t nu, Dims d,
                            const float* __restrict__ dist, const int32_t* __restrict__ tid, float* __restrict__ nd,
                            int32_t* __restrict__ nt) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nu) return;
  const int64_t u = upd[q];
  const int64_t n = d.size();
  int i, j, k;
  unflat(d, u, i, j, k);
  float best = dist[u];
  int32_t bt = tid[u];
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        if (a == 0 && b == 0 && c == 0) continue;
        int64_t nb = u + a * d.sx() + b * d.sy() + c;
        nb = nb < 0 ? 0 : (nb > n - 1 ? n - 1 : nb);
        int32_t t = tid[nb];
        if (t < 0) continue;
        float dd = tri_distance(T, t, i, j, k);
        if (dd < best) {
          best = dd;
          bt = t;
        }
      }
  nd[q] = best;
  nt[q] = bt;
}

__global__ void k_apply_list(const int32_t* __restrict__ upd, int64_t nu, const float* __restrict__ nd,
                             const int32_t* __restrict__ nt, float* __restrict__ dist, int32_t* __restrict__ tid,
                             uint8_t* __restrict__ known) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nu) return;
  const int64_t u = upd[q];
  dist[u] = nd[q];
  tid[u] = nt[q];
  known[u] = 1;
}

struct StarPred {
  const float* dist;
  float lo, hi;
  __device__ bool operator()(int64_t e) const {
    float v = dist[e];
    return v > lo && v <= hi;
  }
};

// Re-test every triangle incident to a vertex of the node's closest triangle.
__global__ void k_star_refine(TriSoup T, const int32_t* __restrict__ nodes, int64_t nn, Dims d,
                              const int32_t* __restrict__ vid, const int32_t* __restrict__ first,
                              const int32_t* __restrict__ deg, const int32_t* __restrict__ star, float* dist,
                              int32_t* tid, int32_t* changed) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nn) return;
  const int64_t u = nodes[q];
  const int32_t t0 = tid[u];
  if (t0 < 0) return;
  int i, j, k;
  unflat(d, u, i, j, k);
  uint64_t best = ~0ull;
  for (int s = 0; s < 3; ++s) {
    int32_t v = vid[3 * (int64_t)t0 + s];
    int32_t f0 = first[v], cnt = deg[v];
    for (int c = 0; c < cnt; ++c) {
      int32_t cand = star[f0 + c];
      uint64_t key = pack_key(tri_distance(T, cand, i, j, k), cand);
      best = key < best ? key : best;
    }
  }
  float bd = __int_as_float((int)(uint32_t)(best >> 32));
  if (bd < dist[u]) {
    dist[u] = bd;
    tid[u] = (int32_t)(uint32_t)(best & 0xFFFFFFFFull);
    atomicAdd(changed, 1);
  }
}

// ---------------------------------------------------------------- sign
// Orientation of (0,0),(x1,y1),(x2,y2) in FP64 with the SoS tie-break of
// svierosion/sdf.py:269-282.  Contraction is off so the products round as in
// the reference's separate multiplications.
__device__ int orient_sos(double x1, double y1, double x2, double y2, double& area) {
#pragma clang fp contract(off)
  double a = x1 * y2 - y1 * x2;
  area = a;
  if (a > 0.0) return 1;
  if (a < 0.0) return -1;
  if (y2 > y1) return -1;
  if (y2 < y1) return 1;
  if (x1 > x2) return -1;
  if (x1 < x2) return 1;
  return 0;
}

// One thread per original face: crossings of the grid lines parallel to `axis`
// that pierce it, accumulated at floor(z)+1 (signed: -orientation; else +1).
__global__ void k_sign_cross(const double* __restrict__ tri, int64_t nf, Dims d, int axis, int signed_rule,
                             int32_t* __restrict__ acc) {
#pragma clang fp contract(off)
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int b_ax = (axis + 1) % 3, c_ax = (axis + 2) % 3;
  const int n[3] = {d.nx, d.ny, d.nz};
  const int na = n[axis], nb = n[b_ax], nc = n[c_ax];
  const double* v = tri + 9 * f;
  double U[3], W[3], Z[3];
  for (int q = 0; q < 3; ++q) {
    U[q] = v[3 * q + b_ax];
    W[q] = v[3 * q + c_ax];
    Z[q] = v[3 * q + axis];
  }
  double umin = fmin(fmin(U[0], U[1]), U[2]), umax = fmax(fmax(U[0], U[1]), U[2]);
  double wmin = fmin(fmin(W[0], W[1]), W[2]), wmax = fmax(fmax(W[0], W[1]), W[2]);
  int64_t j0 = (int64_t)fmin(fmax(ceil(umin), 0.0), (double)nb);
  int64_t j1 = (int64_t)fmin(fmax(floor(umax), -1.0), (double)(nb - 1));
  int64_t k0 = (int64_t)fmin(fmax(ceil(wmin), 0.0), (double)nc);
  int64_t k1 = (int64_t)fmin(fmax(floor(wmax), -1.0), (double)(nc - 1));
  for (int64_t jj = j0; jj <= j1; ++jj)
    for (int64_t kk = k0; kk <= k1; ++kk) {
      double x1 = U[0] - (double)jj, x2 = U[1] - (double)jj, x3 = U[2] - (double)jj;
      double y1 = W[0] - (double)kk, y2 = W[1] - (double)kk, y3 = W[2] - (double)kk;
      double A, B, Cc;
      int sa = orient_sos(x2, y2, x3, y3, A);
      int sb = orient_sos(x3, y3, x1, y1, B);
      int sc = orient_sos(x1, y1, x2, y2, Cc);
      if (sa == 0 || sb != sa || sc != sa) continue;
      double S = A + B + Cc;
      if (S == 0.0) continue;
      double z = (A * Z[0] + B * Z[1] + Cc * Z[2]) / S;
      int64_t i0 = (int64_t)floor(z) + 1;
      i0 = i0 < 0 ? 0 : (i0 > na ? na : i0);
      atomicAdd(&acc[(i0 * nb + jj) * nc + kk], signed_rule ? -sa : 1);
    }
}

// Running sum along each line; votes[node] += inside.
__global__ void k_sign_scan(const int32_t* __restrict__ acc, Dims d, int axis, int signed_rule,
                            uint8_t* __restrict__ votes) {
  const int b_ax = (axis + 1) % 3, c_ax = (axis + 2) % 3;
  const int n[3] = {d.nx, d.ny, d.nz};
  const int na = n[axis], nb = n[b_ax], nc = n[c_ax];
  int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= (int64_t)nb * nc) return;
  const int jj = (int)(line / nc), kk = (int)(line % nc);
  const int64_t st[3] = {d.sx(), d.sy(), 1};
  int64_t node = (int64_t)jj * st[b_ax] + (int64_t)kk * st[c_ax];
  int32_t run = 0;
  for (int i0 = 0; i0 < na; ++i0) {
    run += acc[((int64_t)i0 * nb + jj) * nc + kk];
    bool inside = signed_rule ? (run != 0) : ((run & 1) != 0);
    if (inside) votes[node] += 1;
    node += st[axis];
  }
}

__global__ void k_sdf_finish(const float* __restrict__ dist, const uint8_t* __restrict__ votes, int naxes, int64_t n,
                             float band, float h, float* __restrict__ phi) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float v = fminf(dist[e], band) * h;
  phi[e] = (votes[e] * 2 > naxes) ? -v : v;
}

template <class T>
struct DevBuf {
  // stream-ordered temporaries: no device-wide synchronisation on free, so
  // other jobs' kernels on other streams keep running
  T* p = nullptr;
  hipStream_t st = nullptr;
  ~DevBuf() {
    if (p) hipFreeAsync(p, st);
  }
  hipError_t alloc(int64_t n, hipStream_t s) {
    st = s;
    return hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(T) * (n > 0 ? n : 1), s);
  }
};

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_sdf_build(void* stream, const float* tri_local, const int32_t* tri_org, const int32_t* tri_box,
                    int64_t n_split, const int32_t* tri_vid, const int32_t* star_first, const int32_t* star_deg,
                    const int32_t* star_tris, int64_t n_vertices, const double* faces_idx, int64_t n_faces,
                    int sign_axes_mask, int signed_rule, int nx, int ny, int nz, double h, double band,
                    double exact_band, float* phi_out, float* dist_out, int32_t* tid_out, void* ws,
                    int64_t ws_bytes) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  const int64_t n = d.size();
  SVIMX_REQUIRE(n > 0 && n < (int64_t)INT32_MAX && phi_out, SVIMX_EBADARG, "bad SDF grid or output");
  SVIMX_REQUIRE(n_split >= 0 && n_split < (int64_t)INT32_MAX && n_faces >= 0, SVIMX_EBADARG, "bad triangle count");
  SVIMX_REQUIRE((sign_axes_mask & 7) != 0, SVIMX_EBADARG, "no sign axis");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, n, w)) return SVIMX_EBADARG;
  hipStream_t st = as_stream(stream);
  const dim3 tb(256), gn(blocks_for(n, 256));
  DevBuf<uint64_t> keys;
  DevBuf<float> dist_own, nd;
  DevBuf<int32_t> tid_own, list, prev, nt, acc, changed;
  DevBuf<uint8_t> known, ring;
  SVIMX_CHECK_HIP(keys.alloc(n, st));
  SVIMX_CHECK_HIP(known.alloc(n, st));
  SVIMX_CHECK_HIP(ring.alloc(n, st));
  SVIMX_CHECK_HIP(list.alloc(n, st));
  SVIMX_CHECK_HIP(prev.alloc(n, st));
  SVIMX_CHECK_HIP(nd.alloc(2 * n, st));
  SVIMX_CHECK_HIP(nt.alloc(2 * n, st));
  SVIMX_CHECK_HIP(changed.alloc(1, st));
  float* dist = dist_out;
  int32_t* tid = tid_out;
  if (!dist) {
    SVIMX_CHECK_HIP(dist_own.alloc(n, st));
    dist = dist_own.p;
  }
  if (!tid) {
    SVIMX_CHECK_HIP(tid_own.alloc(n, st));
    tid = tid_own.p;
  }
  TriSoup T{tri_local, tri_org};
  // 1. near field
  hipLaunchKernelGGL(k_fill_u64, gn, tb, 0, st, keys.p, n, (uint64_t)0x7FFFFFFFFFFFFFFFull);
  SVIMX_CHECK_LAUNCH();
  const float eb = (float)(exact_band * (1.0 + 1e-6));
  if (n_split > 0) {
    // triangles binned per brick (count, scan, fill), then one workgroup per brick
    const BrickGrid B{(nx + kSdfBrick - 1) / kSdfBrick, (ny + kSdfBrick - 1) / kSdfBrick,
                      (nz + kSdfBrick - 1) / kSdfBrick};
    const int64_t nbr = B.count();
    SVIMX_REQUIRE(nbr < (int64_t)INT32_MAX, SVIMX_EBADARG, "too many SDF bricks");
    DevBuf<int64_t> boffs, cursor;
    SVIMX_CHECK_HIP(boffs.alloc(nbr + 1, st));
    SVIMX_CHECK_HIP(cursor.alloc(nbr + 1, st));
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(nbr + 1, 256)), tb, 0, st,
                       reinterpret_cast<uint64_t*>(boffs.p), nbr + 1, (uint64_t)0);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_brick_count, dim3(blocks_for(n_split, 256)), tb, 0, st, tri_box, n_split, B,
                       reinterpret_cast<unsigned long long*>(boffs.p));
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, boffs.p, nbr);
    SVIMX_CHECK_LAUNCH();
    int64_t nlist = 0;
    SVIMX_CHECK_HIP(hipMemcpyAsync(&nlist, boffs.p + nbr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SVIMX_CHECK_HIP(hipMemcpyAsync(cursor.p, boffs.p, sizeof(int64_t) * (nbr + 1), hipMemcpyDeviceToDevice, st));
    SVIMX_CHECK_HIP(hipStreamSynchronize(st));
    DevBuf<int32_t> blist;
    SVIMX_CHECK_HIP(blist.alloc(std::max<int64_t>(nlist, 1), st));
    hipLaunchKernelGGL(k_brick_fill, dim3(blocks_for(n_split, 256)), tb, 0, st, tri_box, n_split, B,
                       reinterpret_cast<unsigned long long*>(cursor.p), blist.p);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_near_bricks, dim3((unsigned)nbr), tb, 0, st, T, tri_box, boffs.p, blist.p, B, d, eb,
                       (unsigned long long*)keys.p);
    SVIMX_CHECK_LAUNCH();
  }  // the brick buffers are freed stream-ordered at scope exit
  hipLaunchKernelGGL(k_unpack, gn, tb, 0, st, keys.p, n, dist, tid, known.p);
  SVIMX_CHECK_LAUNCH();
  // 2. ring propagation (Jacobi per ring over the new ring plus the previous one)
  const int rings = std::max((int)std::ceil(band - exact_band) + 1, 0);
  int64_t nprev = 0;
  for (int r = 0; r < rings; ++r) {
    hipLaunchKernelGGL(k_ring_mask, gn, tb, 0, st, known.p, ring.p, d);
    SVIMX_CHECK_LAUNCH();
    int64_t nr = 0;
    int rc = compact(st, U8Pred{ring.p}, n, list.p, n, w, &nr);
    if (rc) return rc;
    if (nr + nprev == 0) break;
    if (nprev > 0)
      SVIMX_CHECK_HIP(hipMemcpyAsync(list.p + nr, prev.p, sizeof(int32_t) * nprev, hipMemcpyDeviceToDevice, st));
    const int64_t nu = nr + nprev;
    hipLaunchKernelGGL(k_propagate, dim3(blocks_for(nu, 256)), tb, 0, st, T, list.p, nu, d, dist, tid, nd.p, nt.p);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_apply_list, dim3(blocks_for(nu, 256)), tb, 0, st, list.p, nu, nd.p, nt.p, dist, tid,
                       known.p);
    SVIMX_CHECK_LAUNCH();
    SVIMX_CHECK_HIP(hipMemcpyAsync(prev.p, list.p, sizeof(int32_t) * nr, hipMemcpyDeviceToDevice, st));
    nprev = nr;
  }
  // 3. vertex-star refinement
  if (n_split > 0 && n_vertices > 0 && star_first && star_deg && star_tris && tri_vid) {
    int64_t ns = 0;
    int rc = compact(st, StarPred{dist, (float)exact_band, (float)(band + 1.0)}, n, list.p, n, w, &ns);
    if (rc) return rc;
    for (int pass = 0; pass < 2 && ns > 0; ++pass) {
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, st, changed.p, (int64_t)1, 0);
      SVIMX_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_star_refine, dim3(blocks_for(ns, 256)), tb, 0, st, T, list.p, ns, d, tri_vid, star_first,
                         star_deg, star_tris, dist, tid, changed.p);
      SVIMX_CHECK_LAUNCH();
      int32_t ch = 0;
      SVIMX_CHECK_HIP(hipMemcpyAsync(&ch, changed.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      SVIMX_CHECK_HIP(hipStreamSynchronize(st));
      if (ch == 0) break;
    }
  }
  // 4. sign
  int naxes = 0;
  hipLaunchKernelGGL(k_fill_u8s, gn, tb, 0, st, ring.p, n, (uint8_t)0);  // ring reused as the vote count
  SVIMX_CHECK_LAUNCH();
  const int dimsv[3] = {nx, ny, nz};
  for (int axis = 0; axis < 3; ++axis) {
    if (!((sign_axes_mask >> axis) & 1)) continue;
    ++naxes;
    const int na = dimsv[axis], nb = dimsv[(axis + 1) % 3], nc = dimsv[(axis + 2) % 3];
    const int64_t nacc = (int64_t)(na + 1) * nb * nc;
    DevBuf<int32_t> accb;
    SVIMX_CHECK_HIP(accb.alloc(nacc, st));
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(nacc, 256)), tb, 0, st, accb.p, nacc, 0);
    SVIMX_CHECK_LAUNCH();
    if (n_faces > 0) {
      hipLaunchKernelGGL(k_sign_cross, dim3(blocks_for(n_faces, 64)), dim3(64), 0, st, faces_idx, n_faces, d, axis,
                         signed_rule, accb.p);
      SVIMX_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_sign_scan, dim3(blocks_for((int64_t)nb * nc, 256)), tb, 0, st, accb.p, d, axis,
                       signed_rule, ring.p);
    SVIMX_CHECK_LAUNCH();
    SVIMX_CHECK_HIP(hipStreamSynchronize(st));  // accb is freed at scope exit
  }
  hipLaunchKernelGGL(k_sdf_finish, gn, tb, 0, st, dist, ring.p, naxes, n, (float)band, (float)h, phi_out);
  SVIMX_CHECK_LAUNCH();
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
