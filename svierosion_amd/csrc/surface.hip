// Zero set of phi as a triangle mesh by marching tetrahedra on the device.
//
// The reference meshes the band with Lewiner marching cubes on the host
// (svierosion/surface.py:16-56).  Here every cell whose lower corner is a band
// node is split into the 6 Kuhn tetrahedra (corner 0 -> 7 along each axis
// permutation), a split that is conforming across cells, so the surface is
// watertight by construction and has no ambiguous cases.  Inside = phi < 0.
//
// Vertices live on tetrahedron edges.  Every edge of the Kuhn split is owned
// by its lower end node and one of 7 directions (+x, +y, +z, +x+y, +x+z, +y+z,
// +x+y+z); a vertex on that edge is bit `dir` of the owner's 8-bit slot mask.
// A vertex that falls exactly on a node (phi == 0 there) is bit 7 of that
// node's mask, so coincident vertices are shared.  Vertex ids are the
// exclusive scan of the mask popcounts over band positions plus the rank of
// the bit inside the mask, so ids follow (band position, direction) order and
// do not depend on scheduling.  Triangles with a repeated vertex are dropped.
// Orientation is combinatorial: a positively oriented tetrahedron listed as
// (inside..., outside...) gives outward triangles in canonical order, an odd
// arrangement flips them.
//
// Passes: positions (dense node -> band position), count (triangles per cell,
// slot masks by atomicOr), two scans, vertices (+ one capped Newton step on
// the tricubic interpolant), emit (faces at the scanned offsets).
#include "common.h"
#include "compact.h"

namespace svimx {

// corner c = 4*x + 2*y + z; tetrahedron t walks 0 -> 7 along permutation t.
__constant__ int kTet[6][4] = {{0, 4, 6, 7}, {0, 4, 5, 7}, {0, 2, 6, 7}, {0, 2, 3, 7}, {0, 1, 5, 7}, {0, 1, 3, 7}};
// sign of det(v1 - v0, v2 - v0, v3 - v0) of each tetrahedron
__constant__ int kTetSign[6] = {1, -1, -1, 1, 1, -1};
// corner offset bits of direction 0..6 (x = 4, y = 2, z = 1)
__constant__ int kDirBits[7] = {4, 2, 1, 6, 5, 3, 7};

__device__ __forceinline__ int edge_dir(int dbits) {
  switch (dbits) {
    case 4: return 0;
    case 2: return 1;
    case 1: return 2;
    case 6: return 3;
    case 5: return 4;
    case 3: return 5;
    default: return 6;
  }
}

struct MtArgs {
  const float* phi;
  const int32_t* band;
  int64_t nb;
  Dims d;
  const int32_t* pos;  // dense node -> band position (-1 off the band)
};

__device__ __forceinline__ int64_t corner_node(const Dims& d, int64_t n0, int c) {
  return n0 + ((c >> 2) & 1) * d.sx() + ((c >> 1) & 1) * d.sy() + (c & 1);
}

// Slot (8 * band position + bit) of the vertex on edge (a inside, b outside),
// a and b corner ids of the cell at n0; -1 if the owner node is off the band.
__device__ __forceinline__ int64_t vertex_slot(const MtArgs& A, int64_t n0, int a, int b, float fb) {
  if (fb == 0.0f) {
    const int32_t p = A.pos[corner_node(A.d, n0, b)];
    return p < 0 ? -1 : 8 * (int64_t)p + 7;
  }
  // on a Kuhn edge one end's corner bits are a subset of the other's
  const int lo = ((a & b) == a) ? a : b;
  const int32_t p = A.pos[corner_node(A.d, n0, lo)];
  return p < 0 ? -1 : 8 * (int64_t)p + edge_dir(a ^ b);
}

// Triangles of tetrahedron t of the cell at n0: calls emit(s0, s1, s2) with
// vertex slots for each non-degenerate triangle and returns their number.
template <class Emit>
__device__ __forceinline__ int tet_triangles(const MtArgs& A, int64_t n0, const float cv[8], int t, Emit emit) {
  int ins[4], outs[4], order[4], ni = 0, no = 0;
  for (int q = 0; q < 4; ++q) {
    const int c = kTet[t][q];
    if (cv[c] < 0.0f) ins[ni++] = q; else outs[no++] = q;
  }
  if (ni == 0 || ni == 4) return 0;
  for (int q = 0; q < ni; ++q) order[q] = ins[q];
  for (int q = 0; q < no; ++q) order[ni + q] = outs[q];
  int inv = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) inv += order[i] > order[j];
  const bool flip = kTetSign[t] * ((inv & 1) ? -1 : 1) < 0;
  auto slot = [&](int qi, int qo) {
    const int a = kTet[t][qi], b = kTet[t][qo];
    return vertex_slot(A, n0, a, b, cv[b]);
  };
  auto tri = [&](int64_t s0, int64_t s1, int64_t s2) {
    if (s0 < 0 || s1 < 0 || s2 < 0 || s0 == s1 || s1 == s2 || s0 == s2) return 0;
    if (flip) emit(s0, s2, s1); else emit(s0, s1, s2);
    return 1;
  };
  if (ni == 1) return tri(slot(ins[0], outs[0]), slot(ins[0], outs[1]), slot(ins[0], outs[2]));
  if (ni == 3) return tri(slot(ins[0], outs[0]), slot(ins[1], outs[0]), slot(ins[2], outs[0]));
  const int64_t s0 = slot(ins[0], outs[0]), s1 = slot(ins[0], outs[1]), s2 = slot(ins[1], outs[1]),
                s3 = slot(ins[1], outs[0]);
  return tri(s0, s1, s2) + tri(s0, s2, s3);
}

__device__ __forceinline__ bool load_cell(const MtArgs& A, int64_t n0, float cv[8]) {
  int i, j, k;
  unflat(A.d, n0, i, j, k);
  if (i >= A.d.nx - 1 || j >= A.d.ny - 1 || k >= A.d.nz - 1) return false;
  int nin = 0;
  for (int c = 0; c < 8; ++c) {
    cv[c] = A.phi[corner_node(A.d, n0, c)];
    nin += cv[c] < 0.0f;
  }
  return nin > 0 && nin < 8;
}

__global__ void k_mt_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_mt_positions(const int32_t* band, int64_t nb, int32_t* pos, uint32_t* masks) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  pos[band[q]] = (int32_t)q;
  masks[q] = 0u;
}

__global__ void k_mt_count(MtArgs A, int32_t* ntri, uint32_t* masks) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.nb) return;
  const int64_t n0 = A.band[q];
  float cv[8];
  int cnt = 0;
  if (load_cell(A, n0, cv)) {
    auto mark = [&](int64_t s0, int64_t s1, int64_t s2) {
      atomicOr(&masks[s0 >> 3], 1u << (s0 & 7));
      atomicOr(&masks[s1 >> 3], 1u << (s1 & 7));
      atomicOr(&masks[s2 >> 3], 1u << (s2 & 7));
    };
    for (int t = 0; t < 6; ++t) cnt += tet_triangles(A, n0, cv, t, mark);
  }
  ntri[q] = cnt;
}

__global__ void k_mt_popcount(const uint32_t* masks, int64_t nb, int32_t* cnt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nb) cnt[q] = __popc(masks[q]);
}

// Index-space position of every vertex owned by band position q.
__global__ void k_mt_vertices(const float* phi, const int32_t* band, int64_t nb, Dims d, const uint32_t* masks,
                              const int64_t* voffs, float* xyz) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  uint32_t m = masks[q];
  if (!m) return;
  const int64_t na = band[q];
  int i, j, k;
  unflat(d, na, i, j, k);
  const float fa = phi[na];
  int64_t v = voffs[q];
  for (int dir = 0; dir < 8; ++dir) {
    if (!((m >> dir) & 1u)) continue;
    float x = (float)i, y = (float)j, z = (float)k;
    if (dir != 7) {
      const int b = kDirBits[dir];
      const float fb = phi[corner_node(d, na, b)];
      const float t = fa / (fa - fb);  // the ends have opposite signs, fb != 0
      x += t * (float)((b >> 2) & 1);
      y += t * (float)((b >> 1) & 1);
      z += t * (float)(b & 1);
    }
    xyz[3 * v] = x;
    xyz[3 * v + 1] = y;
    xyz[3 * v + 2] = z;
    ++v;
  }
}

// One capped Newton step onto the zero set of the tricubic interpolant of phi
// (svierosion/surface.py:44-51); points are first clamped to where the
// 4-point stencil fits.
__global__ void k_mt_project(const float* phi, Dims d, float* xyz, int64_t nv, float max_shift) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float hi[3] = {(float)d.nx - 3.001f, (float)d.ny - 3.001f, (float)d.nz - 3.001f};
  float p[3];
  for (int a = 0; a < 3; ++a) p[a] = fminf(fmaxf(xyz[3 * v + a], 1.0f), hi[a]);
  float g[3];
  const float val = tricubic_grad(phi, d, p, g);
  const float g2 = fmaxf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2], 1e-30f);
  float s[3];
  for (int a = 0; a < 3; ++a) s[a] = (val / g2) * g[a];
  const float len = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  const float scale = len > max_shift ? max_shift / len : 1.0f;
  for (int a = 0; a < 3; ++a) xyz[3 * v + a] = p[a] - s[a] * scale;
}

__device__ __forceinline__ int32_t slot_id(const uint32_t* masks, const int64_t* voffs, int64_t s) {
  const int64_t q = s >> 3;
  const uint32_t below = masks[q] & ((1u << (s & 7)) - 1u);
  return (int32_t)(voffs[q] + __popc(below));
}

__global__ void k_mt_emit(MtArgs A, const int64_t* toffs, const uint32_t* masks, const int64_t* voffs,
                          int32_t* faces) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.nb) return;
  const int64_t n0 = A.band[q];
  float cv[8];
  if (!load_cell(A, n0, cv)) return;
  int64_t o = toffs[q];
  auto put = [&](int64_t s0, int64_t s1, int64_t s2) {
    faces[3 * o] = slot_id(masks, voffs, s0);
    faces[3 * o + 1] = slot_id(masks, voffs, s1);
    faces[3 * o + 2] = slot_id(masks, voffs, s2);
    ++o;
  };
  for (int t = 0; t < 6; ++t) tet_triangles(A, n0, cv, t, put);
}

// stream-ordered device arrays (no device-wide synchronisation on free)
template <class T>
struct Buf {
  T* p = nullptr;
  hipStream_t st = nullptr;
  ~Buf() {
    if (p) hipFreeAsync(p, st);
  }
  hipError_t alloc(int64_t n, hipStream_t s) {
    st = s;
    return hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(T) * (n > 0 ? n : 1), s);
  }
  T* release() {
    T* r = p;
    p = nullptr;
    return r;
  }
};

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_extract_surface(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                          double max_shift, float** xyz_out, int64_t* nv_out, int32_t** faces_out,
                          int64_t* nf_out) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(phi && band && xyz_out && nv_out && faces_out && nf_out, SVIMX_EBADARG, "null argument");
  SVIMX_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && nb >= 0, SVIMX_EBADARG, "bad grid or band size");
  SVIMX_REQUIRE(4 * nb < (int64_t)INT32_MAX, SVIMX_EBADARG, "band too large for int32 vertex ids");
  const Dims d{nx, ny, nz};
  const int64_t n = d.size();
  hipStream_t st = as_stream(stream);
  *xyz_out = nullptr;
  *faces_out = nullptr;
  *nv_out = *nf_out = 0;
  if (nb == 0) return SVIMX_OK;
  Buf<int32_t> pos, ntri, vcnt;
  Buf<uint32_t> masks;
  Buf<int64_t> toffs, voffs;
  SVIMX_CHECK_HIP(pos.alloc(n, st));
  SVIMX_CHECK_HIP(ntri.alloc(nb, st));
  SVIMX_CHECK_HIP(vcnt.alloc(nb, st));
  SVIMX_CHECK_HIP(masks.alloc(nb, st));
  SVIMX_CHECK_HIP(toffs.alloc(nb + 1, st));
  SVIMX_CHECK_HIP(voffs.alloc(nb + 1, st));
  const unsigned gb = (unsigned)blocks_for(nb, 256);
  hipLaunchKernelGGL(k_mt_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, st, pos.p, n, -1);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mt_positions, dim3(gb), dim3(256), 0, st, band, nb, pos.p, masks.p);
  SVIMX_CHECK_LAUNCH();
  const MtArgs A{phi, band, nb, d, pos.p};
  hipLaunchKernelGGL(k_mt_count, dim3(gb), dim3(256), 0, st, A, ntri.p, masks.p);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mt_popcount, dim3(gb), dim3(256), 0, st, masks.p, nb, vcnt.p);
  SVIMX_CHECK_LAUNCH();
  int rc = scan_to_offsets(st, ntri.p, nb, toffs.p);
  if (rc) return rc;
  rc = scan_to_offsets(st, vcnt.p, nb, voffs.p);
  if (rc) return rc;
  int64_t nt = 0, nv = 0;
  SVIMX_CHECK_HIP(hipMemcpyAsync(&nt, toffs.p + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipMemcpyAsync(&nv, voffs.p + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  if (nt == 0) return SVIMX_OK;
  Buf<float> xyz;
  Buf<int32_t> faces;
  SVIMX_CHECK_HIP(xyz.alloc(3 * nv, st));
  SVIMX_CHECK_HIP(faces.alloc(3 * nt, st));
  hipLaunchKernelGGL(k_mt_vertices, dim3(gb), dim3(256), 0, st, phi, band, nb, d, masks.p, voffs.p, xyz.p);
  SVIMX_CHECK_LAUNCH();
  if (max_shift > 0) {
    hipLaunchKernelGGL(k_mt_project, dim3(blocks_for(nv, 256)), dim3(256), 0, st, phi, d, xyz.p, nv,
                       (float)max_shift);
    SVIMX_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_mt_emit, dim3(gb), dim3(256), 0, st, A, toffs.p, masks.p, voffs.p, faces.p);
  SVIMX_CHECK_LAUNCH();
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  *nv_out = nv;
  *nf_out = nt;
  *xyz_out = xyz.release();
  *faces_out = faces.release();
  return SVIMX_OK;
  SVIMX_TRY_END
}

// Both are ordered on the caller's stream, never on the legacy null stream:
// another thread may be capturing a HIP graph on its own stream meanwhile (the
// batch engine), and null-stream work is the kind of cross-stream dependency
// that can invalidate a capture.
int svimx_free(void* stream, void* p) {
  if (p) SVIMX_CHECK_HIP(hipFreeAsync(p, as_stream(stream)));
  return SVIMX_OK;
}

int svimx_copy_to_host(void* stream, void* dst, const void* src, int64_t bytes) {
  if (bytes <= 0) return SVIMX_OK;
  SVIMX_REQUIRE(dst && src, SVIMX_EBADARG, "null argument");
  SVIMX_CHECK_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  SVIMX_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
  return SVIMX_OK;
}

// Timing events: no system-scope fence when they are recorded, so bracketing a
// launch with them does not write back and invalidate the caches between that
// launch and its neighbours (what a default event does on this device).
int svimx_timing_event_create(void** event) {
  SVIMX_REQUIRE(event, SVIMX_EBADARG, "null argument");
  hipEvent_t e = nullptr;
  SVIMX_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  *event = e;
  return SVIMX_OK;
}

int svimx_timing_event_record(void* event, void* stream) {
  SVIMX_REQUIRE(event, SVIMX_EBADARG, "null event");
  SVIMX_CHECK_HIP(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)));
  return SVIMX_OK;
}

int svimx_timing_event_elapsed(void* start, void* end, float* ms) {
  SVIMX_REQUIRE(start && end && ms, SVIMX_EBADARG, "null argument");
  SVIMX_CHECK_HIP(hipEventSynchronize(static_cast<hipEvent_t>(end)));
  SVIMX_CHECK_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
  return SVIMX_OK;
}

int svimx_timing_event_destroy(void* event) {
  if (event) SVIMX_CHECK_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return SVIMX_OK;
}

}  // extern "C"
