// PDE reinitialisation (Min 2010) on the dilated band, followed by the band
// rebuild (svierosion/levelset.py:154-183).
//
// Work set: the band scattered into a byte mask, dilated by a (2r+1)^3 cube
// with three separable line passes, minus `margin` layers at the grid faces,
// then compacted in node order -- svierosion/levelset.py:166-174.
//
// Geometry (subcell crossings and local dtau, svierosion/kernels.py:335-359)
// is computed once from phi0.  Only nodes next to the interface have a
// crossing; they keep dtau and the six subcell distances in their slot of
// `geo` (slot = work index), every other node only its 4-byte info word (sign,
// no crossing, dtau = cfl*h), so a sweep streams ~12 bytes for most nodes.
//
// Each TVD-RK2 iteration is two Jacobi sweeps ping-ponging between phi and a
// scratch field:  scratch = phi + dtau R(phi);  phi = (phi + scratch +
// dtau R(scratch)) / 2  (svierosion/levelset.py:178-182).  The second sweep
// reads phi only at its own node, so writing phi in place is race-free.
//
// Rebuild: every band node lies in the work set (the work set is the band
// dilated) and no node outside the work set changed, so clamping, the copies
// into the other two stage buffers, the occupancy reset and the new band are
// all computed over the work set alone; the count of inside nodes outside the
// band follows from counts over the old band and the work set.  Counts and the
// band extent use per-block partials and the deterministic finalisation.
#include "common.h"
#include "compact.h"

namespace svimx {

__global__ void k_u8_fill(uint8_t* p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_scatter_mark(uint8_t* mask, const int32_t* idx, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mask[idx[i]] = 1;
}

// out = OR over offsets -r..r along `axis` of in; the last pass also clears
// `margin` layers at every face.
__global__ void k_dilate_axis(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, Dims d, int axis, int r,
                              int margin) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.size()) return;
  int c[3];
  unflat(d, e, c[0], c[1], c[2]);
  const int n[3] = {d.nx, d.ny, d.nz};
  const int64_t st[3] = {d.sx(), d.sy(), 1};
  int lo = c[axis] - r < 0 ? 0 : c[axis] - r;
  int hi = c[axis] + r > n[axis] - 1 ? n[axis] - 1 : c[axis] + r;
  uint8_t v = 0;
  for (int q = lo; q <= hi && !v; ++q) v = in[e + (int64_t)(q - c[axis]) * st[axis]];
  if (margin > 0) {
    for (int a = 0; a < 3; ++a)
      if (c[a] < margin || c[a] > n[a] - 1 - margin) v = 0;
  }
  out[e] = v;
}

struct MaskPred {
  const uint8_t* m;
  __device__ bool operator()(int64_t e) const { return m[e] != 0; }
};

// partials[b][slot] = #{i in block b's share : phi[list[i]] < 0}
__global__ void __launch_bounds__(kBandThreads)
    k_count_neg(const float* __restrict__ phi, const int32_t* __restrict__ list, int64_t n,
                double* __restrict__ partials, int slot) {
  __shared__ double red[4];
  double c = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += phi[list[i]] < 0.0f ? 1.0 : 0.0;
  c = block_reduce<0>(c, red);
  if (threadIdx.x == 0) partials[blockIdx.x * kPartialSlots + slot] = c;
}

__global__ void k_copy_list(const float* __restrict__ src, float* __restrict__ dst, const int32_t* __restrict__ idx,
                            int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[idx[i]];
}

// info word: bits 0-5 crossing flags, bits 6-7 sign code.  Nodes with a
// crossing store dtau and the six distances at geo[8 i].
__global__ void k_reinit_geo(const float* __restrict__ phi, const int32_t* __restrict__ idx, int64_t n, Dims d,
                             float h, float cfl, uint32_t* __restrict__ info, float* __restrict__ geo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ReinitGeo g = reinit_geometry(phi, idx[i], d, h, cfl);
  if (g.flags & 63u) {
    float* o = geo + 8 * i;
    o[0] = g.dtau;
    for (int q = 0; q < 6; ++q) o[1 + q] = g.x[q];
  }
  info[i] = g.flags & 0xFFu;
}

// mode 0: out = in + dtau R(in);  mode 1: out = (out + in + dtau R(in)) / 2.
__global__ void __launch_bounds__(256) k_reinit_sweep(int mode, const float* __restrict__ in, float* out,
                                                      const int32_t* __restrict__ idx, int64_t n, Dims d, float h,
                                                      float dtau0, const uint32_t* __restrict__ info,
                                                      const float* __restrict__ geo) {
  // XCD-contiguous slices of the (sorted) work list keep stencil neighbours in one L2
  const int lb = (blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds;
  const int64_t i = (int64_t)lb * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t node = idx[i];
  const uint32_t word = info[i];
  float dtau = dtau0;
  float x[6] = {0, 0, 0, 0, 0, 0};
  if (word & 63u) {
    const float* g = geo + 8 * i;
    dtau = g[0];
    for (int q = 0; q < 6; ++q) x[q] = g[1 + q];
  }
  const float r = reinit_rhs(in, node, d, h, word, x);
  if (mode == 0) {
    out[node] = in[node] + dtau * r;
  } else {
    out[node] = 0.5f * (out[node] + in[node] + dtau * r);
  }
}

// Clamp the work set, copy it into the other stage buffers and reset its
// occupancy.  Partials: slot 2 inside-and-outside-the-new-band count, slots
// 3..5 / 6..8 min / max i, j, k of the new band.
__global__ void __launch_bounds__(kBandThreads)
    k_rebuild_work(float* __restrict__ phi, float* __restrict__ c1, float* __restrict__ c2, float* __restrict__ occ,
                   const int32_t* __restrict__ idx, int64_t n, Dims d, float cap, float thr,
                   double* __restrict__ partials) {
  __shared__ double red[4];
  const OccLayout OL = OccLayout::of(d.nx, d.ny, d.nz);
  double outneg = 0.0;
  double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t node = idx[i];
    const float v = fminf(fmaxf(phi[node], -cap), cap);
    phi[node] = v;
    c1[node] = v;
    c2[node] = v;
    int a, b, c;
    unflat(d, node, a, b, c);
    occ[OL.at(a, b, c)] = v < 0.0f ? 1.0f : 0.0f;
    const bool inb = fabsf(v) < thr;
    if (!inb && v < 0.0f) outneg += 1.0;
    if (inb) {
      lo[0] = fmin(lo[0], (double)a); hi[0] = fmax(hi[0], (double)a);
      lo[1] = fmin(lo[1], (double)b); hi[1] = fmax(hi[1], (double)b);
      lo[2] = fmin(lo[2], (double)c); hi[2] = fmax(hi[2], (double)c);
    }
  }
  double r = block_reduce<0>(outneg, red);
  if (threadIdx.x == 0) partials[blockIdx.x * kPartialSlots + 2] = r;
  for (int q = 0; q < 3; ++q) {
    double l = block_reduce<2>(lo[q], red);
    if (threadIdx.x == 0) partials[blockIdx.x * kPartialSlots + 3 + q] = l;
    double u = block_reduce<1>(hi[q], red);
    if (threadIdx.x == 0) partials[blockIdx.x * kPartialSlots + 6 + q] = u;
  }
}

struct WorkBandPred {
  const float* phi;
  const int32_t* idx;
  float thr;
  __device__ bool operator()(int64_t i) const { return fabsf(phi[idx[i]]) < thr; }
};

// Result scratch slots 19..27 (outside the public SVIMX_R_* range).
constexpr int kRsOldNeg = 19, kRsWorkNeg = 20, kRsOutNeg = 21, kRsLo = 22, kRsHi = 25;

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_reinit(void* stream, float* phi, float* scratch, float* phi_copy, float* occupancy, int32_t* band,
                 int64_t nb, int64_t inside_outside, int nx, int ny, int nz, double h, double width, int iterations,
                 double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b, int32_t* work_idx,
                 int32_t* work_info, float* geo, int64_t work_cap, void* ws, int64_t ws_bytes, int64_t* info) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  const int64_t n = d.size();
  SVIMX_REQUIRE(phi && scratch && phi_copy && occupancy && band && mask_a && mask_b && work_idx && work_info && geo &&
                    info,
                SVIMX_EBADARG, "null reinit buffer");
  SVIMX_REQUIRE(phi != scratch && phi != phi_copy && scratch != phi_copy, SVIMX_EBADARG, "aliased reinit buffers");
  SVIMX_REQUIRE(iterations >= 0 && dilation >= 0 && margin >= 0, SVIMX_EBADARG, "bad reinit parameters");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, n, w)) return SVIMX_EBADARG;
  hipStream_t st = as_stream(stream);
  const dim3 gn(blocks_for(n, 256)), tb(256);
  ReduceSpec spec;
  auto one_slot = [&](int slot, int op, int dst) {
    for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
    spec.op[slot] = (int8_t)op;
    spec.dst[slot] = (int8_t)dst;
  };
  hipLaunchKernelGGL(k_u8_fill, gn, tb, 0, st, mask_a, n, (uint8_t)0);
  SVIMX_CHECK_LAUNCH();
  int nblk = reduce_grid(nb);
  hipLaunchKernelGGL(k_count_neg, dim3(nblk), tb, 0, st, phi, band, nb, w.partials, 0);
  SVIMX_CHECK_LAUNCH();
  one_slot(0, 0, kRsOldNeg);
  int rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  if (nb > 0) {
    hipLaunchKernelGGL(k_scatter_mark, dim3(blocks_for(nb, 256)), tb, 0, st, mask_a, band, nb);
    SVIMX_CHECK_LAUNCH();
  }
  // x, y, z passes: a -> b -> a -> b; margin cleared on the last
  hipLaunchKernelGGL(k_dilate_axis, gn, tb, 0, st, mask_a, mask_b, d, 0, dilation, 0);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_dilate_axis, gn, tb, 0, st, mask_b, mask_a, d, 1, dilation, 0);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_dilate_axis, gn, tb, 0, st, mask_a, mask_b, d, 2, dilation, margin);
  SVIMX_CHECK_LAUNCH();
  int64_t nw = 0;
  rc = compact(st, MaskPred{mask_b}, n, work_idx, work_cap, w, &nw);
  info[9] = nw;
  if (rc) return rc;
  const float capf = (float)(width * h);
  const float thr = (float)(width * h * (1.0 - 1e-5));
  nblk = reduce_grid(nw);
  hipLaunchKernelGGL(k_count_neg, dim3(nblk), tb, 0, st, phi, work_idx, nw, w.partials, 1);
  SVIMX_CHECK_LAUNCH();
  one_slot(1, 0, kRsWorkNeg);
  rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  if (nw > 0) {
    const int gwb = (int)((blocks_for(nw, 256) + kXcds - 1) / kXcds * kXcds);
    const dim3 gw(blocks_for(nw, 256));
    hipLaunchKernelGGL(k_copy_list, gw, tb, 0, st, phi, scratch, work_idx, nw);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reinit_geo, gw, tb, 0, st, phi, work_idx, nw, d, (float)h, (float)cfl,
                       (uint32_t*)work_info, geo);
    SVIMX_CHECK_LAUNCH();
    const float dtau0 = (float)cfl * fmaxf((float)h, 1e-3f * (float)h);
    for (int it = 0; it < iterations; ++it) {
      hipLaunchKernelGGL(k_reinit_sweep, dim3(gwb), tb, 0, st, 0, phi, scratch, work_idx, nw, d, (float)h, dtau0,
                         (const uint32_t*)work_info, geo);
      SVIMX_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_reinit_sweep, dim3(gwb), tb, 0, st, 1, scratch, phi, work_idx, nw, d, (float)h, dtau0,
                         (const uint32_t*)work_info, geo);
      SVIMX_CHECK_LAUNCH();
    }
  }
  hipLaunchKernelGGL(k_rebuild_work, dim3(nblk), tb, 0, st, phi, scratch, phi_copy, occupancy, work_idx, nw, d, capf,
                     thr, w.partials);
  SVIMX_CHECK_LAUNCH();
  for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
  spec.op[2] = 0; spec.dst[2] = kRsOutNeg;
  for (int q = 0; q < 3; ++q) {
    spec.op[3 + q] = 2; spec.dst[3 + q] = (int8_t)(kRsLo + q);
    spec.op[6 + q] = 1; spec.dst[6 + q] = (int8_t)(kRsHi + q);
  }
  rc = launch_finalize(st, w.partials, nblk, w.results, spec);
  if (rc) return rc;
  int64_t nbn = 0;
  rc = compact(st, WorkBandPred{phi, work_idx, thr}, nw, band, n, w, &nbn, ListMap{work_idx});
  info[0] = nbn;
  if (rc) return rc;
  double r[9];
  SVIMX_CHECK_HIP(hipMemcpyAsync(r, w.results + kRsOldNeg, sizeof(r), hipMemcpyDeviceToHost, st));
  SVIMX_CHECK_HIP(hipStreamSynchronize(st));
  // inside nodes outside the band: (all negatives before) - (work negatives before) + (new outside-band negatives)
  info[1] = inside_outside + (int64_t)r[0] - (int64_t)r[1] + (int64_t)r[2];
  const int dims[3] = {nx, ny, nz};
  bool ok = true;
  for (int a = 0; a < 3; ++a) {
    info[3 + a] = nbn ? (int64_t)r[3 + a] : 0;
    info[6 + a] = nbn ? (int64_t)r[6 + a] : 0;
    if (nbn && (r[3 + a] < margin || r[6 + a] > dims[a] - 1 - margin)) ok = false;
  }
  info[2] = ok ? 1 : 0;
  if (!ok) {
    set_error("narrow band reaches the grid boundary; enlarge the padding");
    return SVIMX_EMARGIN;
  }
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
