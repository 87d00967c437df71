// PDE reinitialisation (Min 2010) on the dilated band.
//
// Work set: band scattered into a byte mask, dilated by a (2r+1)^3 cube with
// three separable line passes, minus `margin` layers at the grid faces, then
// compacted (sorted) -- svierosion/levelset.py:166-174.  Geometry (subcell
// crossings, local dtau) is computed once from phi0 and kept as 8 floats per
// node.  Each TVD-RK2 iteration is two Jacobi sweeps ping-ponging between phi
// and a scratch field:  scratch = phi + dtau R(phi);  phi = (phi + scratch +
// dtau R(scratch)) / 2  (svierosion/levelset.py:178-182).  The second sweep
// reads phi only at its own node, so writing phi in place is race-free.
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

__global__ void k_copy_list(const float* __restrict__ src, float* __restrict__ dst, const int32_t* __restrict__ idx,
                            int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[idx[i]];
}

// Geometry in structure-of-arrays form: geo[s * nwork + i], s = 0 dtau,
// 1..6 subcell distances, 7 flags (as float).
__global__ void k_reinit_geo(const float* __restrict__ phi, const int32_t* __restrict__ idx, int64_t n, Dims d,
                             float h, float cfl, float* __restrict__ geo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ReinitGeo g = reinit_geometry(phi, idx[i], d, h, cfl);
  geo[i] = g.dtau;
  for (int s = 0; s < 6; ++s) geo[(1 + s) * n + i] = g.x[s];
  geo[7 * n + i] = (float)g.flags;
}

// mode 0: out = in + dtau R(in);  mode 1: out = (out + in + dtau R(in)) / 2.
__global__ void __launch_bounds__(256) k_reinit_sweep(int mode, const float* __restrict__ in, float* out,
                                                      const int32_t* __restrict__ idx, int64_t n, Dims d, float h,
                                                      const float* __restrict__ geo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t node = idx[i];
  float x[6];
  for (int s = 0; s < 6; ++s) x[s] = geo[(1 + s) * n + i];
  const uint32_t flags = (uint32_t)geo[7 * n + i];
  const float dtau = geo[i];
  const float r = reinit_rhs(in, node, d, h, flags, x);
  if (mode == 0) {
    out[node] = in[node] + dtau * r;
  } else {
    out[node] = 0.5f * (out[node] + in[node] + dtau * r);
  }
}

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_reinit(void* stream, float* phi, float* scratch, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                 double h, int iterations, double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b,
                 int32_t* work_idx, float* geo, int64_t work_cap, void* ws, int64_t ws_bytes, int64_t* n_work) {
  SVIMX_TRY_BEGIN
  Dims d{nx, ny, nz};
  const int64_t n = d.size();
  SVIMX_REQUIRE(phi && scratch && phi != scratch && mask_a && mask_b && work_idx && geo && n_work, SVIMX_EBADARG,
                "null or aliased reinit buffer");
  SVIMX_REQUIRE(iterations >= 0 && dilation >= 0 && margin >= 0, SVIMX_EBADARG, "bad reinit parameters");
  Workspace w;
  if (!carve_workspace(ws, ws_bytes, n, w)) return SVIMX_EBADARG;
  hipStream_t st = as_stream(stream);
  const dim3 gn(blocks_for(n, 256)), tb(256);
  hipLaunchKernelGGL(k_u8_fill, gn, tb, 0, st, mask_a, n, (uint8_t)0);
  SVIMX_CHECK_LAUNCH();
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
  int rc = compact(st, MaskPred{mask_b}, n, work_idx, work_cap, w, &nw);
  *n_work = nw;
  if (rc) return rc;
  if (nw == 0) return SVIMX_OK;
  const dim3 gw(blocks_for(nw, 256));
  hipLaunchKernelGGL(k_copy_list, gw, tb, 0, st, phi, scratch, work_idx, nw);
  SVIMX_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_reinit_geo, gw, tb, 0, st, phi, work_idx, nw, d, (float)h, (float)cfl, geo);
  SVIMX_CHECK_LAUNCH();
  for (int it = 0; it < iterations; ++it) {
    hipLaunchKernelGGL(k_reinit_sweep, gw, tb, 0, st, 0, phi, scratch, work_idx, nw, d, (float)h, geo);
    SVIMX_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_reinit_sweep, gw, tb, 0, st, 1, scratch, phi, work_idx, nw, d, (float)h, geo);
    SVIMX_CHECK_LAUNCH();
  }
  return SVIMX_OK;
  SVIMX_TRY_END
}

}  // extern "C"
