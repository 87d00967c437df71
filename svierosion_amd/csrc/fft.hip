// N_r by FFT: the reference's method (svierosion/svi.py:66-93), for balls too
// wide for the direct summation and as a selectable alternative.
//
// N_r = 1 - h^3 (w * f) / |B_r| with the circular convolution computed as
//     conv = IFFT(FFT(f) * K^),   K = ball weights at wrapped offsets.
// K is real and even (the weights depend on |dx|, |dy|, |dz|), so K^ is real,
// stored as one float per node.  The inverse transform is a forward one on the
// conjugate: conv = conj(FFT(conj(FFT(f) K^))) / N, and only Re(conv) is used,
// so six forward line passes do the whole evaluation:
//     z (real f from the padded occupancy) -> y -> x (epilogue: conj(. K^))
//     -> x -> y -> z (epilogue: N_r = 1 + negbv h^3 Re(.) / N)
// over one complex buffer (in place per line tile).
//
// Line transform: mixed-radix Stockham autosort (radices 4, 2, 3, 5, 7; every
// grid extent is a product of 2, 3, 5, 7 by construction) on a tile of T <= 16
// lines (8 of 432 points) held in one LDS buffer (row pitch L+1 against bank
// conflicts), each stage transformed in registers between two block barriers.
// Stage with radix r after m points: butterfly j of L/r (k = j mod m) reads
// x_q = A[j + q L/r] w^(qk / (m r)), forms the r-point DFT y_p, and writes
// B[(j / m) m r + k + p m].  Lines along y and x are tiled with consecutive z
// so global loads and stores coalesce across the lines of a tile.  Twiddles
// come from per-length tables computed in FP64 on the host.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "fft.h"

namespace svimx {


enum FftIn { kInComplex = 0, kInOccupancy = 1 };
enum FftOut { kOutComplex = 0, kOutConjMulK = 1, kOutNr = 2, kOutRealK = 3 };

struct FftPass {
  int axis;        // 0 x, 1 y, 2 z
  int L;           // line length
  int T;           // lines per tile
  int nst;         // radix stages
  int radix[16];
  uint64_t mag_nb[16], mag_m[16];  // floor(2^32 / d) + 1 of L / radix and of the stage's m: exact
                                   // quotients of numerators below 2^16 by (x * mag) >> 32
  int64_t nlines;  // lines of this axis
  const float2* tw;
  Dims d;
  OccLayout OL;
  int in_mode, out_mode;
  float coef;      // kOutNr: negbv * h^3 / N
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// element t of line l: dense offset (line indices fit in 32 bits)
__device__ __forceinline__ int64_t line_offset(const FftPass& P, int64_t l, int t) {
  const Dims& d = P.d;
  if (P.axis == 2) return l * d.nz + t;
  if (P.axis == 1) {
    const uint32_t x = (uint32_t)l / (uint32_t)d.nz, z = (uint32_t)l - x * (uint32_t)d.nz;
    return (int64_t)x * d.sx() + (int64_t)t * d.nz + z;
  }
  return l + (int64_t)t * d.sx();
}

// One radix-R Stockham stage over the tile in place: every thread reads and
// transforms its butterflies into registers (at most fft_kb(R) of them), the
// block barrier orders all reads before any write, then the outputs go back
// to the same LDS buffer.  One buffer instead of two ping-pong halves lets a
// tile hold twice the lines (the y and x passes move 64-byte row segments at
// 432 points, 128-byte ones at 224 or less).
__host__ __device__ constexpr int fft_pitch(int L) { return L + 1; }  // row pitch of a line in LDS

__device__ __forceinline__ int fft_div(int x, uint64_t mag) { return (int)(((uint64_t)x * mag) >> 32); }

// exp(-2 pi i k / R) as FP32 constants (the FP64 values rounded), so the
// unrolled butterflies multiply by immediates and drop the exact zero parts
template <int R>
__device__ __forceinline__ float2 fft_root(int k) {
  constexpr double c3 = -0.5, s3 = 0.86602540378443864676;
  constexpr double c5a = 0.30901699437494742410, s5a = 0.95105651629515357212;
  constexpr double c5b = -0.80901699437494742410, s5b = 0.58778525229247312917;
  constexpr double c7a = 0.62348980185873353053, s7a = 0.78183148246802980871;
  constexpr double c7b = -0.22252093395631440429, s7b = 0.97492791218182360702;
  constexpr double c7c = -0.90096886790241912624, s7c = 0.43388373911755812048;
  if constexpr (R == 2) return k ? make_float2(-1.0f, 0.0f) : make_float2(1.0f, 0.0f);
  if constexpr (R == 4) {
    switch (k) {
      case 1: return make_float2(0.0f, -1.0f);
      case 2: return make_float2(-1.0f, 0.0f);
      case 3: return make_float2(0.0f, 1.0f);
      default: return make_float2(1.0f, 0.0f);
    }
  }
  if constexpr (R == 3) {
    switch (k) {
      case 1: return make_float2((float)c3, (float)-s3);
      case 2: return make_float2((float)c3, (float)s3);
      default: return make_float2(1.0f, 0.0f);
    }
  }
  if constexpr (R == 5) {
    switch (k) {
      case 1: return make_float2((float)c5a, (float)-s5a);
      case 2: return make_float2((float)c5b, (float)-s5b);
      case 3: return make_float2((float)c5b, (float)s5b);
      case 4: return make_float2((float)c5a, (float)s5a);
      default: return make_float2(1.0f, 0.0f);
    }
  }
  switch (k) {  // R == 7
    case 1: return make_float2((float)c7a, (float)-s7a);
    case 2: return make_float2((float)c7b, (float)-s7b);
    case 3: return make_float2((float)c7c, (float)-s7c);
    case 4: return make_float2((float)c7c, (float)s7c);
    case 5: return make_float2((float)c7b, (float)s7b);
    case 6: return make_float2((float)c7a, (float)s7a);
    default: return make_float2(1.0f, 0.0f);
  }
}

template <int R>
__device__ __forceinline__ void fft_stage(float2* A, const float2* tw, int L, int tl, int m, int pitch, uint64_t mnb,
                                          uint64_t mm) {
  constexpr int KB = fft_kb(R);
  const int nb = L / R, mr = m * R, twstep = L / mr, nbt = tl * nb;
  float2 y[KB][R];
#pragma unroll
  for (int i = 0; i < KB; ++i) {
    const int b = threadIdx.x + i * kFftThreads;
    if (b < nbt) {
      const int l = fft_div(b, mnb), j = b - l * nb, k = j - fft_div(j, mm) * m;
      float2 x[R];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const float2 a = A[l * pitch + j + q * nb];
        x[q] = q ? cmul(a, tw[q * k * twstep]) : a;  // q k twstep < L
      }
#pragma unroll
      for (int p = 0; p < R; ++p) {
        float2 acc = x[0];
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const float2 w = fft_root<R>((p * q) % R);
          const float re = (w.x != 0.0f ? x[q].x * w.x : 0.0f) - (w.y != 0.0f ? x[q].y * w.y : 0.0f);
          const float im = (w.y != 0.0f ? x[q].x * w.y : 0.0f) + (w.x != 0.0f ? x[q].y * w.x : 0.0f);
          acc.x += re;
          acc.y += im;
        }
        y[i][p] = acc;
      }
    }
  }
  __syncthreads();  // every read of A is done before any write
#pragma unroll
  for (int i = 0; i < KB; ++i) {
    const int b = threadIdx.x + i * kFftThreads;
    if (b < nbt) {
      const int l = fft_div(b, mnb), j = b - l * nb, jm = fft_div(j, mm), k = j - jm * m;
      const int base = jm * mr + k;
#pragma unroll
      for (int p = 0; p < R; ++p) A[l * pitch + base + p * m] = y[i][p];
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kFftThreads) k_fft_lines(FftPass P, float2* __restrict__ data,
                                                           const float* __restrict__ occ,
                                                           const float* __restrict__ khat, float* __restrict__ out,
                                                           float* __restrict__ khat_out) {
  extern __shared__ float2 lds[];
  const int L = P.L, T = P.T, pitch = fft_pitch(L);
  float2* A = lds;
  float2* TW = lds + T * pitch;  // the axis' twiddles, read by every stage
  for (int t = threadIdx.x; t < L; t += blockDim.x) TW[t] = P.tw[t];
  const int64_t l0 = (int64_t)blockIdx.x * T;
  const int tl = (int)min<int64_t>(T, P.nlines - l0);
  const int n = tl * L;
  // load: z lines element-major within a line, y/x lines line-major (coalesced z)
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    int l, t;
    if (P.axis == 2) { l = e / L; t = e - l * L; } else { t = e / tl; l = e - t * tl; }
    const int64_t off = line_offset(P, l0 + l, t);
    float2 v;
    if (P.in_mode == kInOccupancy) {
      const uint32_t line = (uint32_t)(l0 + l);  // axis z: line = x * ny + y
      const int x = (int)(line / (uint32_t)P.d.ny), y = (int)(line - (uint32_t)x * (uint32_t)P.d.ny);
      v = make_float2(occ[P.OL.at(x, y, t)], 0.0f);
    } else {
      v = data[off];
    }
    A[l * pitch + t] = v;
  }
  __syncthreads();
  int m = 1;
  for (int s = 0; s < P.nst; ++s) {
    const int r = P.radix[s];  // uniform
    const uint64_t mnb = P.mag_nb[s], mm = P.mag_m[s];
    switch (r) {
      case 2: fft_stage<2>(A, TW, L, tl, m, pitch, mnb, mm); break;
      case 3: fft_stage<3>(A, TW, L, tl, m, pitch, mnb, mm); break;
      case 4: fft_stage<4>(A, TW, L, tl, m, pitch, mnb, mm); break;
      case 5: fft_stage<5>(A, TW, L, tl, m, pitch, mnb, mm); break;
      default: fft_stage<7>(A, TW, L, tl, m, pitch, mnb, mm); break;
    }
    m *= r;
  }
  // store with the pass epilogue
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    int l, t;
    if (P.axis == 2) { l = e / L; t = e - l * L; } else { t = e / tl; l = e - t * tl; }
    const int64_t off = line_offset(P, l0 + l, t);
    const float2 v = A[l * pitch + t];
    switch (P.out_mode) {
      case kOutConjMulK: {
        const float k = khat[off];
        data[off] = make_float2(v.x * k, -v.y * k);
        break;
      }
      case kOutNr: out[off] = 1.0f + P.coef * v.x; break;
      case kOutRealK: khat_out[off] = v.x; break;
      default: data[off] = v;
    }
  }
}

__global__ void k_fft_zero(float2* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float2(0.0f, 0.0f);
}

// ball weights at wrapped offsets: K[(dx mod nx, dy mod ny, dz mod nz)] = w(dx, dy, dz)
__global__ void k_fft_place_kernel(float2* C, const float* __restrict__ w, int m, Dims d) {
  const int W = 2 * m + 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)W * W * W) return;
  const int c = (int)(i % W), b = (int)((i / W) % W), a = (int)(i / ((int64_t)W * W));
  const int x = ((a - m) % d.nx + d.nx) % d.nx, y = ((b - m) % d.ny + d.ny) % d.ny, z = ((c - m) % d.nz + d.nz) % d.nz;
  C[((int64_t)x * d.ny + y) * d.nz + z] = make_float2(w[i], 0.0f);
}

static std::vector<int> fft_radices(int L) {
  std::vector<int> r;
  for (int q : {4, 2, 3, 5, 7})
    while (L % q == 0) {
      r.push_back(q);
      L /= q;
    }
  if (L != 1) r.clear();
  return r;
}

static FftPass make_pass(const FftPlan& F, int axis, int in_mode, int out_mode, float coef) {
  FftPass P{};
  P.axis = axis;
  P.L = F.len[axis];
  const int L = P.L;
  P.T = std::max(1, std::min(16, (kFftTileElems / fft_pitch(L))));
  const std::vector<int> r = fft_radices(L);
  P.nst = (int)r.size();
  int m = 1;
  for (int s = 0; s < P.nst; ++s) {
    P.radix[s] = r[s];
    P.mag_nb[s] = (1ull << 32) / (uint64_t)(L / r[s]) + 1;
    P.mag_m[s] = (1ull << 32) / (uint64_t)m + 1;
    m *= r[s];
  }
  const Dims& d = F.d;
  P.nlines = axis == 2 ? (int64_t)d.nx * d.ny : (axis == 1 ? (int64_t)d.nx * d.nz : (int64_t)d.ny * d.nz);
  P.tw = F.tw[axis];
  P.d = d;
  P.OL = OccLayout::of(d.nx, d.ny, d.nz);
  P.in_mode = in_mode;
  P.out_mode = out_mode;
  P.coef = coef;
  return P;
}

static int run_pass(hipStream_t st, const FftPass& P, float2* data, const float* occ, const float* khat, float* out,
                    float* khat_out) {
  const unsigned grid = (unsigned)((P.nlines + P.T - 1) / P.T);
  const size_t lds = sizeof(float2) * ((size_t)P.T * fft_pitch(P.L) + (size_t)P.L);
  hipLaunchKernelGGL(k_fft_lines, dim3(grid), dim3(kFftThreads), lds, st, P, data, occ, khat, out, khat_out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int fft_plan_create(FftPlan& F, const double* weights, int m, Dims d) {
  F = FftPlan{};
  F.d = d;
  F.len[0] = d.nx;
  F.len[1] = d.ny;
  F.len[2] = d.nz;
  for (int a = 0; a < 3; ++a) {
    SVIMX_REQUIRE(F.len[a] >= 1 && fft_pitch(F.len[a]) <= kFftTileElems && !fft_radices(F.len[a]).empty(),
                  SVIMX_EBADARG, "FFT needs grid extents that are products of 2, 3, 5, 7 and below %d (got %d)",
                  kFftTileElems - 1,
                  F.len[a]);
  }
  const int64_t n = d.size();
  SVIMX_CHECK_HIP(hipMalloc(&F.C, sizeof(float2) * n));
  SVIMX_CHECK_HIP(hipMalloc(&F.khat, sizeof(float) * n));
  for (int a = 0; a < 3; ++a) {
    const int L = F.len[a];
    std::vector<float2> tw(L);
    for (int t = 0; t < L; ++t) {
      const double ang = -2.0 * M_PI * t / L;
      tw[t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    SVIMX_CHECK_HIP(hipMalloc(&F.tw[a], sizeof(float2) * L));
    SVIMX_CHECK_HIP(hipMemcpy(F.tw[a], tw.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
  }
  // K^ = Re FFT(K): place the weights, three passes, keep the real part
  const int W = 2 * m + 1;
  std::vector<float> wf((size_t)W * W * W);
  for (size_t i = 0; i < wf.size(); ++i) wf[i] = (float)weights[i];
  float* dw = nullptr;
  SVIMX_CHECK_HIP(hipMalloc(&dw, sizeof(float) * wf.size()));
  hipError_t e = hipMemcpy(dw, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(dw);
    SVIMX_CHECK_HIP(e);
  }
  hipStream_t st = nullptr;
  hipLaunchKernelGGL(k_fft_zero, dim3(blocks_for(n, 256)), dim3(256), 0, st, F.C, n);
  hipLaunchKernelGGL(k_fft_place_kernel, dim3(blocks_for((int64_t)W * W * W, 256)), dim3(256), 0, st, F.C, dw, m, d);
  int rc = run_pass(st, make_pass(F, 2, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 1, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 0, kInComplex, kOutRealK, 0.0f), F.C, nullptr, nullptr, nullptr, F.khat);
  e = hipStreamSynchronize(st);
  hipFree(dw);
  if (rc) return rc;
  SVIMX_CHECK_HIP(e);
  return SVIMX_OK;
}

void fft_plan_destroy(FftPlan& F) {
  hipFree(F.C);
  hipFree(F.khat);
  for (int a = 0; a < 3; ++a) hipFree(F.tw[a]);
  F = FftPlan{};
}

int fft_nr(hipStream_t st, const FftPlan& F, const float* occupancy, float* nr_out, float scale, float negbv) {
  const float coef = (float)((double)negbv * (double)scale / (double)F.d.size());
  int rc = run_pass(st, make_pass(F, 2, kInOccupancy, kOutComplex, 0.0f), F.C, occupancy, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 1, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 0, kInComplex, kOutConjMulK, 0.0f), F.C, nullptr, F.khat, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 0, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 1, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 2, kInComplex, kOutNr, coef), F.C, nullptr, nullptr, nr_out, nullptr);
  return rc;
}

}  // namespace svimx
