// N_r by FFT: the reference's method (svierosion/svi.py:66-93), for balls too
// wide for the direct summation and as a selectable alternative.
//
// N_r = 1 - h^3 (w * f) / |B_r| with the circular convolution computed as
//     conv = IFFT(FFT(f) * K^),   K = ball weights at wrapped offsets.
// K is real and even (the weights depend on |dx|, |dy|, |dz|), so K^ is real,
// stored as one float per kept frequency.  The inverse transform is a forward
// one on the conjugate: conv = conj(FFT(conj(FFT(f) K^))) / N, and only
// Re(conv) is used, so six forward line passes do the whole evaluation.
//
// f and conv are real, so only the z frequencies 0 .. nz/2 are kept (the rest
// follow by Hermitian symmetry): the work array holds nx * ny * (nz/2 + 1)
// complex values.
//   z, forward: two real z-lines a, b form one complex line a + i b; from its
//     transform Z the half spectra are F_a[k] = (Z[k] + conj Z[N-k]) / 2 and
//     F_b[k] = (Z[k] - conj Z[N-k]) / 2i, k = 0 .. nz/2;
//   y, x, forward (epilogue of x: conj(. K^)), then x, y again, on the half
//     spectrum;
//   z, last: the two half lines Q_a, Q_b are extended by Q[N-k] = conj Q[k]
//     into one complex line Q_a + i Q_b, whose transform is N (r_a + i r_b)
//     with the real results r_a, r_b of the two lines.
// Half the bytes of a full complex pipeline and half the line transforms of
// the y and x passes.
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


enum FftIn { kInComplex = 0 };
enum FftOut { kOutComplex = 0, kOutConjMulK = 1, kOutRealK = 2 };

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
  float coef;      // last z pass: negbv * h^3 / N
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
    const float2 v = data[off];
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
      case kOutRealK: khat_out[off] = v.x; break;
      default: data[off] = v;
    }
  }
}

// The z passes over pairs of z-lines: complex tile line l holds z-lines
// a = 2 (p0 + l) and b = a + 1 (b = nx ny is a zero line).  FWD: the real
// lines come from the occupancy (occ) or a dense real array (rin) and their
// half spectra go to `half` (row pitch nzh); else the half lines come from
// `half`, are extended by Hermitian symmetry, and N_r = 1 + coef Re / Im goes
// to `out` for lines a / b.
template <bool FWD>
__global__ void __launch_bounds__(kFftThreads) k_fft_zpair(FftPass P, int nzh, float2* __restrict__ half,
                                                           const float* __restrict__ occ,
                                                           const float* __restrict__ rin, float* __restrict__ out) {
  extern __shared__ float2 lds[];
  const int L = P.L, T = P.T, pitch = fft_pitch(L);
  float2* A = lds;
  float2* TW = lds + T * pitch;
  for (int t = threadIdx.x; t < L; t += blockDim.x) TW[t] = P.tw[t];
  const int64_t nz_lines = (int64_t)P.d.nx * P.d.ny;
  const int64_t p0 = (int64_t)blockIdx.x * T;
  const int tl = (int)min<int64_t>(T, P.nlines - p0);
  for (int e = threadIdx.x; e < tl * L; e += blockDim.x) {
    const int l = e / L, t = e - l * L;
    const int64_t a = 2 * (p0 + l), b = a + 1;
    const bool hb = b < nz_lines;
    float2 v;
    if constexpr (FWD) {
      float fa, fb = 0.0f;
      if (occ) {
        const uint32_t xa = (uint32_t)a / (uint32_t)P.d.ny, ya = (uint32_t)a - xa * (uint32_t)P.d.ny;
        fa = occ[P.OL.at((int)xa, (int)ya, t)];
        if (hb) {
          const uint32_t xb = (uint32_t)b / (uint32_t)P.d.ny, yb = (uint32_t)b - xb * (uint32_t)P.d.ny;
          fb = occ[P.OL.at((int)xb, (int)yb, t)];
        }
      } else {
        fa = rin[a * L + t];
        if (hb) fb = rin[b * L + t];
      }
      v = make_float2(fa, fb);
    } else {
      const bool low = t < nzh;
      const int k = low ? t : L - t;
      const float2 qa = half[a * nzh + k];
      const float2 qb = hb ? half[b * nzh + k] : make_float2(0.0f, 0.0f);
      // Q_a + i Q_b (k = t), or conj(Q_a) + i conj(Q_b) (k = N - t)
      v = low ? make_float2(qa.x - qb.y, qa.y + qb.x) : make_float2(qa.x + qb.y, qb.x - qa.y);
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
  if constexpr (FWD) {
    for (int e = threadIdx.x; e < tl * nzh; e += blockDim.x) {
      const int l = e / nzh, k = e - l * nzh;
      const int64_t a = 2 * (p0 + l), b = a + 1;
      const float2 z = A[l * pitch + k], w = A[l * pitch + (k ? L - k : 0)];
      half[a * nzh + k] = make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
      if (b < nz_lines) half[b * nzh + k] = make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x));
    }
  } else {
    for (int e = threadIdx.x; e < tl * L; e += blockDim.x) {
      const int l = e / L, t = e - l * L;
      const int64_t a = 2 * (p0 + l), b = a + 1;
      const float2 v = A[l * pitch + t];
      out[a * L + t] = 1.0f + P.coef * v.x;
      if (b < nz_lines) out[b * L + t] = 1.0f + P.coef * v.y;
    }
  }
}


// ball weights at wrapped offsets: K[(dx mod nx, dy mod ny, dz mod nz)] = w(dx, dy, dz)
__global__ void k_fft_place_kernel(float* K, const float* __restrict__ w, int m, Dims d) {
  const int W = 2 * m + 1;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)W * W * W) return;
  const int c = (int)(i % W), b = (int)((i / W) % W), a = (int)(i / ((int64_t)W * W));
  const int x = ((a - m) % d.nx + d.nx) % d.nx, y = ((b - m) % d.ny + d.ny) % d.ny, z = ((c - m) % d.nz + d.nz) % d.nz;
  K[((int64_t)x * d.ny + y) * d.nz + z] = w[i];
}

__global__ void k_fft_zero_real(float* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
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

// Pass over the lines of one axis.  y and x passes run on the half spectrum
// (dims nx, ny, nzh); the z passes go over pairs of z-lines of the full grid.
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
  const Dims dh{d.nx, d.ny, F.nzh};
  if (axis == 2) {
    P.d = d;
    P.nlines = ((int64_t)d.nx * d.ny + 1) / 2;  // complex lines = pairs of z-lines
  } else {
    P.d = dh;
    P.nlines = axis == 1 ? (int64_t)dh.nx * dh.nz : (int64_t)dh.ny * dh.nz;
  }
  P.tw = F.tw[axis];
  P.OL = OccLayout::of(d.nx, d.ny, d.nz);
  P.in_mode = in_mode;
  P.out_mode = out_mode;
  P.coef = coef;
  return P;
}

static size_t pass_lds(const FftPass& P) {
  return sizeof(float2) * ((size_t)P.T * fft_pitch(P.L) + (size_t)P.L);
}

static int run_pass(hipStream_t st, const FftPass& P, float2* data, const float* occ, const float* khat, float* out,
                    float* khat_out) {
  const unsigned grid = (unsigned)((P.nlines + P.T - 1) / P.T);
  hipLaunchKernelGGL(k_fft_lines, dim3(grid), dim3(kFftThreads), pass_lds(P), st, P, data, occ, khat, out, khat_out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

// z pass over line pairs: forward from occ (occupancy layout) or rin (dense
// real) into the half spectrum, or the last pass from it into out.
static int run_zpair(hipStream_t st, const FftPlan& F, bool fwd, const float* occ, const float* rin, float* out,
                     float coef) {
  const FftPass P = make_pass(F, 2, 0, 0, coef);
  const unsigned grid = (unsigned)((P.nlines + P.T - 1) / P.T);
  if (fwd)
    hipLaunchKernelGGL(k_fft_zpair<true>, dim3(grid), dim3(kFftThreads), pass_lds(P), st, P, F.nzh, F.C, occ, rin,
                       out);
  else
    hipLaunchKernelGGL(k_fft_zpair<false>, dim3(grid), dim3(kFftThreads), pass_lds(P), st, P, F.nzh, F.C, occ, rin,
                       out);
  SVIMX_CHECK_LAUNCH();
  return SVIMX_OK;
}

int fft_plan_create(FftPlan& F, const double* weights, int m, Dims d) {
  F = FftPlan{};
  F.d = d;
  F.len[0] = d.nx;
  F.len[1] = d.ny;
  F.len[2] = d.nz;
  F.nzh = d.nz / 2 + 1;
  for (int a = 0; a < 3; ++a) {
    SVIMX_REQUIRE(F.len[a] >= 1 && fft_pitch(F.len[a]) <= kFftTileElems && !fft_radices(F.len[a]).empty(),
                  SVIMX_EBADARG, "FFT needs grid extents that are products of 2, 3, 5, 7 and below %d (got %d)",
                  kFftTileElems - 1,
                  F.len[a]);
  }
  const int64_t n = d.size();
  const int64_t nh = (int64_t)d.nx * d.ny * F.nzh;
  SVIMX_CHECK_HIP(hipMalloc(&F.C, sizeof(float2) * nh));
  SVIMX_CHECK_HIP(hipMalloc(&F.khat, sizeof(float) * nh));
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
  // K^ = Re FFT(K) on the half spectrum: place the weights in a dense real
  // array, three passes, keep the real part
  const int W = 2 * m + 1;
  std::vector<float> wf((size_t)W * W * W);
  for (size_t i = 0; i < wf.size(); ++i) wf[i] = (float)weights[i];
  float* dw = nullptr;
  float* K = nullptr;
  hipError_t e = hipMalloc(&dw, sizeof(float) * wf.size());
  if (e == hipSuccess) e = hipMalloc(&K, sizeof(float) * n);
  if (e == hipSuccess) e = hipMemcpy(dw, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(dw);
    hipFree(K);
    SVIMX_CHECK_HIP(e);
  }
  hipStream_t st = nullptr;
  hipLaunchKernelGGL(k_fft_zero_real, dim3(blocks_for(n, 256)), dim3(256), 0, st, K, n);
  hipLaunchKernelGGL(k_fft_place_kernel, dim3(blocks_for((int64_t)W * W * W, 256)), dim3(256), 0, st, K, dw, m, d);
  int rc = run_zpair(st, F, true, nullptr, K, nullptr, 0.0f);
  if (!rc) rc = run_pass(st, make_pass(F, 1, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 0, kInComplex, kOutRealK, 0.0f), F.C, nullptr, nullptr, nullptr, F.khat);
  e = hipStreamSynchronize(st);
  hipFree(dw);
  hipFree(K);
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
  int rc = run_zpair(st, F, true, occupancy, nullptr, nullptr, 0.0f);
  if (!rc) rc = run_pass(st, make_pass(F, 1, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 0, kInComplex, kOutConjMulK, 0.0f), F.C, nullptr, F.khat, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 0, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_pass(st, make_pass(F, 1, kInComplex, kOutComplex, 0.0f), F.C, nullptr, nullptr, nullptr, nullptr);
  if (!rc) rc = run_zpair(st, F, false, nullptr, nullptr, nr_out, coef);
  return rc;
}

}  // namespace svimx
