// FFT evaluation of N_r (see fft.hip): plan state and entry points used by
// the N_r plan in nr.hip.
#pragma once

#include "common.h"

namespace svimx {

// LDS slots per line tile (28 KB): lines * fft_pitch(L), 8 lines of 432 points;
// with 512 threads per tile three tiles share a CU (16 lines of 1024 threads,
// one tile per CU, were 12% slower)
constexpr int kFftTileElems = 3584;
constexpr int kFftThreads = 512;
// butterflies per thread in a radix-R stage of a full tile
__host__ __device__ constexpr int fft_kb(int R) { return (kFftTileElems / R + kFftThreads - 1) / kFftThreads; }

struct FftPlan {
  Dims d{0, 0, 0};
  int len[3] = {0, 0, 0};
  int nzh = 0;               // kept z frequencies, nz / 2 + 1
  float2* C = nullptr;       // complex work array, nx*ny*nzh (the half spectrum)
  float* khat = nullptr;     // real spectrum of the ball kernel, nx*ny*nzh
  float2* tw[3] = {nullptr, nullptr, nullptr};  // twiddles exp(-2 pi i t / L) per axis
};

int fft_plan_create(FftPlan& F, const double* weights, int m, Dims d);
void fft_plan_destroy(FftPlan& F);
// N_r = 1 + negbv * scale * conv(f, K) at every node (dense nr_out).
int fft_nr(hipStream_t st, const FftPlan& F, const float* occupancy, float* nr_out, float scale, float negbv);

}  // namespace svimx
