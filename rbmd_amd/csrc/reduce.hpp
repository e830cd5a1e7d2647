// reduce.hpp -- deterministic wavefront / workgroup reductions (no atomics).
#pragma once
#include "dev.hpp"

namespace rbmd {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum NV values over the workgroup (kBlock threads) and write one row of
// S_NSLOT doubles: the NV sums go to `slots`, every other slot gets zero.
template <int NV>
__device__ __forceinline__ void block_write_row(const double (&vals)[NV], const int (&slots)[NV], double* row) {
  __shared__ double lds[kBlock / kWave][NV > 0 ? NV : 1];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = wave_sum(vals[k]);
    if (lane == 0) lds[wave][k] = t;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < S_NSLOT; s += blockDim.x) {
    double out = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (slots[k] == s) {
        double acc = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) acc += lds[w][k];
        out = acc;
      }
    row[s] = out;
  }
}

}  // namespace rbmd
