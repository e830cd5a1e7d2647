// reduce.hpp -- deterministic wavefront / workgroup reductions (no atomics).
#pragma once
#include "dev.hpp"

namespace rbmd {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum NV values over the workgroup (NW wavefronts) and write them as one row
// of NV doubles (fixed-order sums: wavefront shuffle tree, then wavefronts in
// order).  Every thread of the workgroup must call this.
template <int NV, int NW = kBlock / kWave>
__device__ __forceinline__ void block_write_row(const double (&vals)[NV], double* row) {
  __shared__ double lds[NW][NV];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double t = wave_sum(vals[k]);
    if (lane == 0) lds[wave][k] = t;
  }
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    double acc = 0.0;
    for (int w = 0; w < NW; ++w) acc += lds[w][threadIdx.x];
    row[threadIdx.x] = acc;
  }
}

// BinPointsWorklet (ContPointLocator.cxx:38-66): ijk = trunc((x - min) / dxdydz),
// clamped to [0, dims-1]; a NaN coordinate lands in cell 0 (and is flagged elsewhere)
__device__ __forceinline__ int bin_axis(double x, double lo, double d, int dims) {
  double t = (x - lo) / d;
  int k = 0;
  if (t == t) k = (int)fmax(fmin(t, 2.0e9), -2.0e9);  // trunc toward zero like the Vec3f->Id3 cast
  k = k < 0 ? 0 : k;
  k = k > dims - 1 ? dims - 1 : k;
  return k;
}

// Count one atom per thread into cell_count.  With few cells the workgroup
// first aggregates in LDS, so each cell takes one global atomic per workgroup
// instead of one per atom.  Every thread of the workgroup must call this
// (cell < 0: nothing to count).
// count_cells for three cells per thread (-1: none)
__device__ __forceinline__ void count_cells3(const SysView& s, const int (&cell)[3], int* lds) {
  if (s.lds_bins) {
    for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) lds[c] = 0;
    __syncthreads();
    for (int t = 0; t < 3; ++t)
      if (cell[t] >= 0) atomicAdd(&lds[cell[t]], 1);
    __syncthreads();
    for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) {
      int v = lds[c];
      if (v) atomicAdd(&s.cell_count[c], v);
    }
  } else {
    for (int t = 0; t < 3; ++t)
      if (cell[t] >= 0) atomicAdd(&s.cell_count[cell[t]], 1);
  }
}

__device__ __forceinline__ void count_cells(const SysView& s, int cell, int* lds) {
  if (s.lds_bins) {
    for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) lds[c] = 0;
    __syncthreads();
    if (cell >= 0) atomicAdd(&lds[cell], 1);
    __syncthreads();
    for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) {
      int v = lds[c];
      if (v) atomicAdd(&s.cell_count[c], v);
    }
  } else if (cell >= 0) {
    atomicAdd(&s.cell_count[cell], 1);
  }
}

}  // namespace rbmd
