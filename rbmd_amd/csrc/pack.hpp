// pack.hpp -- device side of packed launches (see PackLaunch in dev.hpp).
#pragma once
#include "dev.hpp"

namespace rbmd {

// Packed kernels take the table as separate restrict-qualified parameters,
// so the compiler may read the (never written) table with scalar loads.
#define PACK_PARAMS const PackSys* __restrict__ pk_sys, const int* __restrict__ pk_map, const int* __restrict__ pk_first
#define PACK_ARGS(L) (L).sys, (L).map, (L).first

// the system of this block and the block's index within that system's grid
#define PACK_ENTRY(e, bid)                              \
  const int pk_k_ = pk_map[blockIdx.x];                 \
  const int bid = (int)blockIdx.x - pk_first[pk_k_];    \
  const PackSys& e = pk_sys[pk_k_];

// number of blocks of this block's system (its own grid size)
#define PACK_GRID() (pk_first[pk_k_ + 1] - pk_first[pk_k_])
#define PACK_INDEX() (pk_k_)

}  // namespace rbmd
