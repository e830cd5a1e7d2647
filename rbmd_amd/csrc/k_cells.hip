// k_cells.hip -- cell list (integer counting sort) and neighbour lists.
//
// Reference: ContPointLocator::Build (framework/locator/ContPointLocator.cxx:70-92)
// bins with BinPointsWorklet (:38-66) and sorts atom ids by cell with a
// comparison SortByKey.  Here binning is a counting sort: per-cell histogram
// with atomics, one exclusive scan, an atomic scatter, then a per-cell rank
// sort that restores the canonical within-cell order (ascending atom id), so
// the result is bit-exact and independent of atomic ordering.
//
// Neighbour lists follow ComputeNeighboursWorklet (RunWorklet.cpp:296-355) and
// ComputeRBLNeighboursOnceWorklet (:906-980): the (2nc+1)^3 stencil around
// InWhichCell(p_i) in x-outer / z-inner order, cells in sorted order, accept
// when rc^2 - d^2 > 1e-4 and d^2 > 1e-4 (RBL: core d < rs, every pice_num-th
// shell atom).  One wavefront owns 64 i-atoms of a cell; the j loop is
// wave-uniform, so j positions come through the scalar cache (s_load) and are
// broadcast to all lanes.  Lists are column-major [slot][n] so that the pair
// kernel's per-lane reads of slot k are coalesced.
#include "dev.hpp"

namespace rbmd {

__global__ void k_fill_int(int* p, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_double(double* p, double v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_int(int* p, int v, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, p, v, n);
}
void launch_fill_double(double* p, double v, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_double, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, p, v, n);
}

// BinPointsWorklet: ijk = trunc((x - min) / dxdydz), clamped to [0, dims-1]
__device__ __forceinline__ int bin_axis(double x, double lo, double d, int dims) {
  double t = (x - lo) / d;
  int k = 0;
  if (t == t) k = (int)fmax(fmin(t, 2.0e9), -2.0e9);  // trunc toward zero like the Vec3f->Id3 cast
  k = k < 0 ? 0 : k;
  k = k > dims - 1 ? dims - 1 : k;
  return k;
}

// InWhichCell (ExecPointLocator.h:271-274): unclamped trunc
__device__ __forceinline__ int cell_axis(double x, double lo, double d) {
  double t = (x - lo) / d;
  if (!(t == t)) return 0;
  return (int)fmax(fmin(t, 2.0e9), -2.0e9);
}

__global__ __launch_bounds__(kBlock) void k_bin(SysView s, const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  double px = x[i], py = y[i], pz = z[i];
  if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) atomicOr(s.status, 1);
  int a = bin_axis(px, s.lo[0], s.dxdydz[0], s.dims[0]);
  int b = bin_axis(py, s.lo[1], s.dxdydz[1], s.dims[1]);
  int c = bin_axis(pz, s.lo[2], s.dxdydz[2], s.dims[2]);
  int cell = a + b * s.dims[0] + c * s.dims[0] * s.dims[1];
  s.cell_of[i] = cell;
  atomicAdd(&s.cell_count[cell], 1);
}

void launch_bin(const SysView& s, hipStream_t st, const double* x, const double* y, const double* z) {
  hipLaunchKernelGGL(k_bin, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, x, y, z);
}

// block-wide exclusive scan of one int per thread (1024 threads = 16 waves)
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds_wave, int* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  int inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int t = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += t;
  }
  if (lane == kWave - 1) lds_wave[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    int nw = blockDim.x / kWave;
    for (int w = 0; w < nw; ++w) {
      int t = lds_wave[w];
      lds_wave[w] = acc;
      acc += t;
    }
    lds_wave[16] = acc;
  }
  __syncthreads();
  int ex = inc - v + lds_wave[wave];
  *total = lds_wave[16];
  __syncthreads();
  return ex;
}

// one workgroup: cell_start = exclusive scan of counts; cursor = cell_start;
// work items (cell, 64-atom chunk); counts reset to zero for the next binning
__global__ __launch_bounds__(1024) void k_scan(SysView s) {
  __shared__ int lds[17];
  int carry = 0, icarry = 0;
  for (int base = 0; base < s.ncell; base += blockDim.x) {
    int c = base + threadIdx.x;
    int cnt = c < s.ncell ? s.cell_count[c] : 0;
    int nch = (cnt + kChunk - 1) / kChunk;
    int tot, itot;
    int ex = block_exclusive_scan(cnt, lds, &tot);
    int iex = block_exclusive_scan(nch, lds, &itot);
    if (c < s.ncell) {
      s.cell_start[c] = carry + ex;
      s.cursor[c] = carry + ex;
      s.cell_count[c] = 0;
      for (int k = 0; k < nch; ++k) {
        int it = icarry + iex + k;
        if (it < s.n_items) s.items[it] = (c << 8) | (k & 255);
      }
      if (nch > 256) atomicOr(s.status, 4);
    }
    carry += tot;
    icarry += itot;
  }
  if (threadIdx.x == 0) {
    s.cell_start[s.ncell] = carry;
    s.n_items_dev[0] = icarry < s.n_items ? icarry : s.n_items;
    if (icarry > s.n_items) atomicOr(s.status, 4);
  }
}

void launch_scan(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, s);
}

__global__ __launch_bounds__(kBlock) void k_scatter(SysView s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  int pos = atomicAdd(&s.cursor[s.cell_of[i]], 1);
  s.point_ids[pos] = i;
}

void launch_scatter(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s);
}

constexpr int kSortCap = 4096;

// one workgroup per cell: rank each id by the number of smaller ids in the cell
__global__ __launch_bounds__(kBlock) void k_cell_sort(SysView s) {
  __shared__ int ids[kSortCap];
  int c = blockIdx.x;
  int start = s.cell_start[c];
  int cnt = s.cell_start[c + 1] - start;
  if (cnt > kSortCap) {
    if (threadIdx.x == 0) atomicOr(s.status, 4);
    return;
  }
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) ids[t] = s.point_ids[start + t];
  __syncthreads();
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    int me = ids[t];
    int r = 0;
    for (int k = 0; k < cnt; ++k) r += ids[k] < me;
    s.point_ids[start + r] = me;
    s.rank[me] = start + r;
  }
}

void launch_cell_sort(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_sort, dim3(s.ncell), dim3(kBlock), 0, st, s);
}

// sorted copies of positions / charges / types; misfit = current position is in
// a different cell than the binned one (possible after SHAKE moves atoms, or at
// x == hi where binning clamps and InWhichCell does not)
__global__ __launch_bounds__(kBlock) void k_gather(SysView s) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= s.n) return;
  int id = s.point_ids[p];
  double px = s.x[id], py = s.y[id], pz = s.z[id];
  s.spos[p] = make_double4(px, py, pz, s.q[id]);
  s.stype[p] = s.type[id];
  int cell = s.cell_of[id];
  int a = cell % s.dims[0];
  int b = (cell / s.dims[0]) % s.dims[1];
  int c = cell / (s.dims[0] * s.dims[1]);
  int ua = cell_axis(px, s.lo[0], s.dxdydz[0]);
  int ub = cell_axis(py, s.lo[1], s.dxdydz[1]);
  int uc = cell_axis(pz, s.lo[2], s.dxdydz[2]);
  s.misfit[p] = (ua != a || ub != b || uc != c) ? 1 : 0;
}

void launch_gather(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s);
}

__device__ __forceinline__ int wrap_axis(int v, int dims) { return (v + dims) % dims; }

__device__ __forceinline__ int pack_entry(int j, int sa, int sb, int sc) {
  return (j << kShiftBits) | ((sa + 3) | ((sb + 3) << 3) | ((sc + 3) << 6));
}

// Record one accepted candidate for lane i.  VERLET: slot = count.  RBL: core
// slots [0, rs_num), kept shell slots from rs_num (every pice_num-th shell atom).
struct ListState {
  int cnt, core, shell_seen, shell_kept;
};

template <bool RBL>
__device__ __forceinline__ void consider(const SysView& s, ListState& L, int i, int entry, double d2) {
  if (!RBL) {
    if (s.cut2 - d2 > 0.0001 && d2 > 0.0001) {
      if (L.cnt < s.cap)
        s.nbr[(size_t)L.cnt * s.n + i] = entry;
      else
        atomicOr(s.status, 4);
      L.cnt++;
    }
  } else {
    if (s.cut2 - d2 > 0.0001 && d2 - s.rs2 > 0.0001) {
      if (L.shell_seen % s.pice_num == 0) {
        int slot = s.rs_num + L.shell_kept;
        if (slot < s.cap)
          s.nbr[(size_t)slot * s.n + i] = entry;
        else
          atomicOr(s.status, 4);
        L.shell_kept++;
      }
      L.shell_seen++;
    } else if (s.rs2 - d2 > 0.0001 && d2 > 0.0001) {
      if (L.core < s.rs_num)
        s.nbr[(size_t)L.core * s.n + i] = entry;
      else
        atomicOr(s.status, 4);
      L.core++;
    }
  }
}

template <bool RBL>
__device__ __forceinline__ void finish(const SysView& s, const ListState& L, int i) {
  if (!RBL) {
    s.nbr_num[i] = L.cnt < s.cap ? L.cnt : s.cap;
  } else {
    s.nbr_num[2 * i] = L.core < s.rs_num ? L.core : s.rs_num;
    int kept = L.shell_kept;
    if (s.rs_num + kept > s.cap) kept = s.cap - s.rs_num;
    s.nbr_num[2 * i + 1] = kept;
  }
}

// one wavefront per work item (64 i-atoms of one cell), 4 items per workgroup
template <bool RBL>
__global__ __launch_bounds__(kBlock) void k_nbr_build(SysView s) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  const int item_idx = blockIdx.x * (kBlock / kWave) + wave;
  if (item_idx >= s.n_items_dev[0]) return;
  const int item = __builtin_amdgcn_readfirstlane(s.items[item_idx]);
  const int c = item >> 8;
  const int chunk = item & 255;
  const int start = s.cell_start[c];
  const int cnt = s.cell_start[c + 1] - start;
  const int local = chunk * kChunk + lane;
  const int i = start + local;
  bool active = local < cnt;
  double4 pi = make_double4(0, 0, 0, 0);
  if (active) {
    pi = s.spos[i];
    if (s.misfit[i]) active = false;  // handled by k_nbr_misfit
  }
  const int ca = c % s.dims[0];
  const int cb = (c / s.dims[0]) % s.dims[1];
  const int cc = c / (s.dims[0] * s.dims[1]);
  ListState L = {0, 0, 0, 0};
  const int nc = s.nc;
  for (int a = -nc; a <= nc; ++a) {
    for (int b = -nc; b <= nc; ++b) {
      for (int e = -nc; e <= nc; ++e) {
        int na = ca + a, nb = cb + b, ne = cc + e;
        int wa = wrap_axis(na, s.dims[0]), wb = wrap_axis(nb, s.dims[1]), we = wrap_axis(ne, s.dims[2]);
        if (wa < 0 || wb < 0 || we < 0) {  // stencil wider than two boxes: unsupported geometry
          if (lane == 0) atomicOr(s.status, 4);
          continue;
        }
        double ox = s.dxdydz[0] * (double)(wa - na);
        double oy = s.dxdydz[1] * (double)(wb - nb);
        double oz = s.dxdydz[2] * (double)(we - ne);
        int sa = (wa - na) / s.dims[0], sb = (wb - nb) / s.dims[1], sc = (we - ne) / s.dims[2];
        int wcell = wa + wb * s.dims[0] + we * s.dims[0] * s.dims[1];
        int js = s.cell_start[wcell], je = s.cell_start[wcell + 1];
        for (int j = js; j < je; ++j) {
          double4 pj = s.spos[j];  // wave-uniform address -> scalar load
          double rx = (pj.x - ox) - pi.x;
          double ry = (pj.y - oy) - pi.y;
          double rz = (pj.z - oz) - pi.z;
          double d2 = rx * rx + ry * ry + rz * rz;
          if (active) consider<RBL>(s, L, i, pack_entry(j, sa, sb, sc), d2);
        }
      }
    }
  }
  if (local < cnt && active) finish<RBL>(s, L, i);
}

// atoms whose current cell differs from the binned one: own stencil, one thread each
template <bool RBL>
__global__ __launch_bounds__(kBlock) void k_nbr_misfit(SysView s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n || !s.misfit[i]) return;
  double4 pi = s.spos[i];
  int ca = cell_axis(pi.x, s.lo[0], s.dxdydz[0]);
  int cb = cell_axis(pi.y, s.lo[1], s.dxdydz[1]);
  int cc = cell_axis(pi.z, s.lo[2], s.dxdydz[2]);
  ListState L = {0, 0, 0, 0};
  const int nc = s.nc;
  for (int a = -nc; a <= nc; ++a)
    for (int b = -nc; b <= nc; ++b)
      for (int e = -nc; e <= nc; ++e) {
        int na = ca + a, nb = cb + b, ne = cc + e;
        int wa = wrap_axis(na, s.dims[0]), wb = wrap_axis(nb, s.dims[1]), we = wrap_axis(ne, s.dims[2]);
        if (wa < 0 || wb < 0 || we < 0) {  // far outside the box: PeriodicIndexOffset assumes >= -dims
          atomicOr(s.status, 1);
          continue;
        }
        double ox = s.dxdydz[0] * (double)(wa - na);
        double oy = s.dxdydz[1] * (double)(wb - nb);
        double oz = s.dxdydz[2] * (double)(we - ne);
        int sa = (wa - na) / s.dims[0], sb = (wb - nb) / s.dims[1], sc = (we - ne) / s.dims[2];
        if (sa < -3 || sa > 3 || sb < -3 || sb > 3 || sc < -3 || sc > 3) {
          atomicOr(s.status, 1);
          continue;
        }
        int wcell = wa + wb * s.dims[0] + we * s.dims[0] * s.dims[1];
        int js = s.cell_start[wcell], je = s.cell_start[wcell + 1];
        for (int j = js; j < je; ++j) {
          double4 pj = s.spos[j];
          double rx = (pj.x - ox) - pi.x;
          double ry = (pj.y - oy) - pi.y;
          double rz = (pj.z - oz) - pi.z;
          double d2 = rx * rx + ry * ry + rz * rz;
          consider<RBL>(s, L, i, pack_entry(j, sa, sb, sc), d2);
        }
      }
  finish<RBL>(s, L, i);
}

void launch_nbr_build(const SysView& s, hipStream_t st) {
  int grid = (s.n_items + (kBlock / kWave) - 1) / (kBlock / kWave);
  int gm = (s.n + kBlock - 1) / kBlock;
  if (s.rbl) {
    hipLaunchKernelGGL(k_nbr_build<true>, dim3(grid), dim3(kBlock), 0, st, s);
    hipLaunchKernelGGL(k_nbr_misfit<true>, dim3(gm), dim3(kBlock), 0, st, s);
  } else {
    hipLaunchKernelGGL(k_nbr_build<false>, dim3(grid), dim3(kBlock), 0, st, s);
    hipLaunchKernelGGL(k_nbr_misfit<false>, dim3(gm), dim3(kBlock), 0, st, s);
  }
}

}  // namespace rbmd
