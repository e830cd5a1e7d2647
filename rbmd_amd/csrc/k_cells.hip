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
// shell atom), with d^2 rounded as the oracle rounds it (dist2).  One
// wavefront owns 64 i-atoms of a cell and one stencil column (VERLET) or
// the whole stencil (RBL); candidates are loaded coalesced and broadcast
// through LDS.  Lists are row-major (one row of segments per atom), so the
// pair kernel's lanes of one atom read consecutive entries.
#include "dev.hpp"
#include "pack.hpp"
#include "reduce.hpp"
#include "cells.hpp"

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

__global__ __launch_bounds__(kBlock) void k_bin(SysView s, const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z) {
  extern __shared__ int lds_cells[];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cell = -1;
  if (i < s.n) {
    double px = x[i], py = y[i], pz = z[i];
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) atomicOr(s.status, 1);
    int a = bin_axis(px, s.lo[0], s.dxdydz[0], s.dims[0]);
    int b = bin_axis(py, s.lo[1], s.dxdydz[1], s.dims[1]);
    int c = bin_axis(pz, s.lo[2], s.dxdydz[2], s.dims[2]);
    cell = a + b * s.dims[0] + c * s.dims[0] * s.dims[1];
    s.cell_of[i] = cell;
  }
  count_cells(s, cell, lds_cells);
}

size_t bin_lds_bytes(const SysView& s) { return s.lds_bins ? (size_t)s.ncell * sizeof(int) : 0; }

// Displacement of a coordinate since the last build, as the minimum image:
// image flags cannot unwrap it (SHAKE re-applies the periodic wrap to
// positions the drift already wrapped, so a flag can count one crossing twice,
// as the reference's does), and between builds every displacement is below
// skin / 2 plus one step's motion, far under half a box edge.
__device__ __forceinline__ double skin_disp(double x, double x0, double box) {
  const double d = x - x0;
  return d - box * rint(d / box);
}

// Skin lists: after the step's positions are final (after SHAKE), flag a
// rebuild when any atom is more than skin / 2 from its position at the last
// build (a non-finite distance counts as moved), and bin every atom with the
// current positions, as the rebuild's scan needs.  On a step without a
// rebuild the scan kernel clears the counts again.  One atomic per wavefront
// that saw a moved atom.
__global__ __launch_bounds__(kBlock) void k_skin_check(SysView s) {
  extern __shared__ int lds_cells[];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cell = -1;
  bool moved = false;
  if (i < s.n) {
    const double px = s.x[i], py = s.y[i], pz = s.z[i];
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) atomicOr(s.status, 1);
    const double ux = skin_disp(px, s.x0[i], s.box[0]);
    const double uy = skin_disp(py, s.y0[i], s.box[1]);
    const double uz = skin_disp(pz, s.z0[i], s.box[2]);
    moved = !(ux * ux + uy * uy + uz * uz <= s.half_skin2);
    const int a = bin_axis(px, s.lo[0], s.dxdydz[0], s.dims[0]);
    const int b = bin_axis(py, s.lo[1], s.dxdydz[1], s.dims[1]);
    const int c = bin_axis(pz, s.lo[2], s.dxdydz[2], s.dims[2]);
    cell = a + b * s.dims[0] + c * s.dims[0] * s.dims[1];
    s.cell_of[i] = cell;
  }
  if (__ballot(moved) != 0ull && (threadIdx.x & (kWave - 1)) == 0)
    atomicMax((unsigned long long*)s.rebuild_at, (unsigned long long)*s.cur_step);
  count_cells(s, cell, lds_cells);
}

void launch_skin_check(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_skin_check, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), bin_lds_bytes(s), st, s);
}

__global__ void k_skin_request(int64_t* rebuild_at, const int64_t* cur_step, int ahead) {
  if (threadIdx.x == 0) *rebuild_at = *cur_step + ahead;
}

void launch_skin_request(const SysView& s, int ahead, hipStream_t st) {
  hipLaunchKernelGGL(k_skin_request, dim3(1), dim3(kWave), 0, st, s.rebuild_at, s.cur_step, ahead);
}

// a step without a rebuild: the counts binned this step are dropped
__device__ __forceinline__ void clear_counts(const SysView& s) {
  for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) s.cell_count[c] = 0;
}

void launch_bin(const SysView& s, hipStream_t st, const double* x, const double* y, const double* z) {
  hipLaunchKernelGGL(k_bin, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), bin_lds_bytes(s), st, s, x, y, z);
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

// Exclusive scan of the cell counts by one workgroup.  With OUT it writes
// cell_start, the work items (cell, 64-atom chunk), the misfit count and the
// item count; with KERNEL (the scan kernel) also cursor = cell_start and the
// counts reset to zero for the next binning.  start (LDS, or null) receives
// every cell's start.
template <bool KERNEL>
__device__ __forceinline__ void cell_scan(const SysView& s, int* start, const bool out) {
  __shared__ int lds[17];
  int carry = 0, icarry = 0;
  for (int base = 0; base < s.ncell; base += blockDim.x) {
    int c = base + threadIdx.x;
    int cnt = c < s.ncell ? s.cell_count[c] : 0;
    int tot;
    int ex = block_exclusive_scan(cnt, lds, &tot);
    if (start && c < s.ncell) start[c] = carry + ex;
    if (out) {  // workgroup-uniform
      int nch = (cnt + kChunk - 1) / kChunk;
      int itot;
      int iex = block_exclusive_scan(nch, lds, &itot);
      if (c < s.ncell) {
        s.cell_start[c] = carry + ex;
        if (KERNEL) {
          s.cursor[c] = carry + ex;
          s.cell_count[c] = 0;
        }
        for (int k = 0; k < nch; ++k) {
          int it = icarry + iex + k;
          if (it < s.n_items) s.items[it] = (c << 8) | (k & 255);
        }
        if (nch > 256) atomicOr(s.status, 4);
      }
      icarry += itot;
    }
    carry += tot;
  }
  if (out && threadIdx.x == 0) {
    s.misfit_count[0] = 0;
    s.cell_start[s.ncell] = carry;
    s.n_items_dev[0] = icarry < s.n_items ? icarry : s.n_items;
    if (icarry > s.n_items) atomicOr(s.status, 4);
  }
}

__device__ __forceinline__ void scan_body(const SysView& s, const int bid) {
  if (!list_rebuild(s)) {
    clear_counts(s);
    return;
  }
  cell_scan<true>(s, nullptr, true);
}

__global__ __launch_bounds__(1024) void k_scan(SysView s) {
  scan_body(s, (int)blockIdx.x);
}

void launch_scan(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, s);
}

// Scatter atom ids into their cells.  With LDS aggregation each workgroup
// reserves one range per cell with a single global atomic and places its atoms
// by their LDS rank; the order inside a cell is fixed afterwards by k_cell_sort.
__device__ __forceinline__ void scatter_body(const SysView& s, const int bid) {
  extern __shared__ int lds[];
  if (!list_rebuild(s)) return;
  int i = bid * blockDim.x + threadIdx.x;
  int cell = i < s.n ? s.cell_of[i] : -1;
  if (!s.lds_bins) {
    if (cell >= 0) s.point_ids[atomicAdd(&s.cursor[cell], 1)] = i;
    return;
  }
  int* cnt = lds;
  int* base = lds + s.ncell;
  for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) cnt[c] = 0;
  __syncthreads();
  int local = cell >= 0 ? atomicAdd(&cnt[cell], 1) : 0;
  __syncthreads();
  for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) {
    int v = cnt[c];
    base[c] = v ? atomicAdd(&s.cursor[c], v) : 0;
  }
  __syncthreads();
  if (cell >= 0) s.point_ids[base[cell] + local] = i;
}

__global__ __launch_bounds__(kBlock) void k_scatter(SysView s) {
  scatter_body(s, (int)blockIdx.x);
}

void launch_scatter(const SysView& s, hipStream_t st) {
  size_t lds = s.lds_bins ? 2 * (size_t)s.ncell * sizeof(int) : 0;
  hipLaunchKernelGGL(k_scatter, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), lds, st, s);
}

// Scan and scatter in one launch, for few cells (scan_in_scatter): every
// workgroup scans the cell counts itself (block 0 also writes the scan's
// outputs) and reserves its range in each cell with one atomic on cursor,
// which starts at zero here (the cell sort resets it and the counts).  The
// order inside a cell is fixed afterwards by k_cell_sort, as with k_scatter.
__global__ __launch_bounds__(kBlock) void k_scan_scatter(SysView s) {
  extern __shared__ int lds[];
  int* start = lds;
  int* cnt = lds + s.ncell;
  int* base = cnt + s.ncell;
  const int bid = (int)blockIdx.x;
  if (!list_rebuild(s)) {
    if (bid == 0) clear_counts(s);
    return;
  }
  cell_scan<false>(s, start, bid == 0);
  const int i = bid * blockDim.x + threadIdx.x;
  const int cell = i < s.n ? s.cell_of[i] : -1;
  for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) cnt[c] = 0;
  __syncthreads();
  const int local = cell >= 0 ? atomicAdd(&cnt[cell], 1) : 0;
  __syncthreads();
  for (int c = threadIdx.x; c < s.ncell; c += blockDim.x) {
    const int v = cnt[c];
    base[c] = v ? start[c] + atomicAdd(&s.cursor[c], v) : 0;
  }
  __syncthreads();
  if (cell >= 0) s.point_ids[base[cell] + local] = i;
}

bool scan_in_scatter(const SysView& s) { return s.lds_bins && s.ncell <= kScanScatterCells; }

void launch_scan_scatter(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_scan_scatter, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock),
                     3 * (size_t)s.ncell * sizeof(int), st, s);
}

constexpr int kSortCap = 4096;

// One workgroup per cell: rank each id by the number of smaller ids in the
// cell, then write the atom's sorted copies at its sorted position p: x, y, z,
// q, type, and the misfit flag (the current position lies in a different
// cell than the binned one: possible after SHAKE moves atoms, or at x == hi,
// where binning clamps and InWhichCell does not).  Runs after the step's
// positions are final.
__device__ __forceinline__ void cell_sort_body(const SysView& s, const int bid) {
  __shared__ int ids[kSortCap];
  if (!list_rebuild(s)) {
    // skin list between builds: the build's sorted order is kept; the sorted
    // copy of each position is its build position plus its displacement, so
    // a list entry's image offset still applies
    const int stride = s.ncell * blockDim.x;
    for (int p = bid * blockDim.x + threadIdx.x; p < s.n; p += stride) {
      const int me = s.point_ids[p];
      const double x0 = s.x0[me], y0 = s.y0[me], z0 = s.z0[me];
      s.spos[p] = make_double4(x0 + skin_disp(s.x[me], x0, s.box[0]), y0 + skin_disp(s.y[me], y0, s.box[1]),
                               z0 + skin_disp(s.z[me], z0, s.box[2]), s.q[me]);
    }
    return;
  }
  int c = bid;
  int start = s.cell_start[c];
  int cnt = s.cell_start[c + 1] - start;
  if (threadIdx.x == 0) {  // read by the binning and scatter kernels before this one
    s.cell_count[c] = 0;
    s.cursor[c] = 0;
    if (c == 0 && s.list_builds) atomicAdd(s.list_builds, 1);
  }
  if (cnt > kSortCap) {
    if (threadIdx.x == 0) atomicOr(s.status, 4);
    return;
  }
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) ids[t] = s.point_ids[start + t];
  __syncthreads();
  const int a = c % s.dims[0];
  const int b = (c / s.dims[0]) % s.dims[1];
  const int e = c / (s.dims[0] * s.dims[1]);
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int me = ids[t];
    int r = 0;
    for (int k = 0; k < cnt; ++k) r += ids[k] < me;
    const int p = start + r;
    s.point_ids[p] = me;
    s.rank[me] = p;
    const double px = s.x[me], py = s.y[me], pz = s.z[me];
    s.spos[p] = make_double4(px, py, pz, s.q[me]);
    s.stype[p] = s.type[me];
    if (s.skin_on) {  // the build's positions, for the skin check and the gathers until the next build
      s.x0[me] = px;
      s.y0[me] = py;
      s.z0[me] = pz;
    }
    const int ua = cell_axis(px, s.lo[0], s.dxdydz[0]);
    const int ub = cell_axis(py, s.lo[1], s.dxdydz[1]);
    const int uc = cell_axis(pz, s.lo[2], s.dxdydz[2]);
    const int mf = (ua != a || ub != b || uc != e) ? 1 : 0;
    s.misfit[p] = mf;
    if (mf) s.misfit_list[atomicAdd(s.misfit_count, 1)] = p;
  }
}

__global__ __launch_bounds__(kBlock) void k_cell_sort(SysView s) {
  cell_sort_body(s, (int)blockIdx.x);
}

void launch_cell_sort(const SysView& s, hipStream_t st) {
  hipLaunchKernelGGL(k_cell_sort, dim3(s.ncell), dim3(kBlock), 0, st, s);
}

// VERLET list build.  Three workgroups per work item (64 i-atoms of one
// cell), three wavefronts each: wavefront w = 3 (a+1) + (b+1) of the nine owns
// the column (a, b) of the 3x3x3 stencil (its three cells e = -1, 0, 1) and
// writes list segment w, so the concatenated segments follow the reference's
// order (x offset outermost, then y, then z, then the cell's sorted order).
// Small workgroups spread evenly over the CUs: with one 9-wavefront
// workgroup per item, a 30,000-atom launch put three items on some CUs and
// two on others, and the fullest CUs set the kernel time.
//
// The three cells of a column form one candidate stream.  A batch of up to
// 128 candidates is loaded at once (image shift applied); candidates farther
// than the cutoff from the i-chunk's bounding box are dropped by a ballot and
// the rest staged in LDS in stream order.  Every lane then tests the staged
// candidates against its own i-atom (FP32, with the reference's FP64
// predicate near the thresholds) and appends each accepted entry to its own
// row, so the row stays in stream order.
//
// Workgroups past the work items take the misfit atoms (whose current cell is
// not the binned one, e.g. after SHAKE): one workgroup per misfit atom, the 9
// wavefronts again own the 9 columns, lanes test 64 candidates at a time and
// accepted ones are written at their ballot prefix position (stream order).
// wavefront `wave` (its stencil column) of misfit slot `slot`: misfit atoms
// slot, slot + kMisfitBlocks, ...
__device__ void misfit_block(const SysView& s, int wave, int lane, int slot) {
  const int count = s.misfit_count[0];
  const int a = wave / 3 - 1, b = wave % 3 - 1;
  const int seg0 = wave * s.seg_cap;
  for (int m = slot; m < count; m += kMisfitBlocks) {
    const int i = __builtin_amdgcn_readfirstlane(s.misfit_list[m]);
    const double4 pi = s.spos[i];
    const int ca = cell_axis(pi.x, s.lo[0], s.dxdydz[0]);
    const int cb = cell_axis(pi.y, s.lo[1], s.dxdydz[1]);
    const int cc = cell_axis(pi.z, s.lo[2], s.dxdydz[2]);
    Column col;
    make_column(s, ca, cb, cc, a, b, lane, col);
    const int total = col.total;
    int cnt = 0;
    for (int base = 0; base < total; base += kWave) {
      const int t = base + lane;
      bool ok = false;
      int entry = 0;
      if (t < total) {
        double4 pj;
        COLUMN_CANDIDATE(t, pj, entry);
        double rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
        double d2 = dist2(rx, ry, rz);
        ok = s.list_cut2 - d2 > 0.0001 && d2 > 0.0001;
      }
      const unsigned long long bal = __ballot(ok);
      if (ok) {
        int slot = cnt + __popcll(bal & ((1ull << lane) - 1ull));
        if (slot < s.seg_cap)
          s.nbr[nbr_index(seg0 + slot, i, s.cap)] = entry;
        else
          atomicOr(s.status, 4);
      }
      cnt += __popcll(bal);
    }
    if (lane == 0) s.nbr_num[wave * s.n + i] = cnt < s.seg_cap ? cnt : s.seg_cap;
  }
}

// One workgroup: kNbrWaves columns of one work item (or of one misfit slot).
// The grid is kNbrParts passes over the work items (and misfit slots); pass 0
// holds part 1 of every item -- the centre column and two edge columns, whose
// box culls keep the most candidates (about 2.8, 2.2 and 2.2 cells' worth
// against 1.4 for a corner column) -- so the longest-lived wavefronts are
// dispatched first; passes 1 and 2 hold parts 2 and 0.
static_assert(kNbrParts == 3, "the pass order below names three parts");
__device__ __forceinline__ void nbr_build_verlet_body(const SysView& s, const int bid) {
  __shared__ float4 stage[kNbrWaves][kStageCap];  // FP32 position from the chunk's box corner, entry bits in w
  __shared__ double stage64[kNbrWaves][kStageCap][3];  // the same candidates' shifted FP64 positions (re-tests only)
  __shared__ int ring[kNbrWaves][kWave * kListRingStride];  // each lane's last accepted entries
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  const int per_pass = s.n_items + kMisfitBlocks;
  const int pass = bid / per_pass;
  const int wid = bid - pass * per_pass;
  const int part = pass == kNbrParts - 1 ? 0 : pass + 1;
  const int wave = part * kNbrWaves + w;  // the stencil column this wavefront owns
  if (!list_rebuild(s)) return;  // skin list between builds
  if (wid >= s.n_items) {
    misfit_block(s, wave, lane, wid - s.n_items);
    return;
  }
  const int item_idx = wid;
  if (item_idx >= s.n_items_dev[0]) return;
  const int item = __builtin_amdgcn_readfirstlane(s.items[item_idx]);
  const int c = item >> 8;
  const int chunk = item & 255;
  const int start = s.cell_start[c];
  const int cnt_c = s.cell_start[c + 1] - start;
  const int local = chunk * kChunk + lane;
  const int i = start + local;
  bool active = local < cnt_c;
  double4 pi = make_double4(0, 0, 0, 0);
  if (active) {
    pi = s.spos[i];
    if (s.misfit[i]) active = false;  // a misfit workgroup lists it
  }
  // bounding box of the active i-atoms (wave reductions)
  double bl[3], bh[3];
  {
    double v[3] = {pi.x, pi.y, pi.z};
    for (int d = 0; d < 3; ++d) {
      double lo = active ? v[d] : 1e300, hi = active ? v[d] : -1e300;
      for (int off = 1; off < kWave; off <<= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, kWave));
        hi = fmax(hi, __shfl_xor(hi, off, kWave));
      }
      // wave-uniform: scalar registers
      bl[d] = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(lo)),
                               __builtin_amdgcn_readfirstlane(__double2loint(lo)));
      bh[d] = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(hi)),
                               __builtin_amdgcn_readfirstlane(__double2loint(hi)));
    }
  }
  const int ca = c % s.dims[0];
  const int cb = (c / s.dims[0]) % s.dims[1];
  const int cc = c / (s.dims[0] * s.dims[1]);
  const int a = wave / 3 - 1, b = wave % 3 - 1;
  const int seg0 = wave * s.seg_cap;
  const int seg_cap = s.seg_cap;
  int* row = s.nbr + nbr_index(seg0, active ? i : 0, s.cap);  // this lane's segment
  Column col;
  make_column(s, ca, cb, cc, a, b, lane, col);
  const int total = col.total;
  const double cut2 = s.list_cut2;  // rc^2, or (rc + skin)^2 for a skin list
  // FP32 test with an exact FP64 re-test near the thresholds.  Coordinates are
  // taken from the chunk's box corner bl: every staged candidate lies within
  // the list cutoff of the box, so each coordinate is below 3 rc in size and
  // carries at most 2^-24 * 3 rc of conversion error; with the FP32 difference
  // and sum of squares, |d2_f32 - d2_f64| < 1.7e-6 rc^2 where d2 is near rc^2.
  // Candidates within 1e-5 rc^2 of rc^2 - 1e-4, or at or below 1e-3, are
  // decided by the FP64 predicate of the per-candidate path (the reference's
  // rc^2 - d^2 > 1e-4 && d^2 > 1e-4 on the same rounded vector), all others by
  // the FP32 comparison, which then agrees with it.  An FP32 d^2 of exactly
  // 0 is rejected without a re-test: it means FP32-equal coordinates, which
  // differ by at most one FP32 ulp in FP64, so their FP64 d^2 is below
  // 3 (2^-23 * 3 rc)^2, far under 1e-4.  That is the atom itself (same cell,
  // no image shift), which every centre column meets once per i-atom.  A lane
  // without an i-atom sits at 1e30: its d^2 is +inf, and neither predicate
  // holds.
  const float cut2m_f = (float)(cut2 - 0.0001);
  const float band_f = (float)(1e-5 * cut2);
  // accepted in FP32: 1e-3 < d2 < cut2m - band, as one unsigned range test on
  // the bits (non-negative floats order as their bit patterns)
  const unsigned in_lo = __float_as_uint(1e-3f) + 1u;
  const unsigned in_w = __float_as_uint(cut2m_f - band_f) - in_lo;
  const float out_f = cut2m_f + band_f;  // rejected in FP32: d2 >= out_f
  // not rejected in FP32: 0 < d2 < out_f, as one unsigned range test (d2 = 0
  // wraps to the top)
  const unsigned nr_w = __float_as_uint(out_f) - 1u;
  const float lix = active ? (float)(pi.x - bl[0]) : 1e30f;
  const float liy = active ? (float)(pi.y - bl[1]) : 1e30f;
  const float liz = active ? (float)(pi.z - bl[2]) : 1e30f;
  // Accepted entries go through a ring of kListRing slots per lane in LDS
  // and leave as one 16-byte store per four entries (a segment starts on a
  // 64-byte boundary and holds a multiple of 16 slots), so the row is written
  // in whole 16-byte pieces rather than one scattered 4-byte store per entry.
  // Every lane writes every candidate's entry at its count and advances the
  // count only if it accepted, so a rejected entry is overwritten by the next
  // one: no select and no branch per candidate.  The lanes' rings are
  // kListRingStride words apart (odd), so a write of the whole wavefront at
  // equal counts hits distinct banks.
  int* const my = &ring[w][lane * kListRingStride];
  int cnt = 0;
  auto push = [&](bool ok, int e) {
    my[cnt & (kListRing - 1)] = e;
    cnt += ok ? 1 : 0;
  };
  // after at most four pushes since the count was cnt0 the count has passed
  // at most one multiple of four: the four entries before it are complete in
  // one half of the ring and go to the row.  Past the capacity the last four
  // slots are overwritten: the list is then flagged as overflowed (status 4)
  // and not used.
  auto flush = [&](int cnt0) {
    if ((cnt ^ cnt0) & ~3) {
      const int end = cnt & ~3;
      const int h = (end - 4) & (kListRing - 1);
      *reinterpret_cast<int4*>(row + (min(end, seg_cap) - 4)) = make_int4(my[h], my[h + 1], my[h + 2], my[h + 3]);
    }
  };
  // the FP32 predicate of one staged candidate: ok accepted, nx not rejected
  // (accepted, or decided in FP64 if not accepted)
  auto test = [&](const float4 q, bool& ok, bool& nx) {
    const float rx = q.x - lix, ry = q.y - liy, rz = q.z - liz;
    const float d2 = rx * rx + ry * ry + rz * rz;
    ok = __float_as_uint(d2) - in_lo < in_w;
    nx = __float_as_uint(d2) - 1u < nr_w;
  };
  // the per-candidate FP64 predicate on the build's own vector (the staged
  // shifted position; no global load in the test loop, where a load would
  // make every iteration wait for the row stores before it)
  auto exact = [&](int slot) -> bool {
    const double rx = stage64[w][slot][0] - pi.x;
    const double ry = stage64[w][slot][1] - pi.y;
    const double rz = stage64[w][slot][2] - pi.z;
    const double d2 = dist2(rx, ry, rz);
    return (cut2 - d2 > 0.0001) & (d2 > 0.0001);
  };
  // Batches of up to kStageRounds x 64 candidates: the batch's loads are
  // issued back to back, the survivors of the box cull are staged in stream
  // order, then every lane tests them all.  Loads and the entry stores are
  // counted by one in-order counter, so waiting for a load also waits for
  // every store issued before it: with whole batches there is one such wait
  // per batch, not one per 64 candidates.
  for (int base = 0; base < total; base += kStageRounds * kWave) {
    double4 raw[kStageRounds];
    double ozr[kStageRounds];
    int er[kStageRounds];
#pragma unroll
    for (int r = 0; r < kStageRounds; ++r) {  // straight-line loads (past the end: the last candidate again)
      const int t = base + r * kWave + lane;
      int e;
      raw[r] = s.spos[COLUMN_LOCATE(t < total ? t : total - 1, e, ozr[r])];
      er[r] = t < total ? e : -1;
    }
    int kc = 0;
#pragma unroll
    for (int r = 0; r < kStageRounds; ++r) {
      bool keep = false;
      double4 pj = make_double4(0, 0, 0, 0);
      if (er[r] >= 0) {
        pj = make_double4(raw[r].x - col.ox, raw[r].y - col.oy, raw[r].z - ozr[r], raw[r].w);
        double dx = fmax(0.0, fmax(bl[0] - pj.x, pj.x - bh[0]));
        double dy = fmax(0.0, fmax(bl[1] - pj.y, pj.y - bh[1]));
        double dz = fmax(0.0, fmax(bl[2] - pj.z, pj.z - bh[2]));
        keep = dx * dx + dy * dy + dz * dz < cut2;  // superset of every lane's acceptance
      }
      const unsigned long long bal = __ballot(keep);
      if (keep) {
        const int slot = kc + __popcll(bal & ((1ull << lane) - 1ull));
        stage[w][slot] = make_float4((float)(pj.x - bl[0]), (float)(pj.y - bl[1]), (float)(pj.z - bl[2]),
                                     __int_as_float(er[r]));
        stage64[w][slot][0] = pj.x;
        stage64[w][slot][1] = pj.y;
        stage64[w][slot][2] = pj.z;
      }
      kc += __popcll(bal);
    }
    if (kc == 0) continue;
    wave_sync();
    // every staged candidate is tested by every lane; an accepting lane
    // appends the candidate's entry to its own row in stream order.  Four
    // candidates are read per LDS round trip, into two sets of registers
    // used in turn: the next group's reads are issued before this group's
    // test and appends, so their latency overlaps them.
    auto group = [&](const float4 (&q)[4], int u) {
      bool ok[4], nx[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) test(q[r], ok[r], nx[r]);
      // two copies of the appends, so that in the common one the FP32
      // decisions stay lane masks; a lane needs the FP64 test where nx and ok
      // differ (ok implies nx), found on the two masks without a per-lane not
      const unsigned long long retest = (__ballot(nx[0]) ^ __ballot(ok[0])) | (__ballot(nx[1]) ^ __ballot(ok[1])) |
                                        (__ballot(nx[2]) ^ __ballot(ok[2])) | (__ballot(nx[3]) ^ __ballot(ok[3]));
      const int cnt0 = cnt;
      if (retest == 0ull) {
#pragma unroll
        for (int r = 0; r < 4; ++r) push(ok[r], __float_as_int(q[r].w));
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) push(nx[r] && !ok[r] ? exact(u + r) : ok[r], __float_as_int(q[r].w));
      }
      flush(cnt0);
    };
    auto read4 = [&](float4 (&q)[4], int u) {
#pragma unroll
      for (int r = 0; r < 4; ++r) q[r] = stage[w][u + r];
    };
    int u = 0;
    float4 qa[4], qb[4];
    if (kc >= 4) read4(qa, 0);
    for (; u + 8 <= kc; u += 8) {
      read4(qb, u + 4);
      group(qa, u);
      if (u + 12 <= kc) read4(qa, u + 8);
      group(qb, u + 4);
    }
    if (u + 4 <= kc) {  // qa holds the group at u
      group(qa, u);
      u += 4;
    }
    for (; u < kc; ++u) {
      const float4 q = stage[w][u];
      bool ok, nx;
      test(q, ok, nx);
      if (nx && !ok) ok = exact(u);
      const int cnt0 = cnt;
      push(ok, __float_as_int(q.w));
      flush(cnt0);
    }
    wave_sync();
  }
  {  // the last (cnt & 3) entries, in the ring from slot first & (kListRing - 1)
    const int k = cnt & 3;
    const int first = cnt - k;
    const int h = first & (kListRing - 1);
    if (k > 0 && first < seg_cap) row[first] = my[h];
    if (k > 1 && first + 1 < seg_cap) row[first + 1] = my[h + 1];
    if (k > 2 && first + 2 < seg_cap) row[first + 2] = my[h + 2];
  }
  if (cnt > seg_cap) atomicOr(s.status, 4);
  if (active) s.nbr_num[wave * s.n + i] = cnt < seg_cap ? cnt : seg_cap;
}

__global__ __launch_bounds__(kNbrWaves * kWave) void k_nbr_build_verlet(SysView s) {
  nbr_build_verlet_body(s, (int)blockIdx.x);
}

// RBL list state: core slots [0, rs_num), kept shell slots from rs_num (every
// pice_num-th shell atom in traversal order, so one lane walks the whole stencil)
struct RblState {
  int core, shell_seen, shell_kept;
};

__device__ __forceinline__ void rbl_consider(const SysView& s, RblState& L, int i, int entry, double d2) {
  if (s.cut2 - d2 > 0.0001 && d2 - s.rs2 > 0.0001) {
    if (L.shell_seen % s.pice_num == 0) {
      int slot = s.rs_num + L.shell_kept;
      if (slot < s.cap)
        s.nbr[nbr_index(slot, i, s.cap)] = entry;
      else
        atomicOr(s.status, 4);
      L.shell_kept++;
    }
    L.shell_seen++;
  } else if (s.rs2 - d2 > 0.0001 && d2 > 0.0001) {
    if (L.core < s.rs_num)
      s.nbr[nbr_index(L.core, i, s.cap)] = entry;
    else
      atomicOr(s.status, 4);
    L.core++;
  }
}

__device__ __forceinline__ void rbl_finish(const SysView& s, const RblState& L, int i) {
  s.nbr_num[2 * i] = L.core < s.rs_num ? L.core : s.rs_num;
  int kept = L.shell_kept;
  if (s.rs_num + kept > s.cap) kept = s.cap - s.rs_num;
  s.nbr_num[2 * i + 1] = kept;
}

// RBL list build: one wavefront per work item; the j loop is wave-uniform
// (scalar loads broadcast to all lanes)
__global__ __launch_bounds__(kBlock) void k_nbr_build_rbl(SysView s) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  const int item_idx = blockIdx.x * (kBlock / kWave) + wave;
  if (item_idx >= s.n_items_dev[0]) return;
  const int item = __builtin_amdgcn_readfirstlane(s.items[item_idx]);
  const int c = item >> 8;
  const int chunk = item & 255;
  const int start = s.cell_start[c];
  const int cnt_c = s.cell_start[c + 1] - start;
  const int local = chunk * kChunk + lane;
  const int i = start + local;
  bool active = local < cnt_c;
  double4 pi = make_double4(0, 0, 0, 0);
  if (active) {
    pi = s.spos[i];
    if (s.misfit[i]) active = false;
  }
  const int ca = c % s.dims[0];
  const int cb = (c / s.dims[0]) % s.dims[1];
  const int cc = c / (s.dims[0] * s.dims[1]);
  RblState L = {0, 0, 0};
  const int nc = s.nc;
  for (int a = -nc; a <= nc; ++a)
    for (int b = -nc; b <= nc; ++b)
      for (int e = -nc; e <= nc; ++e) {
        StencilCell t = stencil_cell(s, ca, cb, cc, a, b, e);
        if (!t.ok) {
          if (lane == 0) atomicOr(s.status, 4);
          continue;
        }
        const int code = pack_entry(0, t.sa, t.sb, t.sc);
        for (int j = t.js; j < t.je; ++j) {
          double4 pj = s.spos[j];  // wave-uniform address -> scalar load
          double rx = (pj.x - t.ox) - pi.x;
          double ry = (pj.y - t.oy) - pi.y;
          double rz = (pj.z - t.oz) - pi.z;
          double d2 = dist2(rx, ry, rz);
          if (active) rbl_consider(s, L, i, (j << kShiftBits) | code, d2);
        }
      }
  if (active) rbl_finish(s, L, i);
}

// RBL misfit atoms, one wavefront each: core entries at their ballot prefix;
// shell candidates get their ordinal in traversal order from the ballot prefix
// and every pice_num-th one is kept, again at its ballot prefix among the kept
__global__ __launch_bounds__(kBlock) void k_nbr_misfit_rbl(SysView s) {
  const int lane = threadIdx.x & (kWave - 1);
  const int gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const int nw = gridDim.x * blockDim.x / kWave;
  const int count = s.misfit_count[0];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int m = gw; m < count; m += nw) {
    const int i = __builtin_amdgcn_readfirstlane(s.misfit_list[m]);
    const double4 pi = s.spos[i];
    const int ca = cell_axis(pi.x, s.lo[0], s.dxdydz[0]);
    const int cb = cell_axis(pi.y, s.lo[1], s.dxdydz[1]);
    const int cc = cell_axis(pi.z, s.lo[2], s.dxdydz[2]);
    int core = 0, seen = 0, kept = 0;
    const int nc = s.nc;
    for (int a = -nc; a <= nc; ++a)
      for (int b = -nc; b <= nc; ++b)
        for (int e = -nc; e <= nc; ++e) {
          StencilCell t = stencil_cell(s, ca, cb, cc, a, b, e);
          if (!t.ok) {
            if (lane == 0) atomicOr(s.status, 1);
            continue;
          }
          const int code = pack_entry(0, t.sa, t.sb, t.sc);
          for (int base = t.js; base < t.je; base += kWave) {
            const int j = base + lane;
            bool is_shell = false, is_core = false;
            if (j < t.je) {
              double4 pj = s.spos[j];
              double rx = (pj.x - t.ox) - pi.x;
              double ry = (pj.y - t.oy) - pi.y;
              double rz = (pj.z - t.oz) - pi.z;
              double d2 = dist2(rx, ry, rz);
              is_shell = s.cut2 - d2 > 0.0001 && d2 - s.rs2 > 0.0001;
              is_core = !is_shell && s.rs2 - d2 > 0.0001 && d2 > 0.0001;
            }
            const int entry = (j << kShiftBits) | code;
            const unsigned long long bc = __ballot(is_core);
            if (is_core) {
              int slot = core + __popcll(bc & below);
              if (slot < s.rs_num)
                s.nbr[nbr_index(slot, i, s.cap)] = entry;
              else
                atomicOr(s.status, 4);
            }
            core += __popcll(bc);
            const unsigned long long bs = __ballot(is_shell);
            const bool keep = is_shell && ((seen + __popcll(bs & below)) % s.pice_num == 0);
            const unsigned long long bk = __ballot(keep);
            if (keep) {
              int slot = s.rs_num + kept + __popcll(bk & below);
              if (slot < s.cap)
                s.nbr[nbr_index(slot, i, s.cap)] = entry;
              else
                atomicOr(s.status, 4);
            }
            seen += __popcll(bs);
            kept += __popcll(bk);
          }
        }
    if (lane == 0) {
      RblState L = {core, seen, kept};
      rbl_finish(s, L, i);
    }
  }
}

void launch_nbr_build(const SysView& s, hipStream_t st) {
  if (s.rbl) {
    int grid = (s.n_items + (kBlock / kWave) - 1) / (kBlock / kWave);
    hipLaunchKernelGGL(k_nbr_build_rbl, dim3(grid), dim3(kBlock), 0, st, s);
    int waves = s.n < 1024 ? s.n : 1024;
    hipLaunchKernelGGL(k_nbr_misfit_rbl, dim3((waves * kWave + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s);
  } else {
    hipLaunchKernelGGL(k_nbr_build_verlet, dim3(kNbrParts * (s.n_items + kMisfitBlocks)), dim3(kNbrWaves * kWave), 0,
                       st, s);
  }
}

// ------------------------------------------------------------------ packs

__global__ __launch_bounds__(1024) void kp_scan(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  scan_body(e.v, bid);
}

__global__ __launch_bounds__(kBlock) void kp_scatter(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  scatter_body(e.v, bid);
}

__global__ __launch_bounds__(kBlock) void kp_cell_sort(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  cell_sort_body(e.v, bid);
}

__global__ __launch_bounds__(kNbrWaves * kWave) void kp_nbr_build(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  nbr_build_verlet_body(e.v, bid);
}

void launch_pack_cells(int kernel, const PackLaunch& L, int nblocks, size_t lds, hipStream_t st) {
  if (nblocks <= 0) return;
  const dim3 grid(nblocks);
  switch (kernel) {
    case PK_SCAN:
      hipLaunchKernelGGL(kp_scan, grid, dim3(1024), 0, st, PACK_ARGS(L));
      break;
    case PK_SCATTER:
      hipLaunchKernelGGL(kp_scatter, grid, dim3(kBlock), lds, st, PACK_ARGS(L));
      break;
    case PK_CELL_SORT:
      hipLaunchKernelGGL(kp_cell_sort, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      break;
    case PK_NBR:
      hipLaunchKernelGGL(kp_nbr_build, grid, dim3(kNbrWaves * kWave), 0, st, PACK_ARGS(L));
      break;
    default:
      break;
  }
}

}  // namespace rbmd
