// cells.hpp -- device helpers shared by the neighbour-list build and the
// fused cell-stencil force kernel: cell coordinates, the 3x3x3 stencil with
// periodic images, and the candidate stream of one stencil column.
#pragma once
#include "dev.hpp"

namespace rbmd {

constexpr int kMisfitBlocks = 64;  // workgroups that take the misfit atoms
constexpr int kColWaves = 9;        // stencil columns of the VERLET list build
constexpr int kScanScatterCells = 2048;  // most cells for the one-launch scan + scatter
constexpr int kNbrWaves = 3;        // columns (wavefronts) per list-build workgroup
constexpr int kNbrParts = kColWaves / kNbrWaves;  // list-build workgroups per work item
constexpr int kStageRounds = 2;    // 64-candidate rounds staged per list-build batch
constexpr int kStageCap = kStageRounds * 64;  // staged candidates per wavefront (one batch)
constexpr int kListRing = 8;        // list-build LDS ring of accepted entries per lane (two groups of four)
constexpr int kListRingStride = 9;  // words between two lanes' rings (odd: distinct banks)

// InWhichCell (ExecPointLocator.h:271-274): unclamped trunc
__device__ __forceinline__ int cell_axis(double x, double lo, double d) {
  double t = (x - lo) / d;
  if (!(t == t)) return 0;
  return (int)fmax(fmin(t, 2.0e9), -2.0e9);
}

__device__ __forceinline__ int wrap_axis(int v, int dims) { return (v + dims) % dims; }

__device__ __forceinline__ int pack_entry(int j, int sa, int sb, int sc) {
  return (j << kShiftBits) | ((sa + 3) | ((sb + 3) << 3) | ((sc + 3) << 6));
}

// a packed list entry's sorted index and image offset (the offset the build
// applied, bit for bit: the same integer multiple of the cell width)
__device__ __forceinline__ void decode(const SysView& s, int e, int& j, double& ox, double& oy, double& oz) {
  j = e >> kShiftBits;
  int code = e & ((1 << kShiftBits) - 1);
  ox = s.dxdydz[0] * (double)(((code & 7) - 3) * s.dims[0]);
  oy = s.dxdydz[1] * (double)((((code >> 3) & 7) - 3) * s.dims[1]);
  oz = s.dxdydz[2] * (double)((((code >> 6) & 7) - 3) * s.dims[2]);
}

// wavefront-local LDS ordering (writes of all lanes visible before reads)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One stencil cell seen from cell (ca,cb,cc) at offset (a,b,e): wrapped cell,
// image offset dxdydz * (wrapped - unwrapped) (PeriodicIndexOffset /
// PeriodicCoordOffset, ExecPointLocator.h:298-319) and the packed shift code.
struct StencilCell {
  int js, je, sa, sb, sc;
  double ox, oy, oz;
  bool ok;
};

__device__ __forceinline__ StencilCell stencil_cell(const SysView& s, int ca, int cb, int cc, int a, int b, int e) {
  StencilCell t;
  int na = ca + a, nb = cb + b, ne = cc + e;
  int wa = wrap_axis(na, s.dims[0]), wb = wrap_axis(nb, s.dims[1]), we = wrap_axis(ne, s.dims[2]);
  t.ok = wa >= 0 && wb >= 0 && we >= 0;
  if (!t.ok) {
    t.js = t.je = 0;
    return t;
  }
  t.ox = s.dxdydz[0] * (double)(wa - na);
  t.oy = s.dxdydz[1] * (double)(wb - nb);
  t.oz = s.dxdydz[2] * (double)(we - ne);
  t.sa = (wa - na) / s.dims[0];
  t.sb = (wb - nb) / s.dims[1];
  t.sc = (we - ne) / s.dims[2];
  if (t.sa < -3 || t.sa > 3 || t.sb < -3 || t.sb > 3 || t.sc < -3 || t.sc > 3) {
    t.ok = false;
    t.js = t.je = 0;
    return t;
  }
  int wcell = wa + wb * s.dims[0] + we * s.dims[0] * s.dims[1];
  t.js = s.cell_start[wcell];
  t.je = s.cell_start[wcell + 1];
  return t;
}

// The three cells (e = -1, 0, 1) of one stencil column as a candidate stream:
// candidate t < n0 is in cell 0, t < n01 in cell 1, else in cell 2; jb is the
// sorted index of the stream position 0 of each cell (start - prefix).  The
// cells of a column share their x and y image offsets.
struct Column {
  int jb0, jb1, jb2, n0, n01, total, c0, c1, c2;
  double ox, oy, oz0, oz1, oz2;
};

__device__ __forceinline__ StencilCell column_cell(const SysView& s, int ca, int cb, int cc, int a, int b, int e,
                                                   int lane) {
  StencilCell t = stencil_cell(s, ca, cb, cc, a, b, e);
  if (!t.ok) {
    if (lane == 0) atomicOr(s.status, 4);
    t.js = t.je = 0;
    t.sa = t.sb = t.sc = 0;
    t.ox = t.oy = t.oz = 0;
  }
  return t;
}

__device__ __forceinline__ void make_column(const SysView& s, int ca, int cb, int cc, int a, int b, int lane,
                                            Column& col) {
  const StencilCell t0 = column_cell(s, ca, cb, cc, a, b, -1, lane);
  const StencilCell t1 = column_cell(s, ca, cb, cc, a, b, 0, lane);
  const StencilCell t2 = column_cell(s, ca, cb, cc, a, b, 1, lane);
  col.n0 = t0.je - t0.js;
  col.n01 = col.n0 + (t1.je - t1.js);
  col.total = col.n01 + (t2.je - t2.js);
  col.jb0 = t0.js;
  col.jb1 = t1.js - col.n0;
  col.jb2 = t2.js - col.n01;
  col.c0 = pack_entry(0, t0.sa, t0.sb, t0.sc);
  col.c1 = pack_entry(0, t1.sa, t1.sb, t1.sc);
  col.c2 = pack_entry(0, t2.sa, t2.sb, t2.sc);
  col.ox = t1.ox;
  col.oy = t1.oy;
  col.oz0 = t0.oz;
  col.oz1 = t1.oz;
  col.oz2 = t2.oz;
}

// candidate t of a column stream: shifted position (w: charge) and packed entry.  The
// column fields come in by value: selects over fields read through a
// reference become one load through a selected address, which puts the
// column in private scratch memory.
__device__ __forceinline__ void column_candidate(const SysView& s, int t, int n0, int n01, int jb0, int jb1, int jb2,
                                                 int c0, int c1, int c2, double ox, double oy, double oz0, double oz1,
                                                 double oz2, double4& pj, int& entry) {
  const bool q1 = t >= n0, q2 = t >= n01;
  const int j = t + (q2 ? jb2 : (q1 ? jb1 : jb0));
  const double oz = q2 ? oz2 : (q1 ? oz1 : oz0);
  const int code = q2 ? c2 : (q1 ? c1 : c0);
  double4 p = s.spos[j];
  pj = make_double4(p.x - ox, p.y - oy, p.z - oz, p.w);
  entry = (j << kShiftBits) | code;
}

// where candidate t of a column stream lives, without loading it: its sorted
// index, packed entry and z image offset (x and y are the column's), so a
// caller can issue several loads back to back before using any of them
__device__ __forceinline__ int column_locate(int t, int n0, int n01, int jb0, int jb1, int jb2, int c0, int c1,
                                             int c2, double oz0, double oz1, double oz2, int& entry, double& oz) {
  const bool q1 = t >= n0, q2 = t >= n01;
  const int j = t + (q2 ? jb2 : (q1 ? jb1 : jb0));
  oz = q2 ? oz2 : (q1 ? oz1 : oz0);
  entry = (j << kShiftBits) | (q2 ? c2 : (q1 ? c1 : c0));
  return j;
}

#define COLUMN_LOCATE(t, entry, oz) \
  column_locate(t, col.n0, col.n01, col.jb0, col.jb1, col.jb2, col.c0, col.c1, col.c2, col.oz0, col.oz1, col.oz2, entry, oz)

#define COLUMN_CANDIDATE(t, pj, entry)                                                                             \
  column_candidate(s, t, col.n0, col.n01, col.jb0, col.jb1, col.jb2, col.c0, col.c1, col.c2, col.ox, col.oy, \
                   col.oz0, col.oz1, col.oz2, pj, entry)

}  // namespace rbmd
