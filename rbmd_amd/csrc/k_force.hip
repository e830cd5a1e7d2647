// k_force.hip -- fused cell-stencil real-space force: the neighbour search and
// the pair term in one pass, with no stored neighbour list.
//
// Pair set: exactly the VERLET list's (ComputeNeighboursWorklet,
// RunWorklet.cpp:296-355): the 3x3x3 stencil around InWhichCell(p_i) with the
// periodic image offsets of ExecPointLocator.h:298-319, accepting a candidate
// when rc^2 - d^2 > 1e-4 and d^2 > 1e-4 in FP64.  Pair term: the one of
// NearForceVerlet(Weight)Virial (RunWorklet.cpp:2767-2985), shared with the
// list kernels (pair.hpp).  Only the order of the force sums differs from the
// list path.
//
// Workgroup = (work item, stencil x offset a): 3 wavefronts, wavefront w owns
// the stencil column (a, b = w - 1); the work item's 64 i-atoms (one cell
// chunk in sorted order) are held in LDS.  Per round of 64 candidates of the
// column stream (coalesced loads, the next round prefetched):
//   1. candidates farther than rc from the chunk's bounding box are dropped by
//      a wave-uniform ballot; the rest are compacted into LDS (shifted
//      position, charge, sorted index, type);
//   2. lane l tests the staged candidates against i-atom l; every accepted
//      (i, candidate) pair is appended to the wavefront's LDS pair queue at
//      its ballot rank (candidate-major order);
//   3. whenever 64 pairs are queued, each lane evaluates one of them (full
//      lanes whatever the spread of the per-atom counts), adds the force to
//      the wavefront's per-atom LDS accumulator with an LDS FP64 atomic add
//      (one wavefront per accumulator: the adds follow the queue order) and
//      keeps the energy / virial terms in its own registers.
// The three column accumulators of each atom are added in wavefront order and
// stored as the slab-a partial (fn[a + 1][i]); consumers add the three slabs
// in order.  Workgroups past the work items take the misfit atoms (current
// cell != binned cell, e.g. after SHAKE): each a fixed range of sorted
// positions, its misfit atoms one at a time in increasing position, lanes =
// candidates of all 9 columns, and the whole force goes to slab 0.
#include "dev.hpp"
#include "reduce.hpp"
#include "cells.hpp"
#include "pair.hpp"

namespace rbmd {

constexpr int kSlabWaves = 3;
constexpr int kSlabThreads = kSlabWaves * kWave;
constexpr int kRing = 512;  // pair queue per wavefront (>= 63 + 4 * 64 pending pairs)

__device__ __forceinline__ double uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// per-lane energy / virial accumulators
struct ThermoAcc {
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double elj = 0, ec = 0;
};

// special-bond weight of partner j in a list of SP entries (last match wins,
// RunWorklet.cpp:2827-2857); unused entries hold -1
template <int SP>
__device__ __forceinline__ double special_weight(const int* ids, const double* ws, int j) {
  double w = 1.0;
#pragma unroll
  for (int k = 0; k < SP; ++k)
    if (ids[k] == j) w = ws[k];
  return w;
}

template <int SP>
__device__ __forceinline__ void load_specials(const SysView& s, int i, int* ids, double* ws) {
  const int id = s.point_ids[i];
  const int b = s.sp_off[id];
  int spn = s.sp_off[id + 1] - b;
  if (spn > SP) {
    atomicOr(s.status, 4);
    spn = SP;
  }
#pragma unroll
  for (int k = 0; k < SP; ++k) {
    ids[k] = k < spn ? s.rank[s.sp_ids[b + k]] : -1;
    ws[k] = k < spn ? s.sp_w[b + k] : 1.0;
  }
}

// r = p_j (shifted) - p_i, then the minimum image as the list pair kernel
// applies it (a no-op for accepted pairs when rc <= L/2, skipped there)
__device__ __forceinline__ void pair_vector(const SysView& s, double& rx, double& ry, double& rz) {
  if (s.use_min_image) {
    if (fabs(rx) > s.box[0] * 0.5) rx -= (rx > 0 ? s.box[0] : -s.box[0]);
    if (fabs(ry) > s.box[1] * 0.5) ry -= (ry > 0 ? s.box[1] : -s.box[1]);
    if (fabs(rz) > s.box[2] * 0.5) rz -= (rz > 0 ? s.box[2] : -s.box[2]);
  }
}

// ------------------------------------------------------------------ misfits

// Misfit atom i: the block's 3 wavefronts take the columns (a, b = w - 1) for
// a = -1, 0, 1; lanes are candidates.  Totals: lanes (shuffle tree), then the
// wavefronts in order through `buf` (>= kSlabWaves * kWave doubles).
template <int MODE, bool THERMO, bool PERATOM, int SP>
__device__ void misfit_force(const SysView& s, const double2* ctab, const double* teps, const double* tsig, int wave,
                             int lane, double* buf, ThermoAcc& tacc) {
#pragma clang fp contract(off)  // fused multiply-adds are written out (see pair_term)
  constexpr bool EN = THERMO || PERATOM;
  constexpr int NV = PERATOM ? 11 : 3;
  if (s.misfit_count[0] == 0) return;
  const int b = wave - 1;
  const int nt = s.ntypes;
  // this workgroup's range of sorted positions; its misfit atoms are taken in
  // increasing position, so the thermo sums it accumulates have a fixed order
  // (the misfit list itself is appended in atomic order)
  const int per = (s.n + kMisfitBlocks - 1) / kMisfitBlocks;
  const int p0 = ((int)blockIdx.x - kSlabWaves * s.n_items) * per;
  const int p1 = p0 + per < s.n ? p0 + per : s.n;
  for (int base = p0; base < p1; base += kWave) {
   unsigned long long flagged = __ballot(base + lane < p1 && s.misfit[base + lane] != 0);
   while (flagged) {
    const int i = base + __builtin_ctzll(flagged);
    flagged &= flagged - 1ull;
    const double4 pi = s.spos[i];
    const int ti = s.stype[i];
    int spj[SP > 0 ? SP : 1];
    double spw[SP > 0 ? SP : 1];
    if (MODE == 1) load_specials<SP>(s, i, spj, spw);
    const int ca = cell_axis(pi.x, s.lo[0], s.dxdydz[0]);
    const int cb = cell_axis(pi.y, s.lo[1], s.dxdydz[1]);
    const int cc = cell_axis(pi.z, s.lo[2], s.dxdydz[2]);
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int a = -1; a <= 1; ++a) {
      Column col;
      make_column(s, ca, cb, cc, a, b, lane, col);
      for (int base = 0; base < col.total; base += kWave) {
        const int t = base + lane;
        if (t >= col.total) continue;
        double4 pj;
        int entry;
        COLUMN_CANDIDATE(t, pj, entry);
        double rx = pj.x - pi.x, ry = pj.y - pi.y, rz = pj.z - pi.z;
        const double dd = dist2(rx, ry, rz);
        if (!(s.cut2 - dd > 0.0001 && dd > 0.0001)) continue;
        const int j = entry >> kShiftBits;
        pair_vector(s, rx, ry, rz);
        const double d2 = dist2(rx, ry, rz);
        const int tix = ti * nt + s.stype[j];
        const double w = MODE == 1 ? special_weight<SP>(spj, spw, j) : 1.0;
        double elj, ec;
        const double sc = pair_term<MODE, EN>(s, ctab, d2, pi.w * pj.w, teps[tix], tsig[tix], w, elj, ec);
        // one fused rounding per force term
        v[0] = fma(sc, rx, v[0]);
        v[1] = fma(sc, ry, v[1]);
        v[2] = fma(sc, rz, v[2]);
        if (EN) {
          const double px = sc * rx, py = sc * ry, pz = sc * rz;
          double vv[6] = {0, 0, 0, 0, 0, 0};
          vir_add(vv, rx, ry, rz, px, py, pz, MODE == 1 ? -0.5 : 0.5);
          for (int k = 0; k < 6; ++k) tacc.vir[k] += vv[k];
          tacc.elj += elj;
          tacc.ec += ec;
          if (PERATOM) {
            for (int k = 0; k < 6; ++k) v[3 + k] += vv[k];
            v[9] += elj;
            v[10] += ec;
          }
        }
      }
    }
    // the atom's totals: lanes (shuffle tree), then wavefronts in order
    for (int k = 0; k < NV; ++k) {
      const double t = wave_sum(v[k]);
      __syncthreads();
      if (lane == 0) buf[wave] = t;
      __syncthreads();
      if (wave == 0 && lane == 0) {
        double acc = buf[0];
        for (int w2 = 1; w2 < kSlabWaves; ++w2) acc += buf[w2];
        v[k] = acc;
      }
    }
    if (wave == 0 && lane == 0) {
      const size_t n = s.n;
      s.fnx[i] = v[0];
      s.fny[i] = v[1];
      s.fnz[i] = v[2];
      for (int sl = 1; sl < kSlabWaves; ++sl) s.fnx[sl * n + i] = s.fny[sl * n + i] = s.fnz[sl * n + i] = 0.0;
      if (PERATOM)
        for (int sl = 0; sl < kSlabWaves; ++sl) {
          for (int k = 0; k < 6; ++k) s.vir_near[(sl * n + i) * 6 + k] = sl == 0 ? v[3 + k] : 0.0;
          s.e_lj[sl * n + i] = sl == 0 ? v[9] : 0.0;
          s.e_coul[sl * n + i] = sl == 0 ? v[10] : 0.0;
        }
    }
   }
  }
}

// ------------------------------------------------------------------ main path

template <int MODE, bool THERMO, bool PERATOM, int SP>
__global__ __launch_bounds__(kSlabThreads) void k_force_cells(SysView s, double* rows) {
#pragma clang fp contract(off)  // fused multiply-adds are written out (see pair_term)
  constexpr bool EN = THERMO || PERATOM;
  constexpr int SPD = SP > 0 ? SP : 1;
  constexpr int NACC = PERATOM ? 11 : 3;  // per-atom accumulators: force (+ virial, energies)
  // the work item's i-atoms (shared by the three wavefronts)
  __shared__ double4 ipos[kWave];
  __shared__ int ityp[kWave];
  __shared__ int isp_j[kWave][SPD];
  __shared__ double isp_w[kWave][SPD];
  // per wavefront: staged candidates, pair queue, per-atom accumulators
  __shared__ double4 stage[kSlabWaves][kWave];
  __shared__ int stage_j[kSlabWaves][kWave];
  __shared__ int stage_t[kSlabWaves][kWave];
  __shared__ unsigned short ring[kSlabWaves][kRing];
  __shared__ double facc[kSlabWaves][NACC][kWave];
  // tables
  __shared__ double teps[kMaxTypes * kMaxTypes], tsig[kMaxTypes * kMaxTypes];
  __shared__ double2 ctab[MODE == 2 ? 1 : kCoulTab];
  const int nt = s.ntypes;
  for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) {
    teps[k] = s.eps_ij[k];
    tsig[k] = s.sig6_ij[k];
  }
  if (MODE != 2)
    for (int k = threadIdx.x; k < kCoulTab; k += blockDim.x) ctab[k] = s.ctab[k];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  ThermoAcc tacc;
  if ((int)blockIdx.x >= kSlabWaves * s.n_items) {
    __syncthreads();
    misfit_force<MODE, THERMO, PERATOM, SP>(s, ctab, teps, tsig, wave, lane, &facc[0][0][0], tacc);
  } else if ((int)blockIdx.x / kSlabWaves < s.n_items_dev[0]) {
    const int item = __builtin_amdgcn_readfirstlane(s.items[blockIdx.x / kSlabWaves]);
    const int slab = blockIdx.x % kSlabWaves;
    const int a = slab - 1, b = wave - 1;
    const int c = item >> 8;
    const int start = s.cell_start[c];
    const int local = (item & 255) * kChunk + lane;
    const int i = start + local;
    bool active = local < s.cell_start[c + 1] - start;
    double4 pi = make_double4(0, 0, 0, 0);
    if (active) {
      pi = s.spos[i];
      if (s.misfit[i]) active = false;  // a misfit workgroup takes it
    }
    if (wave == 0) {
      ipos[lane] = pi;
      ityp[lane] = active ? s.stype[i] : 0;
      if (MODE == 1) {
        int ids[SPD];
        double ws[SPD];
        if (active) {
          load_specials<SP>(s, i, ids, ws);
        } else {
#pragma unroll
          for (int k = 0; k < SPD; ++k) {
            ids[k] = -1;
            ws[k] = 1.0;
          }
        }
#pragma unroll
        for (int k = 0; k < SPD; ++k) {
          isp_j[lane][k] = ids[k];
          isp_w[lane][k] = ws[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NACC; ++k) facc[wave][k][lane] = 0.0;
    // bounding box of the active i-atoms (wave-uniform)
    double bl[3], bh[3];
    {
      const double v[3] = {pi.x, pi.y, pi.z};
      for (int d = 0; d < 3; ++d) {
        double lo = active ? v[d] : 1e300, hi = active ? v[d] : -1e300;
        for (int off = 1; off < kWave; off <<= 1) {
          lo = fmin(lo, __shfl_xor(lo, off, kWave));
          hi = fmax(hi, __shfl_xor(hi, off, kWave));
        }
        bl[d] = uniform(lo);
        bh[d] = uniform(hi);
      }
    }
    __syncthreads();
    const int ca = c % s.dims[0];
    const int cb = (c / s.dims[0]) % s.dims[1];
    const int cc = c / (s.dims[0] * s.dims[1]);
    Column col;
    make_column(s, ca, cb, cc, a, b, lane, col);
    const int total = col.total;
    const double cut2 = s.cut2;
    unsigned short* rq = ring[wave];
    int head = 0, tail = 0;  // queue positions (wave-uniform)

    // one batch: lane l evaluates queued pair head + l (l < cnt)
    auto batch = [&](int cnt) {
      if (lane < cnt) {
        const int e = rq[(head + lane) & (kRing - 1)];
        const int ii = e & (kWave - 1), u = e >> 6;
        const double4 p = ipos[ii];
        const double4 q = stage[wave][u];
        const int j = stage_j[wave][u];
        const int tix = ityp[ii] * nt + stage_t[wave][u];
        double rx = q.x - p.x, ry = q.y - p.y, rz = q.z - p.z;
        pair_vector(s, rx, ry, rz);
        const double d2 = dist2(rx, ry, rz);
        const double w = MODE == 1 ? special_weight<SP>(isp_j[ii], isp_w[ii], j) : 1.0;
        double elj, ec;
        const double sc = pair_term<MODE, EN>(s, ctab, d2, p.w * q.w, teps[tix], tsig[tix], w, elj, ec);
        const double px = sc * rx, py = sc * ry, pz = sc * rz;
        atomicAdd(&facc[wave][0][ii], px);
        atomicAdd(&facc[wave][1][ii], py);
        atomicAdd(&facc[wave][2][ii], pz);
        if (EN) {
          double vv[6] = {0, 0, 0, 0, 0, 0};
          vir_add(vv, rx, ry, rz, px, py, pz, MODE == 1 ? -0.5 : 0.5);
#pragma unroll
          for (int k = 0; k < 6; ++k) tacc.vir[k] += vv[k];
          tacc.elj += elj;
          tacc.ec += ec;
          if (PERATOM) {
#pragma unroll
            for (int k = 0; k < 6; ++k) atomicAdd(&facc[wave][3 + k][ii], vv[k]);
            atomicAdd(&facc[wave][9][ii], elj);
            atomicAdd(&facc[wave][10][ii], ec);
          }
        }
      }
      head += cnt < kWave ? cnt : kWave;
    };

    double4 nxt = make_double4(0, 0, 0, 0);
    int nxt_e = 0, nxt_t = 0;
    if (lane < total) {
      COLUMN_CANDIDATE(lane, nxt, nxt_e);
      nxt_t = s.stype[nxt_e >> kShiftBits];
    }
    for (int base = 0; base < total; base += kWave) {
      const int t = base + lane;
      const double4 pj = nxt;
      const int entry = nxt_e, tj = nxt_t;
      if (t + kWave < total) {
        COLUMN_CANDIDATE(t + kWave, nxt, nxt_e);
        nxt_t = s.stype[nxt_e >> kShiftBits];
      }
      bool keep = false;
      if (t < total) {
        const double dx = fmax(0.0, fmax(bl[0] - pj.x, pj.x - bh[0]));
        const double dy = fmax(0.0, fmax(bl[1] - pj.y, pj.y - bh[1]));
        const double dz = fmax(0.0, fmax(bl[2] - pj.z, pj.z - bh[2]));
        keep = dx * dx + dy * dy + dz * dz < cut2;  // superset of every lane's acceptance
      }
      const unsigned long long bal = __ballot(keep);
      const int kc = __popcll(bal);
      if (kc == 0) continue;
      wave_sync();  // the previous round's batches have read the stage
      if (keep) {
        const int slot = __popcll(bal & ((1ull << lane) - 1ull));
        stage[wave][slot] = pj;
        stage_j[wave][slot] = entry >> kShiftBits;
        stage_t[wave][slot] = tj;
      }
      wave_sync();
      for (int u0 = 0; u0 < kc; u0 += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int u = u0 + k;
          bool ok = false;
          if (u < kc) {
            const double4 q = stage[wave][u];
            const double rx = q.x - pi.x, ry = q.y - pi.y, rz = q.z - pi.z;
            const double d2 = dist2(rx, ry, rz);
            ok = active & (cut2 - d2 > 0.0001) & (d2 > 0.0001);
          }
          const unsigned long long acc = __ballot(ok);
          if (ok) rq[(tail + __popcll(acc & ((1ull << lane) - 1ull))) & (kRing - 1)] = (unsigned short)(lane | (u << 6));
          tail += __popcll(acc);
        }
        wave_sync();
        while (tail - head >= kWave) batch(kWave);
        wave_sync();
      }
      if (tail > head) {
        batch(tail - head);
        wave_sync();
      }
    }
    // slab partial of each i-atom: the three column accumulators added in order
    __syncthreads();
    if (wave == 0 && active) {
      double v[NACC];
#pragma unroll
      for (int k = 0; k < NACC; ++k) {
        double t = facc[0][k][lane];
        for (int w2 = 1; w2 < kSlabWaves; ++w2) t += facc[w2][k][lane];
        v[k] = t;
      }
      const size_t o = (size_t)slab * s.n + i;
      s.fnx[o] = v[0];
      s.fny[o] = v[1];
      s.fnz[o] = v[2];
      if (PERATOM) {
        for (int k = 0; k < 6; ++k) s.vir_near[o * 6 + k] = v[3 + k];
        s.e_lj[o] = v[9];
        s.e_coul[o] = v[10];
      }
    }
  } else {
    __syncthreads();
  }
  if (THERMO) {
    const double vals[kRowPair] = {tacc.elj,    tacc.ec,     tacc.vir[0], tacc.vir[1],
                                   tacc.vir[2], tacc.vir[3], tacc.vir[4], tacc.vir[5]};
    block_write_row<kRowPair, kSlabWaves>(vals, rows + (size_t)blockIdx.x * kRowPair);
  }
}

int force_cells_blocks(const SysView& s) { return kSlabWaves * s.n_items + kMisfitBlocks; }

template <int MODE, int SP>
void launch_fc(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int grid) {
  if (peratom)
    hipLaunchKernelGGL((k_force_cells<MODE, true, true, SP>), dim3(grid), dim3(kSlabThreads), 0, st, s, rows);
  else if (thermo)
    hipLaunchKernelGGL((k_force_cells<MODE, true, false, SP>), dim3(grid), dim3(kSlabThreads), 0, st, s, rows);
  else
    hipLaunchKernelGGL((k_force_cells<MODE, false, false, SP>), dim3(grid), dim3(kSlabThreads), 0, st, s, rows);
}

// requires ntypes <= kMaxTypes (type-pair tables in LDS) and sp_max <= 16
void launch_force_cells(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int* nblocks) {
  const int grid = force_cells_blocks(s);
  if (nblocks) *nblocks = grid;
  if (s.lj_only)
    launch_fc<2, 0>(s, st, thermo, peratom, rows, grid);
  else if (!s.weighted)
    launch_fc<0, 0>(s, st, thermo, peratom, rows, grid);
  else if (s.sp_max <= 4)
    launch_fc<1, 4>(s, st, thermo, peratom, rows, grid);
  else
    launch_fc<1, 16>(s, st, thermo, peratom, rows, grid);
}

}  // namespace rbmd
