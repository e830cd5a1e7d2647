// k_pair.hip -- real-space LJ + erfc pair forces over the neighbour lists.
//
// One thread per atom in cell-sorted order; each thread walks its own list
// (column-major, coalesced across the wavefront) and gathers neighbour
// positions from the sorted double4 array, which stays resident in L2.
// Forces are gathered, never scattered: no atomics, deterministic sums.
//
// Formulas (all FP64 here):
//   VERLET plain    ComputeNearForceVerletVirialWorklet, RunWorklet.cpp:2890-2985
//   VERLET weighted ComputeNearForceVerletWeightVirialWorklet1, RunWorklet.cpp:2767-2888
//   LJ/CUT only     ComputeLJForceVerletWorklet, RunWorklet.cpp:498-575
//   RBL             ComputeNearForceRBLERF(SpecialBonds)Worklet, RunWorklet.cpp:982-1266
//   LJ      ExecForceFunction.h:412-435 (force), :818-841 (energy)
//   G(r)    ExecForceFunction.h:682-689, erfc(sqrt(alpha) r)/r^2 + 2 sqrt(alpha/pi) e^{-alpha r^2}/r
//   virial  ComputePairVirial_fix (+1/2, plain) / ComputePairVirial (-1/2, weighted, RBL), :304-332
#include "dev.hpp"
#include "reduce.hpp"
#include "pair.hpp"

namespace rbmd {

constexpr int kPairBatch = 2;

// MODE 0: plain LJ+coul, 1: weighted (special bonds), 2: LJ only.
// L = 2^LG consecutive threads share one atom: lane `sub` takes list slots
// sub, sub + L, ... of every segment, kPairBatch slots at a time so that
// several neighbour gathers are in flight per lane, and the partial sums are
// combined with a fixed xor-shuffle tree at the end (deterministic).  SP is
// the register budget for the atom's special list (>= the longest list).
template <int MODE, bool THERMO, bool PERATOM, int LG, int SP>
__global__ __launch_bounds__(kBlock) void k_pair_verlet(SysView s, double* rows) {
  constexpr int L = 1 << LG;
  __shared__ double tab_eps[kMaxTypes * kMaxTypes], tab_sig[kMaxTypes * kMaxTypes];
  __shared__ double ctab[MODE == 2 ? 1 : kCoulM * 2 * (kCoulD + 1)];
  const int nt = s.ntypes;
  const bool lds_tab = nt <= kMaxTypes;
  if (lds_tab)
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) {
      tab_eps[k] = s.eps_ij[k];
      tab_sig[k] = s.sig6_ij[k];
    }
  if (MODE != 2)
    for (int k = threadIdx.x; k < kCoulM * 2 * (kCoulD + 1); k += blockDim.x) ctab[k] = s.ctab[k];
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t >> LG;
  const int sub = t & (L - 1);
  constexpr bool EN = THERMO || PERATOM;
  double fx = 0, fy = 0, fz = 0;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double pelj = 0, pec = 0;
  if (i < s.n) {
    const double4 pi = s.spos[i];
    const int ti = s.stype[i];
    const double qi = pi.w;
    int spj[SP > 0 ? SP : 1];
    double spw[SP > 0 ? SP : 1];
    if (MODE == 1) {
      int id = s.point_ids[i];
      int b = s.sp_off[id], e = s.sp_off[id + 1];
      int spn = e - b;
      if (spn > SP) {
        atomicOr(s.status, 4);
        spn = SP;
      }
#pragma unroll
      for (int k = 0; k < SP; ++k) {
        spj[k] = k < spn ? s.rank[s.sp_ids[b + k]] : -1;
        spw[k] = k < spn ? s.sp_w[b + k] : 1.0;
      }
    }
    for (int seg = 0; seg < s.nseg; ++seg) {
      const int num = s.nbr_num[seg * s.n + i];
      const int base = seg * s.seg_cap;
      for (int k0 = sub; k0 < num; k0 += L * kPairBatch) {
        int ent[kPairBatch];
#pragma unroll
        for (int bb = 0; bb < kPairBatch; ++bb) {
          int k = k0 + bb * L;
          ent[bb] = k < num ? s.nbr[nbr_index(base + k, i, s.n, LG)] : -1;
        }
        double4 pj[kPairBatch];
        int tj[kPairBatch];
#pragma unroll
        for (int bb = 0; bb < kPairBatch; ++bb) {
          int j = ent[bb] >= 0 ? (ent[bb] >> kShiftBits) : i;
          pj[bb] = s.spos[j];
          tj[bb] = s.stype[j];
        }
#pragma unroll
        for (int bb = 0; bb < kPairBatch; ++bb) {
          if (ent[bb] < 0) continue;
          int j;
          double ox, oy, oz;
          decode(s, ent[bb], j, ox, oy, oz);
          double rx = (pj[bb].x - ox) - pi.x;
          double ry = (pj[bb].y - oy) - pi.y;
          double rz = (pj[bb].z - oz) - pi.z;
          min_image(s, rx, ry, rz);
          const double d2 = rx * rx + ry * ry + rz * rz;
          const int tix = ti * nt + tj[bb];
          const double epsij = lds_tab ? tab_eps[tix] : s.eps_ij[tix];
          const double sig6 = lds_tab ? tab_sig[tix] : s.sig6_ij[tix];
          double w = 1.0;
          if (MODE == 1) {
#pragma unroll
            for (int u = 0; u < SP; ++u)
              if (spj[u] == j) w = spw[u];
          }
          double elj, ec;
          const double sc = pair_term<MODE, EN>(s, ctab, d2, qi * pj[bb].w, epsij, sig6, w, elj, ec);
          const double px = sc * rx, py = sc * ry, pz = sc * rz;
          fx += px;
          fy += py;
          fz += pz;
          if (EN) {
            vir_add(vir, rx, ry, rz, px, py, pz, MODE == 1 ? -0.5 : 0.5);
            pelj += elj;
            pec += ec;
          }
        }
      }
    }
  }
  if (THERMO) {
    const double vals[kRowPair] = {pelj, pec, vir[0], vir[1], vir[2], vir[3], vir[4], vir[5]};
    block_write_row<kRowPair>(vals, rows + (size_t)blockIdx.x * kRowPair);
  }
  fx = group_sum<L>(fx);
  fy = group_sum<L>(fy);
  fz = group_sum<L>(fz);
  if (PERATOM) {
    for (int k = 0; k < 6; ++k) vir[k] = group_sum<L>(vir[k]);
    pelj = group_sum<L>(pelj);
    pec = group_sum<L>(pec);
  }
  if (i < s.n && sub == 0) {
    s.fnx[i] = fx;
    s.fny[i] = fy;
    s.fnz[i] = fz;
    if (PERATOM) {
      for (int k = 0; k < 6; ++k) s.vir_near[(size_t)6 * i + k] = vir[k];
      s.e_lj[i] = pelj;
      s.e_coul[i] = pec;
    }
  }
}

template <int MODE, int LG, int SP>
void launch_pair_v(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int grid) {
  if (peratom)
    hipLaunchKernelGGL((k_pair_verlet<MODE, true, true, LG, SP>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  else if (thermo)
    hipLaunchKernelGGL((k_pair_verlet<MODE, true, false, LG, SP>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  else
    hipLaunchKernelGGL((k_pair_verlet<MODE, false, false, LG, SP>), dim3(grid), dim3(kBlock), 0, st, s, rows);
}

template <int MODE, int SP>
void launch_pair_lg(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int grid) {
  if (s.lanes_lg == 2)
    launch_pair_v<MODE, 2, SP>(s, st, thermo, peratom, rows, grid);
  else if (s.lanes_lg == 3)
    launch_pair_v<MODE, 3, SP>(s, st, thermo, peratom, rows, grid);
  else
    launch_pair_v<MODE, 4, SP>(s, st, thermo, peratom, rows, grid);
}

// RBL: per-atom corrected force before the mean removal; the virial keeps the
// reference's assignment of the running-sum force of the last pair (hazard Q5).
template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_pair_rbl(SysView s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  const double4 pi = s.spos[i];
  const int ti = s.stype[i];
  const double qi = pi.w;
  int spn = 0;
  int spj[kMaxSpecial];
  double spw[kMaxSpecial];
  if (WEIGHTED) {
    int id = s.point_ids[i];
    int b = s.sp_off[id], e = s.sp_off[id + 1];
    spn = e - b;
    if (spn > kMaxSpecial) {
      atomicOr(s.status, 4);
      spn = kMaxSpecial;
    }
#pragma unroll
    for (int k = 0; k < kMaxSpecial; ++k) {
      spj[k] = k < spn ? s.rank[s.sp_ids[b + k]] : -1;
      spw[k] = k < spn ? s.sp_w[b + k] : 1.0;
    }
  }
  const double* eps_row = s.eps_ij + ti * s.ntypes;
  const double* sig_row = s.sig6_ij + ti * s.ntypes;
  double lj_rs[3] = {0, 0, 0}, lj_rcs[3] = {0, 0, 0}, el_rs[3] = {0, 0, 0}, el_rcs[3] = {0, 0, 0};
  double vir_rs[6] = {0, 0, 0, 0, 0, 0}, vir_rcs[6] = {0, 0, 0, 0, 0, 0};
  const double pice = (double)s.pice_num;
  for (int part = 0; part < 2; ++part) {
    const int num = s.nbr_num[2 * i + part];
    const int base = part == 0 ? 0 : s.rs_num;
    for (int k = 0; k < num; ++k) {
      int j;
      double ox, oy, oz;
      decode(s, s.nbr[nbr_index(base + k, i, s.n, s.lanes_lg)], j, ox, oy, oz);
      const double4 pj = s.spos[j];
      double rx = (pj.x - ox) - pi.x;
      double ry = (pj.y - oy) - pi.y;
      double rz = (pj.z - oz) - pi.z;
      min_image(s, rx, ry, rz);
      const double d2 = rx * rx + ry * ry + rz * rz;
      const double dis = sqrt(d2);
      const int tj = s.stype[j];
      const double epsij = eps_row[tj];
      const double sig6 = sig_row[tj];
      const double qq = qi * pj.w;
      const double g = s.erfc_table ? gnear_table(dis) : gnear(s.alpha, s.sqrt_alpha, dis);
      double w = 1.0;
      if (WEIGHTED) {
#pragma unroll
        for (int t = 0; t < kMaxSpecial; ++t)
          if (spj[t] == j) w = spw[t];
      }
      double d6 = d2 * d2 * d2;
      double sd6 = sig6 / d6;
      double ljs = -24.0 * epsij * ((2.0 * sd6 - 1.0) * sd6) / d2;
      if (part == 0) {
        if (d2 < s.cut2 && d2 > 0.0001) {
          double e = -qq * g / dis;
          el_rs[0] += e * rx;
          el_rs[1] += e * ry;
          el_rs[2] += e * rz;
        }
        if (d2 < s.rs2 && d2 > 0.0001) {  // LJ core with cutoff rs (hazard Q4)
          lj_rs[0] += w * (ljs * rx);
          lj_rs[1] += w * (ljs * ry);
          lj_rs[2] += w * (ljs * rz);
        }
        for (int t = 0; t < 6; ++t) vir_rs[t] = 0;
        vir_add(vir_rs, rx, ry, rz, lj_rs[0] + s.qqr2e * el_rs[0], lj_rs[1] + s.qqr2e * el_rs[1],
                lj_rs[2] + s.qqr2e * el_rs[2], -0.5);
      } else {
        bool inshell = WEIGHTED ? (d2 < s.cut2 && d2 > s.rs2) : (d2 < s.cut2 && d2 > 0.0001);
        if (inshell) {
          double e = -qq * g / dis;
          el_rcs[0] += pice * (e * rx);
          el_rcs[1] += pice * (e * ry);
          el_rcs[2] += pice * (e * rz);
        }
        if (d2 < s.cut2 && d2 > s.rs2) {
          lj_rcs[0] += w * (pice * (ljs * rx));
          lj_rcs[1] += w * (pice * (ljs * ry));
          lj_rcs[2] += w * (pice * (ljs * rz));
        }
        for (int t = 0; t < 6; ++t) vir_rcs[t] = 0;
        vir_add(vir_rcs, rx, ry, rz, lj_rcs[0] + s.qqr2e * el_rcs[0], lj_rcs[1] + s.qqr2e * el_rcs[1],
                lj_rcs[2] + s.qqr2e * el_rcs[2], -0.5);
      }
    }
  }
  s.fnx[i] = (lj_rs[0] + lj_rcs[0]) + s.qqr2e * (el_rs[0] + el_rcs[0]);
  s.fny[i] = (lj_rs[1] + lj_rcs[1]) + s.qqr2e * (el_rs[1] + el_rcs[1]);
  s.fnz[i] = (lj_rs[2] + lj_rcs[2]) + s.qqr2e * (el_rs[2] + el_rcs[2]);
  if (s.vir_near)
    for (int k = 0; k < 6; ++k) s.vir_near[(size_t)6 * i + k] = vir_rs[k] + vir_rcs[k];
}

// mean removal of the RBL correction force and virial (SumRBLCorrForce /
// SumRBLCorrVirial, ExecutionMD.cpp:663-667): one workgroup sums the per-atom
// values deterministically, then subtracts the means
__global__ __launch_bounds__(1024) void k_rbl_mean(SysView s) {
  __shared__ double lds[16][9];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const bool vir = s.vir_near != nullptr;
  for (int i = threadIdx.x; i < s.n; i += blockDim.x) {
    acc[0] += s.fnx[i];
    acc[1] += s.fny[i];
    acc[2] += s.fnz[i];
    if (vir)
      for (int k = 0; k < 6; ++k) acc[3 + k] += s.vir_near[(size_t)6 * i + k];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 9; ++k) {
    double t = wave_sum(acc[k]);
    if (lane == 0) lds[wave][k] = t;
  }
  __syncthreads();
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int w = 0; w < (int)(blockDim.x / 64); ++w)
    for (int k = 0; k < 9; ++k) m[k] += lds[w][k];
  for (int k = 0; k < 9; ++k) m[k] /= (double)s.n;
  __syncthreads();
  for (int i = threadIdx.x; i < s.n; i += blockDim.x) {
    s.fnx[i] -= m[0];
    s.fny[i] -= m[1];
    s.fnz[i] -= m[2];
    if (vir)
      for (int k = 0; k < 6; ++k) s.vir_near[(size_t)6 * i + k] -= m[3 + k];
  }
}

int pair_blocks(const SysView& s) {
  if (s.rbl && !s.lj_only) return (s.n + kBlock - 1) / kBlock;
  return (int)(((size_t)s.n << s.lanes_lg) + kBlock - 1) / kBlock;
}

void launch_pair(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int* nblocks) {
  const int grid = pair_blocks(s);
  if (nblocks) *nblocks = grid;
  if (s.rbl && !s.lj_only) {
    if (s.weighted)
      hipLaunchKernelGGL(k_pair_rbl<true>, dim3(grid), dim3(kBlock), 0, st, s);
    else
      hipLaunchKernelGGL(k_pair_rbl<false>, dim3(grid), dim3(kBlock), 0, st, s);
    hipLaunchKernelGGL(k_rbl_mean, dim3(1), dim3(1024), 0, st, s);
    return;
  }
  if (s.lj_only)
    launch_pair_lg<2, 0>(s, st, thermo, peratom, rows, grid);
  else if (!s.weighted)
    launch_pair_lg<0, 0>(s, st, thermo, peratom, rows, grid);
  else if (s.sp_max <= 4)
    launch_pair_lg<1, 4>(s, st, thermo, peratom, rows, grid);
  else
    launch_pair_lg<1, 16>(s, st, thermo, peratom, rows, grid);
}

}  // namespace rbmd
