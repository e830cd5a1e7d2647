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

namespace rbmd {

__device__ __forceinline__ double gnear(double alpha, double sqrt_alpha, double dis) {
  return erfc(sqrt_alpha * dis) / (dis * dis) + 2.0 * sqrt_alpha * exp(-alpha * dis * dis) / (sqrt(kPi) * dis);
}

// step table of the RBL path: G at alpha = 0.1 evaluated at the lower edge of
// the FP32 bucket given by 3 exponent + 8 mantissa bits (ExecStaticTable.h:38-39,127-135)
__device__ __forceinline__ double gnear_table(double dis) {
  float fd = (float)dis;
  unsigned bits = __float_as_uint(fd) & 0xFFFF8000u;
  double lo = (double)__uint_as_float(bits);
  return gnear(0.1, sqrt(0.1), lo);
}

__device__ __forceinline__ void min_image(const SysView& s, double& rx, double& ry, double& rz) {
  if (fabs(rx) > s.box[0] * 0.5) rx -= (rx > 0 ? s.box[0] : -s.box[0]);
  if (fabs(ry) > s.box[1] * 0.5) ry -= (ry > 0 ? s.box[1] : -s.box[1]);
  if (fabs(rz) > s.box[2] * 0.5) rz -= (rz > 0 ? s.box[2] : -s.box[2]);
}

__device__ __forceinline__ void decode(const SysView& s, int e, int& j, double& ox, double& oy, double& oz) {
  j = e >> kShiftBits;
  int code = e & ((1 << kShiftBits) - 1);
  ox = s.dxdydz[0] * (double)(((code & 7) - 3) * s.dims[0]);
  oy = s.dxdydz[1] * (double)((((code >> 3) & 7) - 3) * s.dims[1]);
  oz = s.dxdydz[2] * (double)((((code >> 6) & 7) - 3) * s.dims[2]);
}

__device__ __forceinline__ void vir_add(double* v, double rx, double ry, double rz, double fx, double fy, double fz,
                                        double scale) {
  fx *= scale;
  fy *= scale;
  fz *= scale;
  v[0] += rx * fx;
  v[1] += ry * fy;
  v[2] += rz * fz;
  v[3] += rx * fy;
  v[4] += rx * fz;
  v[5] += ry * fz;
}

template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = 1; off < W; off <<= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// One pair term; returns the scalar f such that the force on i is f * r.
// Reference arithmetic: LJ ExecForceFunction.h:412-435 / 818-841, G(r) :682-689,
// special weights RunWorklet.cpp:2827-2857.
template <int MODE, bool EN>
__device__ __forceinline__ double pair_term(const SysView& s, double rx, double ry, double rz, double d2, double qq,
                                            double epsij, double sig6, double w, double& elj, double& ec) {
  const bool in = d2 < s.cut2 && d2 > 0.0001;
  double flj = 0, fc = 0;
  elj = 0;
  ec = 0;
  if (in) {
    double d6 = d2 * d2 * d2;
    double sd6 = sig6 / d6;
    flj = -24.0 * epsij * ((2.0 * sd6 - 1.0) * sd6) / d2;
    if (EN) elj = 0.5 * (4.0 * epsij * (sd6 - 1.0) * sd6);
    if (MODE != 2) {
      double rinv = rsqrt(d2);
      double dis = d2 * rinv;
      double erfcv = erfc(s.sqrt_alpha * dis);
      double g = erfcv * (rinv * rinv) + (2.0 / sqrt(kPi)) * s.sqrt_alpha * exp(-s.alpha * d2) * rinv;
      if (MODE == 0) {
        fc = s.qqr2e * (-qq * g * rinv);
        if (EN) ec = s.qqr2e * (0.5 * qq * erfcv * rinv);
      } else {
        fc = s.qqr2e * (-qq * g * rinv + (1.0 - w) * qq * (rinv * rinv * rinv));
        if (EN) ec = s.qqr2e * (0.5 * qq * erfcv * rinv - (1.0 - w) * (0.5 * qq * rinv));
      }
    }
  }
  if (MODE == 1) {
    flj *= w;
    elj *= w;
  }
  return fc + flj;
}

constexpr int kPairBatch = 4;
constexpr int kMaxTypes = 16;

// MODE 0: plain LJ+coul, 1: weighted (special bonds), 2: LJ only.
// kLanes consecutive threads share one atom: lane `sub` takes list slots
// sub, sub + kLanes, ... of every segment, kPairBatch slots at a time so that
// several neighbour gathers are in flight per lane, and the partial sums are
// combined with a fixed xor-shuffle tree at the end (deterministic).
template <int MODE, bool THERMO, bool PERATOM>
__global__ __launch_bounds__(kBlock) void k_pair_verlet(SysView s, double* rows) {
  __shared__ double tab_eps[kMaxTypes * kMaxTypes], tab_sig[kMaxTypes * kMaxTypes];
  const int nt = s.ntypes;
  const bool lds_tab = nt <= kMaxTypes;
  if (lds_tab)
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) {
      tab_eps[k] = s.eps_ij[k];
      tab_sig[k] = s.sig6_ij[k];
    }
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / kLanes;
  const int sub = t % kLanes;
  constexpr bool EN = THERMO || PERATOM;
  double fx = 0, fy = 0, fz = 0;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double pelj = 0, pec = 0;
  if (i < s.n) {
    const double4 pi = s.spos[i];
    const int ti = s.stype[i];
    const double qi = pi.w;
    int spn = 0;
    int spj[kMaxSpecial];
    double spw[kMaxSpecial];
    if (MODE == 1) {
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
    for (int seg = 0; seg < s.nseg; ++seg) {
      const int num = s.nbr_num[seg * s.n + i];
      const int base = seg * s.seg_cap;
      for (int k0 = sub; k0 < num; k0 += kLanes * kPairBatch) {
        int ent[kPairBatch];
#pragma unroll
        for (int bb = 0; bb < kPairBatch; ++bb) {
          int k = k0 + bb * kLanes;
          ent[bb] = k < num ? s.nbr[nbr_index(base + k, i, s.n)] : -1;
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
            for (int u = 0; u < kMaxSpecial; ++u)
              if (spj[u] == j) w = spw[u];
          }
          double elj, ec;
          const double sc = pair_term<MODE, EN>(s, rx, ry, rz, d2, qi * pj[bb].w, epsij, sig6, w, elj, ec);
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
    const double vals[8] = {pelj, pec, vir[0], vir[1], vir[2], vir[3], vir[4], vir[5]};
    const int slots[8] = {S_ELJ, S_ECOUL, S_VIR_NEAR, S_VIR_NEAR + 1, S_VIR_NEAR + 2,
                          S_VIR_NEAR + 3, S_VIR_NEAR + 4, S_VIR_NEAR + 5};
    block_write_row<8>(vals, slots, rows + (size_t)blockIdx.x * S_NSLOT);
  }
  fx = group_sum<kLanes>(fx);
  fy = group_sum<kLanes>(fy);
  fz = group_sum<kLanes>(fz);
  if (PERATOM) {
    for (int k = 0; k < 6; ++k) vir[k] = group_sum<kLanes>(vir[k]);
    pelj = group_sum<kLanes>(pelj);
    pec = group_sum<kLanes>(pec);
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
      decode(s, s.nbr[nbr_index(base + k, i, s.n)], j, ox, oy, oz);
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

void launch_pair(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int* nblocks) {
  int grid_rbl = (s.n + kBlock - 1) / kBlock;
  int grid = (s.n * kLanes + kBlock - 1) / kBlock;
  if (nblocks) *nblocks = grid;
  if (s.rbl && !s.lj_only) {
    grid = grid_rbl;
    if (nblocks) *nblocks = grid;
    if (s.weighted)
      hipLaunchKernelGGL(k_pair_rbl<true>, dim3(grid), dim3(kBlock), 0, st, s);
    else
      hipLaunchKernelGGL(k_pair_rbl<false>, dim3(grid), dim3(kBlock), 0, st, s);
    hipLaunchKernelGGL(k_rbl_mean, dim3(1), dim3(1024), 0, st, s);
    return;
  }
  const int mode = s.lj_only ? 2 : (s.weighted ? 1 : 0);
#define RBMD_PAIR(M)                                                                                  \
  if (peratom)                                                                                        \
    hipLaunchKernelGGL((k_pair_verlet<M, true, true>), dim3(grid), dim3(kBlock), 0, st, s, rows);     \
  else if (thermo)                                                                                    \
    hipLaunchKernelGGL((k_pair_verlet<M, true, false>), dim3(grid), dim3(kBlock), 0, st, s, rows);    \
  else                                                                                                \
    hipLaunchKernelGGL((k_pair_verlet<M, false, false>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  if (mode == 0) {
    RBMD_PAIR(0)
  } else if (mode == 1) {
    RBMD_PAIR(1)
  } else {
    RBMD_PAIR(2)
  }
#undef RBMD_PAIR
}

}  // namespace rbmd
