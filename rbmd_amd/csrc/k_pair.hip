// k_pair.hip -- real-space LJ + erfc pair forces over the neighbour lists.
//
// L consecutive threads share one atom (cell-sorted order); each lane walks the
// atom's list row, whose 9 column segments it treats as one concatenated list
// (lane `sub` takes positions sub, sub + L, ...), gathers neighbour positions
// from the sorted double4 array (resident in L2) and keeps its partial sums in
// registers; the L partials are combined with a fixed xor-shuffle tree.
// Forces are gathered, never scattered: no atomics, deterministic sums.
//
// Formulas (FP64; the mixed-precision variant evaluates the pair scalar in FP32):
//   VERLET plain    ComputeNearForceVerletVirialWorklet, RunWorklet.cpp:2890-2985
//   VERLET weighted ComputeNearForceVerletWeightVirialWorklet1, RunWorklet.cpp:2767-2888
//   LJ/CUT only     ComputeLJForceVerletWorklet, RunWorklet.cpp:498-575
//   RBL             ComputeNearForceRBLERF(SpecialBonds)Worklet, RunWorklet.cpp:982-1266
//   LJ      ExecForceFunction.h:412-435 (force), :818-841 (energy)
//   G(r)    ExecForceFunction.h:682-689, erfc(sqrt(alpha) r)/r^2 + 2 sqrt(alpha/pi) e^{-alpha r^2}/r
//   virial  ComputePairVirial_fix (+1/2, plain) / ComputePairVirial (-1/2, weighted, RBL), :304-332
#include <type_traits>

#include "dev.hpp"
#include "pack.hpp"
#include "reduce.hpp"
#include "pair.hpp"
#include "cells.hpp"

namespace rbmd {

constexpr int kMaxSeg = 9;  // VERLET list segments per atom (stencil columns)

// Raw buffer loads (32-bit byte offsets from a wave-uniform base, bounds
// checked by the hardware: a load past num_records returns zeros).  The pair
// loop's gathers and list reads take their addresses as one 32-bit shift or
// shift-add instead of 64-bit address arithmetic.
typedef __amdgpu_buffer_rsrc_t BufRsrc;
typedef unsigned BufV4 __attribute__((ext_vector_type(4)));
typedef double BufD2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ BufRsrc buf_rsrc(const void* p, unsigned bytes) {
  // word 3 of a gfx9 buffer resource: 32-bit data format, untyped loads
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double4 buf_double4(BufRsrc r, unsigned off) {
  const BufD2 a = __builtin_bit_cast(BufD2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  const BufD2 b = __builtin_bit_cast(BufD2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 0));
  return make_double4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ int buf_int(BufRsrc r, unsigned off) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// A lane's position in an atom's concatenated segments: `seg` (nseg past the
// end), the offset in it, its length (INT_MAX past the end) and the byte
// offset of the segment's start from the wavefront's first list row.
struct ListCursor {
  int seg, off, num;
  unsigned po;
};

// entry at the cursor (0 past the end: a list entry is never 0, its image
// code is at least 2 + (2 << 3) + (2 << 6)), then one stride forward; crossing
// into the next segment is the rare path.  seg_cnt: the workgroup's segment
// counts in LDS, one column per thread, so a crossing reads the next count at
// an address computed from the segment index.
__device__ __forceinline__ int cursor_take(ListCursor& c, BufRsrc rows, const int (&seg_cnt)[kMaxSeg][kBlock],
                                           int nseg, int seg_cap, int stride) {
  if (__builtin_expect(c.off >= c.num, 0)) {
    do {
      c.off -= c.num;
      ++c.seg;
      // past the last segment the offset lies beyond the buffer (2^31 and
      // more), where loads return 0
      c.po = c.seg < nseg ? c.po + 4u * (unsigned)seg_cap : 0x80000000u;
      c.num = c.seg < nseg ? seg_cnt[c.seg][threadIdx.x] : 0x7fffffff;
    } while (c.off >= c.num);
  }
  const int e = buf_int(rows, c.po + 4u * (unsigned)c.off);
  c.off += stride;
  return e;
}

// MODE 0: plain LJ+coul, 1: weighted (special bonds), 2: LJ only.  SP: the
// register budget for the atom's special list (>= the longest list).  LTAB:
// type-pair tables in LDS (ntypes <= kMaxTypes).  Two entries per lane per
// iteration; the neighbour positions of the next iteration are gathered
// while the current one is evaluated, and its entries one iteration earlier.
// F32 (mixed precision): the pair scalar (1/r, LJ, the Coulomb table) in FP32
// from an FP32 table in LDS; the pair vector, d2, the force and virial sums
// and the energies stay FP64 (each FP32 term is exact in FP64, so the sums
// carry no more rounding than the FP64 path's).
template <int MODE, bool THERMO, bool PERATOM, int SP, bool LTAB, bool F32>
__device__ __forceinline__ void pair_verlet_body(const SysView& s, double* rows, const int bid) {
#pragma clang fp contract(off)  // fused multiply-adds are written out (see pair_term)
  constexpr bool EN = THERMO || PERATOM;
  constexpr int SPD = SP > 0 ? SP : 1;
  using TT = typename std::conditional<F32, float, double>::type;
  using CT = typename std::conditional<F32, float2, double2>::type;
  constexpr int kTab = F32 ? kCoulTab32 : kCoulTab;
  __shared__ TT teps[LTAB ? kMaxTypes * kMaxTypes : 1], tsig[LTAB ? kMaxTypes * kMaxTypes : 1];
  __shared__ CT ctab[MODE == 2 ? 1 : kTab];
  __shared__ int seg_cnt[kMaxSeg][kBlock];  // each thread's list segment counts (one column per thread)
  const int nt = s.ntypes;
  if (LTAB)
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) {
      teps[k] = (TT)s.eps_ij[k];
      tsig[k] = (TT)s.sig6_ij[k];
    }
  if (MODE != 2) {
    const CT* src;
    if constexpr (F32)
      src = s.ctab32;
    else
      src = s.ctab;
    for (int k = threadIdx.x; k < kTab; k += blockDim.x) ctab[k] = src[k];
  }
  __syncthreads();
  const int lg = s.lanes_lg;
  const int L = 1 << lg;
  const int t = bid * blockDim.x + threadIdx.x;
  const int i = t >> lg;
  const int sub = t & (L - 1);
  double fx = 0, fy = 0, fz = 0;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double pelj = 0, pec = 0;
  if (i < s.n) {
    const double4 pi = s.spos[i];
    const int ti = s.stype[i];
    int spj[SPD];
    double spw[SPD];
    if (MODE == 1) {
      const int id = s.point_ids[i];
      const int b = s.sp_off[id];
      int spn = s.sp_off[id + 1] - b;
      if (spn > SP) {
        atomicOr(s.status, 4);
        spn = SP;
      }
#pragma unroll
      for (int k = 0; k < SPD; ++k) {
        spj[k] = k < spn ? s.rank[s.sp_ids[b + k]] : -1;
        spw[k] = k < spn ? s.sp_w[b + k] : 1.0;
      }
    }
#pragma unroll
    for (int g = 0; g < kMaxSeg; ++g) seg_cnt[g][threadIdx.x] = g < s.nseg ? s.nbr_num[g * s.n + i] : 0;
    const int cnt0 = s.nseg > 0 ? s.nbr_num[i] : 0;
    const int nseg = s.nseg, seg_cap = s.seg_cap;
    // the wavefront's list rows (its 64 / L atoms, fewer in the last one)
    // and the sorted positions and types as buffers
    const int i0 = __builtin_amdgcn_readfirstlane(i);
    const int rows_here = min(kWave >> lg, s.n - i0);
    const BufRsrc rrows = buf_rsrc(s.nbr + (size_t)i0 * s.cap, 4u * (unsigned)s.cap * (unsigned)rows_here);
    const BufRsrc rpos = buf_rsrc(s.spos, 32u * (unsigned)s.n);
    const BufRsrc rtyp = buf_rsrc(s.stype, 4u * (unsigned)s.n);
    ListCursor cur{0, sub, cnt0, 4u * (unsigned)(i - i0) * (unsigned)s.cap};
    // pipeline: entries two iterations ahead, positions one iteration ahead
    int ea = cursor_take(cur, rrows, seg_cnt, nseg, seg_cap, L);
    int eb = cursor_take(cur, rrows, seg_cnt, nseg, seg_cap, L);
    // an entry's neighbour as a byte offset into the positions (0 past the
    // end: atom 0, loaded and not used)
    auto jof = [](int e) -> unsigned { return (unsigned)e >> kShiftBits; };
    double4 pa = buf_double4(rpos, 32u * jof(ea));
    double4 pb = buf_double4(rpos, 32u * jof(eb));
    int ta = buf_int(rtyp, 4u * jof(ea));
    int tb = buf_int(rtyp, 4u * jof(eb));
    int na = cursor_take(cur, rrows, seg_cnt, nseg, seg_cap, L);
    int nb = cursor_take(cur, rrows, seg_cnt, nseg, seg_cap, L);
    while (ea > 0) {
      const double4 qa = buf_double4(rpos, 32u * jof(na));
      const double4 qb = buf_double4(rpos, 32u * jof(nb));
      const int ua = buf_int(rtyp, 4u * jof(na));
      const int ub = buf_int(rtyp, 4u * jof(nb));
      const int ma = cursor_take(cur, rrows, seg_cnt, nseg, seg_cap, L);
      const int mb = cursor_take(cur, rrows, seg_cnt, nseg, seg_cap, L);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int e = h == 0 ? ea : eb;
        if (e <= 0) continue;
        const double4 p = h == 0 ? pa : pb;
        const int tj = h == 0 ? ta : tb;
        int j;
        double ox, oy, oz;
        decode1(s, e, j, ox, oy, oz);  // the build's image offset, bit for bit (stencil half width 1)
        double rx = (p.x - ox) - pi.x;
        double ry = (p.y - oy) - pi.y;
        double rz = (p.z - oz) - pi.z;
        if (s.use_min_image) min_image(s, rx, ry, rz);
        const double d2 = dist2(rx, ry, rz);
        const int tix = ti * nt + tj;
        const TT epsij = LTAB ? teps[tix] : (TT)s.eps_ij[tix];
        const TT sig6 = LTAB ? tsig[tix] : (TT)s.sig6_ij[tix];
        double w = 1.0;
        if (MODE == 1) {
#pragma unroll
          for (int u = 0; u < SPD; ++u)
            if (spj[u] == j) w = spw[u];
        }
        double sc, elj, ec;
        // list entries satisfy the acceptance test with this same vector;
        // with the minimum image applied the range is tested again, and a
        // skin list's entries (out to rc + skin) meet the list test here
        if constexpr (F32) {
          float el32, ec32;
          const float qq = (float)(pi.w * p.w);
          if (s.use_min_image)
            sc = pair_term32<MODE, EN, 1>(s, ctab, d2, qq, epsij, sig6, (float)w, el32, ec32);
          else if (s.skin_on)
            sc = pair_term32<MODE, EN, 2>(s, ctab, d2, qq, epsij, sig6, (float)w, el32, ec32);
          else
            sc = pair_term32<MODE, EN, 0>(s, ctab, d2, qq, epsij, sig6, (float)w, el32, ec32);
          elj = el32;
          ec = ec32;
        } else {
          if (s.use_min_image)
            sc = pair_term<MODE, EN, 1>(s, ctab, d2, pi.w * p.w, epsij, sig6, w, elj, ec);
          else if (s.skin_on)
            sc = pair_term<MODE, EN, 2>(s, ctab, d2, pi.w * p.w, epsij, sig6, w, elj, ec);
          else
            sc = pair_term<MODE, EN, 0>(s, ctab, d2, pi.w * p.w, epsij, sig6, w, elj, ec);
        }
        // one fused rounding per force term
        fx = fma(sc, rx, fx);
        fy = fma(sc, ry, fy);
        fz = fma(sc, rz, fz);
        if (EN) {
          const double px = sc * rx, py = sc * ry, pz = sc * rz;
          vir_add(vir, rx, ry, rz, px, py, pz, 1.0);  // the +-1/2 is applied once at the end
          pelj += elj;
          pec += ec;
        }
      }
      ea = na;
      eb = nb;
      pa = qa;
      pb = qb;
      ta = ua;
      tb = ub;
      na = ma;
      nb = mb;
    }
  }
  if (EN) {
    // a power-of-two scale commutes exactly with the sums
    const double vs = MODE == 1 ? -0.5 : 0.5;
#pragma unroll
    for (int k = 0; k < 6; ++k) vir[k] *= vs;
  }
  if (THERMO) {
    const double vals[kRowPair] = {pelj, pec, vir[0], vir[1], vir[2], vir[3], vir[4], vir[5]};
    block_write_row<kRowPair>(vals, rows + (size_t)bid * kRowPair);
  }
  for (int off = 1; off < L; off <<= 1) {
    fx += __shfl_xor(fx, off, kWave);
    fy += __shfl_xor(fy, off, kWave);
    fz += __shfl_xor(fz, off, kWave);
    if (PERATOM) {
      for (int k = 0; k < 6; ++k) vir[k] += __shfl_xor(vir[k], off, kWave);
      pelj += __shfl_xor(pelj, off, kWave);
      pec += __shfl_xor(pec, off, kWave);
    }
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

template <int MODE, bool THERMO, bool PERATOM, int SP, bool LTAB, bool F32>
__global__ __launch_bounds__(kBlock) void k_pair_verlet(SysView s, double* rows) {
  pair_verlet_body<MODE, THERMO, PERATOM, SP, LTAB, F32>(s, rows, (int)blockIdx.x);
}

template <int MODE, int SP, bool LTAB, bool F32>
void launch_pair_v(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int grid) {
  if (peratom)
    hipLaunchKernelGGL((k_pair_verlet<MODE, true, true, SP, LTAB, F32>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  else if (thermo)
    hipLaunchKernelGGL((k_pair_verlet<MODE, true, false, SP, LTAB, F32>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  else
    hipLaunchKernelGGL((k_pair_verlet<MODE, false, false, SP, LTAB, F32>), dim3(grid), dim3(kBlock), 0, st, s,
                       rows);
}

// the mixed-precision kernels are built with LDS type tables only (the host
// refuses mixed precision with more than kMaxTypes types)
template <int MODE, int SP>
void launch_pair_t(const SysView& s, hipStream_t st, int thermo, int peratom, double* rows, int grid) {
  if (s.mixed)
    launch_pair_v<MODE, SP, true, true>(s, st, thermo, peratom, rows, grid);
  else if (s.ntypes <= kMaxTypes)
    launch_pair_v<MODE, SP, true, false>(s, st, thermo, peratom, rows, grid);
  else
    launch_pair_v<MODE, SP, false, false>(s, st, thermo, peratom, rows, grid);
}

// RBL: per-atom corrected force before the mean removal; the virial keeps the
// reference's assignment of the running-sum force of the last pair (hazard Q5).
template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_pair_rbl(SysView s) {
#pragma clang fp contract(off)  // the pair vector and d2 rounded as the list build and the oracle round them
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
      decode(s, s.nbr[nbr_index(base + k, i, s.cap)], j, ox, oy, oz);
      const double4 pj = s.spos[j];
      double rx = (pj.x - ox) - pi.x;
      double ry = (pj.y - oy) - pi.y;
      double rz = (pj.z - oz) - pi.z;
      min_image(s, rx, ry, rz);
      const double d2 = dist2(rx, ry, rz);
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

// RBL energies.  The reference forms them every RBL step over a full Verlet
// list to rc (ExecutionMD::ComputeLJCoulEnergy, ExecutionMD.cpp:670-750:
// ComputeNeighbours over the RBL locator, then LJCoulEnergyVerletVirialWorklet
// RunWorklet.cpp:3145-3221, or SpecialLJCoulEnergyVerletWeightVirialWorklet
// :3047-3143 with special bonds): per pair with rc^2 - d^2 > 1e-4, d^2 > 1e-4
// on the cell-offset vector, the minimum-image vector gives
//   e_lj   += w * 4 eps (s6/d6 - 1) s6/d6 / 2            (ComputePotentialEn_fix)
//   e_coul += qqr2e (q q erfc(sqrt(alpha) r) / 2r - (1 - w) q q / 2r)
// with the analytic erfc at the configured alpha (not the step table).  No
// list is stored: one wavefront per atom walks the (2nc+1)^3 stencil around
// the cell its current position lies in (InWhichCell, so misfit atoms need no
// separate path), lanes over each cell's atoms.  Runs on the steps that form
// thermo sums; its rows also carry the mean-removed RBL virial of each atom.
template <bool WEIGHTED, bool PERATOM>
__global__ __launch_bounds__(kBlock) void k_rbl_energy(SysView s, double* rows) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave);
  double elj = 0.0, ec = 0.0;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  if (i < s.n) {
    const double4 pi = s.spos[i];
    const int ti = s.stype[i];
    int spj[kMaxSpecial];
    double spw[kMaxSpecial];
    if (WEIGHTED) {
      const int id = s.point_ids[i];
      const int b = s.sp_off[id];
      const int spn = min(s.sp_off[id + 1] - b, kMaxSpecial);  // longer lists are refused at creation
#pragma unroll
      for (int k = 0; k < kMaxSpecial; ++k) {
        spj[k] = k < spn ? s.rank[s.sp_ids[b + k]] : -1;
        spw[k] = k < spn ? s.sp_w[b + k] : 1.0;
      }
    }
    const int ca = cell_axis(pi.x, s.lo[0], s.dxdydz[0]);
    const int cb = cell_axis(pi.y, s.lo[1], s.dxdydz[1]);
    const int cc = cell_axis(pi.z, s.lo[2], s.dxdydz[2]);
    const int nc = s.nc;
    for (int a = -nc; a <= nc; ++a)
      for (int b = -nc; b <= nc; ++b)
        for (int e = -nc; e <= nc; ++e) {
          const StencilCell t = stencil_cell(s, ca, cb, cc, a, b, e);
          if (!t.ok) continue;  // an atom more than a box away: flagged by the force pass
          for (int base = t.js; base < t.je; base += kWave) {
            const int j = base + lane;
            if (j >= t.je) continue;
            const double4 pj = s.spos[j];
            double rx = (pj.x - t.ox) - pi.x;
            double ry = (pj.y - t.oy) - pi.y;
            double rz = (pj.z - t.oz) - pi.z;
            double d2 = dist2(rx, ry, rz);
            if (!(s.cut2 - d2 > 0.0001 && d2 > 0.0001)) continue;
            min_image(s, rx, ry, rz);  // MinDistanceVec(p_j, p_i)
            d2 = dist2(rx, ry, rz);
            if (!(d2 < s.cut2 && d2 > 0.0001)) continue;
            const int tix = ti * s.ntypes + s.stype[j];
            const double eps = s.eps_ij[tix], s6 = s.sig6_ij[tix];
            const double d6 = d2 * d2 * d2;
            const double el = 0.5 * (4.0 * eps * (s6 / d6 - 1.0) * (s6 / d6));
            const double dis = sqrt(d2);
            const double qq = pi.w * pj.w;
            const double eco = 0.5 * qq * erfc(s.sqrt_alpha * dis) / dis;
            double w = 1.0;
            if (WEIGHTED) {
#pragma unroll
              for (int u = 0; u < kMaxSpecial; ++u)
                if (spj[u] == j) w = spw[u];
              elj += w * el;
              ec += s.qqr2e * (eco - (1.0 - w) * (0.5 * qq / dis));
            } else {
              elj += el;
              ec += s.qqr2e * eco;
            }
          }
        }
    elj = wave_sum(elj);
    ec = wave_sum(ec);
    if (lane == 0) {
      if (PERATOM) {
        s.e_lj[i] = elj;
        s.e_coul[i] = ec;
      }
      for (int k = 0; k < 6; ++k) vir[k] = s.vir_near[(size_t)6 * i + k];
    } else {
      elj = ec = 0.0;
    }
  }
  const double vals[kRowPair] = {elj, ec, vir[0], vir[1], vir[2], vir[3], vir[4], vir[5]};
  block_write_row<kRowPair>(vals, rows + (size_t)blockIdx.x * kRowPair);
}

int rbl_energy_blocks(const SysView& s) { return (s.n + kBlock / kWave - 1) / (kBlock / kWave); }

void launch_rbl_energy(const SysView& s, hipStream_t st, int peratom, double* rows, int* nblocks) {
  const int grid = rbl_energy_blocks(s);
  if (nblocks) *nblocks = grid;
  if (s.weighted) {
    if (peratom)
      hipLaunchKernelGGL((k_rbl_energy<true, true>), dim3(grid), dim3(kBlock), 0, st, s, rows);
    else
      hipLaunchKernelGGL((k_rbl_energy<true, false>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  } else {
    if (peratom)
      hipLaunchKernelGGL((k_rbl_energy<false, true>), dim3(grid), dim3(kBlock), 0, st, s, rows);
    else
      hipLaunchKernelGGL((k_rbl_energy<false, false>), dim3(grid), dim3(kBlock), 0, st, s, rows);
  }
}

__global__ __launch_bounds__(kBlock) void k_static_table(const double* dis, double* g, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = gnear_table(dis[i]);
}

void launch_static_table(const double* dis, double* g, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_static_table, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, dis, g, n);
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
    launch_pair_t<2, 0>(s, st, thermo, peratom, rows, grid);
  else if (!s.weighted)
    launch_pair_t<0, 0>(s, st, thermo, peratom, rows, grid);
  else if (s.sp_max <= 2)  // rigid water: two partners per atom
    launch_pair_t<1, 2>(s, st, thermo, peratom, rows, grid);
  else if (s.sp_max <= 4)
    launch_pair_t<1, 4>(s, st, thermo, peratom, rows, grid);
  else
    launch_pair_t<1, 16>(s, st, thermo, peratom, rows, grid);
}

// ------------------------------------------------------------------ packs
// per-atom outputs off, LDS type tables (packs need ntypes <= kMaxTypes);
// thermo sums on the steps that report them

template <int MODE, int SP, bool F32, bool THERMO>
__global__ __launch_bounds__(kBlock) void kp_pair(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  pair_verlet_body<MODE, THERMO, false, SP, true, F32>(e.v, e.rows_pair, bid);
}

template <int MODE, int SP, bool F32>
void launch_kp_pair_t(const PackLaunch& L, int nblocks, int thermo, hipStream_t st) {
  if (thermo)
    hipLaunchKernelGGL((kp_pair<MODE, SP, F32, true>), dim3(nblocks), dim3(kBlock), 0, st, PACK_ARGS(L));
  else
    hipLaunchKernelGGL((kp_pair<MODE, SP, F32, false>), dim3(nblocks), dim3(kBlock), 0, st, PACK_ARGS(L));
}

template <int MODE, int SP>
void launch_kp_pair(const PackLaunch& L, int nblocks, int f32, int thermo, hipStream_t st) {
  if (f32)
    launch_kp_pair_t<MODE, SP, true>(L, nblocks, thermo, st);
  else
    launch_kp_pair_t<MODE, SP, false>(L, nblocks, thermo, st);
}

void launch_pack_pair(const PackLaunch& L, int nblocks, const PackSig& g, hipStream_t st) {
  if (nblocks <= 0) return;
  if (g.pair_mode == 2)
    launch_kp_pair<2, 0>(L, nblocks, g.f32, g.thermo, st);
  else if (g.pair_mode == 0)
    launch_kp_pair<0, 0>(L, nblocks, g.f32, g.thermo, st);
  else if (g.pair_sp <= 2)
    launch_kp_pair<1, 2>(L, nblocks, g.f32, g.thermo, st);
  else if (g.pair_sp <= 4)
    launch_kp_pair<1, 4>(L, nblocks, g.f32, g.thermo, st);
  else
    launch_kp_pair<1, 16>(L, nblocks, g.f32, g.thermo, st);
}

}  // namespace rbmd
