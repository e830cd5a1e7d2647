// k_kspace.hip -- random-batch Ewald structure factor and k-space force.
//
// rho(k_l) = sum_i q_i (cos k_l.r_i, sin k_l.r_i)
//   ComputePnumberChargeStructureFactorWorklet + ReduceByKey (RunWorklet.cpp:2209-2291,
//   ExecutionMD.cpp:469-501).  The reference writes a P x N scratch and reduces it by
//   key; here each thread folds 4 atoms, the workgroup reduces with wavefront
//   shuffles (no atomics), and a second pass sums the per-workgroup partials in
//   a fixed order.
// F_i = (S/P) sum_l -(4 pi q_i / (V k_l^2)) (cos(k.r) Im rho - sin(k.r) Re rho) k_l, x qqr2e
//   ComputeNewRBEForceWorklet (RunWorklet.cpp:2086-2141), ComputeRBEForceSum_fix
//   (ExecForceFunction.h:748-762), ComputeKspaceVirial (:397-410), UnitRescale
//   (ExecutionNVT.cpp:718).  With an Ewald prefactor e^{-k^2/4 alpha} and scale 1
//   the same kernel is ComputeNewFarElectrostaticsWorklet (RunWorklet.cpp:2492-2540).
#include "dev.hpp"
#include "pack.hpp"
#include "reduce.hpp"
#include "cells.hpp"
#include "sincos.hpp"

namespace rbmd {

// e^{i k.r} with k = 2 pi m / L is e^{2 pi i t} with the phase t = m.(r/L) in
// turns (several hundred radians at |m| ~ 30): sincos_turns reduces it
// exactly to a quadrant and a remainder of at most 1/8 turn before its
// polynomials, so no large-argument reduction is needed.  F32 (mixed
// precision): the phase reduced exactly in FP64 to at most 1/2 turn, rounded
// to FP32, and the hardware sin / cos (sincos_turns32_hw, |error| <= 1.2e-7);
// the phase, the reduction and every sum stay FP64.
template <bool F32>
__device__ __forceinline__ void phase_sincos(double mx, double my, double mz, double sx, double sy, double sz,
                                             double* sn, double* cs) {
  const double t = fma(mz, sz, fma(my, sy, mx * sx));
  if constexpr (F32) {
    float s32, c32;
    sincos_turns32_hw(t, &s32, &c32);
    *sn = s32;
    *cs = c32;
  } else {
    sincos_turns(t, sn, cs);
  }
}

// m = k L / (2 pi), an integer held in k to ~1e-13: the product with the
// reciprocal rounds to the same integer as the quotient
__device__ __forceinline__ double k_to_m(double k, double box) { return rint(k * box * (0.5 / kPi)); }

// Structure factor partials.  One workgroup: kWave * KB k vectors (KB per
// lane, a tile choice, SysView::rho_kb) against kv.rho_atoms atoms (64 to
// 1024, by system size: rho_atoms_per_block), which its kBlock / kWave
// wavefronts split into contiguous ranges.
// The lanes of a wavefront take the same atom at the same time, so the atom
// is wave-uniform and comes through scalar loads; the wavefronts' partial
// sums are added in wavefront order and written as row ba of rho_part.
// bk: block of k vectors, ba: block of atoms.
template <bool F32, int KB>
__device__ __forceinline__ void rho_partial_body(const SysView& s, const KView& kv, const int bk, const int ba) {
#pragma clang fp contract(off)  // fused multiply-adds are written out: the same rounding in every kernel
  constexpr int NW = kBlock / kWave;
  __shared__ double part[NW][2 * KB][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int kb0 = bk * kWave * KB;
  double mx[KB], my[KB], mz[KB], re[KB], im[KB];
#pragma unroll
  for (int t = 0; t < KB; ++t) {
    const int l = kb0 + t * kWave + lane;
    const bool ok = l < kv.P;
    mx[t] = ok ? k_to_m(kv.kvec[3 * l], s.box[0]) : 0.0;
    my[t] = ok ? k_to_m(kv.kvec[3 * l + 1], s.box[1]) : 0.0;
    mz[t] = ok ? k_to_m(kv.kvec[3 * l + 2], s.box[2]) : 0.0;
    re[t] = im[t] = 0.0;
  }
  const double ibx = 1.0 / s.box[0], iby = 1.0 / s.box[1], ibz = 1.0 / s.box[2];
  if constexpr (F32) {
    // mixed precision: the phase as (m / L).r, so the atom's coordinates are
    // not scaled per atom (a change of ~1e-15 turns in the phase)
#pragma unroll
    for (int t = 0; t < KB; ++t) {
      mx[t] *= ibx;
      my[t] *= iby;
      mz[t] *= ibz;
    }
  }
  const int per = kv.rho_atoms / NW;
  const int a0 = ba * kv.rho_atoms + w * per;
  const int a1 = a0 + per < s.n ? a0 + per : s.n;
  // constant address space: wave-uniform scalar loads (positions and charges
  // are not written during this kernel)
  typedef const __attribute__((address_space(4))) double const_double;
  const const_double* X = (const const_double*)s.x;
  const const_double* Y = (const const_double*)s.y;
  const const_double* Z = (const const_double*)s.z;
  const const_double* Q = (const const_double*)s.q;
#pragma unroll 2
  for (int i = a0; i < a1; ++i) {
    const double sx = F32 ? X[i] : X[i] * ibx, sy = F32 ? Y[i] : Y[i] * iby, sz = F32 ? Z[i] : Z[i] * ibz;
    const double q = Q[i];
#pragma unroll
    for (int t = 0; t < KB; ++t) {
      double sn, cs;
      phase_sincos<F32>(mx[t], my[t], mz[t], sx, sy, sz, &sn, &cs);
      re[t] = fma(q, cs, re[t]);
      im[t] = fma(q, sn, im[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < KB; ++t) {
    part[w][2 * t][lane] = re[t];
    part[w][2 * t + 1][lane] = im[t];
  }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int t = 0; t < KB; ++t) {
      const int l = kb0 + t * kWave + lane;
      if (l < kv.P) {
        double a = part[0][2 * t][lane], b = part[0][2 * t + 1][lane];
        for (int u = 1; u < NW; ++u) {
          a += part[u][2 * t][lane];
          b += part[u][2 * t + 1][lane];
        }
        kv.rho_part[(size_t)ba * 2 * kv.P + 2 * l] = a;
        kv.rho_part[(size_t)ba * 2 * kv.P + 2 * l + 1] = b;
      }
    }
  }
}

// k-vector blocks on x (a converged Ewald grid has millions), atom blocks on y
template <bool F32, int KB>
__global__ __launch_bounds__(kBlock) void k_rho_partial(SysView s, KView kv) {
  rho_partial_body<F32, KB>(s, kv, (int)blockIdx.x, (int)blockIdx.y);
}

int rho_kblocks(const SysView& s, int P) { return (P + kWave * s.rho_kb - 1) / (kWave * s.rho_kb); }

int rho_atoms_per_block(int n, int kblocks) {
  int a = 1024;
  while (a > 64 && (long)kblocks * ((n + a - 1) / a) < 2048) a >>= 1;
  return a;
}

// instantiate F(F32, KB) for the tile choice kb (1, 2 or 4 k vectors per lane)
#define RHO_DISPATCH(kb, f32, F)                 \
  do {                                           \
    if ((kb) == 1) {                             \
      if (f32) F(true, 1); else F(false, 1);     \
    } else if ((kb) == 4) {                      \
      if (f32) F(true, 4); else F(false, 4);     \
    } else {                                     \
      if (f32) F(true, 2); else F(false, 2);     \
    }                                            \
  } while (0)

// rho[t] = sum over the partial rows b of rho_part[b][t].  One workgroup per
// kRedCols columns (one per lane, coalesced rows); wavefront w sums rows w,
// w + kRedWaves, ... (independent loads, in flight together), and the
// kRedWaves partial sums are added in wavefront order: a fixed order, and
// ~300 rows at 300,000 atoms take ~20 loads per wavefront instead of one
// thread's chain of 300.
constexpr int kRedCols = kWave;
constexpr int kRedWaves = 16;

__device__ __forceinline__ void rho_reduce_body(const SysView& s, const KView& kv, const int bid) {
#pragma clang fp contract(off)
  __shared__ double part[kRedWaves][kRedCols];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int n2 = 2 * kv.P;
  const int t = bid * kRedCols + lane;
  double acc = 0.0;
  if (t < n2)
    for (int b = w; b < kv.nblk_rho; b += kRedWaves) acc += kv.rho_part[(size_t)b * n2 + t];
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    double r = 0.0;
    if (t < n2) {
      r = part[0][lane];
      for (int v = 1; v < kRedWaves; ++v) r += part[v][lane];
      kv.rho[t] = r;
    }
    // Im rho from the odd neighbour lane (2P is even: both lanes of a pair
    // are in range or both out); the whole wavefront takes part
    const double r_im = __shfl_xor(r, 1);
    if (kv.krec && t < n2 && !(t & 1)) {
      KRec& rec = kv.krec[t >> 1];
      const int l = t >> 1;
      const double kx = kv.kvec[3 * l], ky = kv.kvec[3 * l + 1], kz = kv.kvec[3 * l + 2];
      const double k2 = kx * kx + ky * ky + kz * kz;
      const double pre = kv.ewald ? exp(-k2 / (4 * s.alpha)) / (s.volume * k2) : 1.0 / (s.volume * k2);
      rec.kx = kx;
      rec.ky = ky;
      rec.kz = kz;
      rec.pre = pre;
      rec.mx = k_to_m(kx, s.box[0]);
      rec.my = k_to_m(ky, s.box[1]);
      rec.mz = k_to_m(kz, s.box[2]);
      rec.rr = kv.f32 ? pre * r : r;
      rec.ri = kv.f32 ? pre * r_im : r_im;
      rec.pad = 0.0;
    }
  }
}

__global__ __launch_bounds__(kRedWaves * kWave) void k_rho_reduce(SysView s, KView kv) {
  rho_reduce_body(s, kv, (int)blockIdx.x);
}

// the k vectors of the current step (device step counter) from the
// double-buffered ring [2][R][3P] into a fixed buffer (step graphs read it)
__device__ __forceinline__ void kvec_slot_body(double* dst, const double* ring, int R, int per, const int64_t* step,
                                               const int bid, const int nblk) {
  const int64_t st = *step;
  const size_t slot = (size_t)((st / R) % 2) * R + (size_t)(st % R);
  for (int t = bid * blockDim.x + threadIdx.x; t < per; t += nblk * blockDim.x) dst[t] = ring[slot * per + t];
}

__global__ __launch_bounds__(kBlock) void k_kvec_slot(double* dst, const double* ring, int R, int per,
                                                      const int64_t* step) {
  kvec_slot_body(dst, ring, R, per, step, (int)blockIdx.x, (int)gridDim.x);
}

void launch_kvec_slot(double* dst, const double* ring, int R, int per, const int64_t* step, hipStream_t st) {
  hipLaunchKernelGGL(k_kvec_slot, dim3((per + kBlock - 1) / kBlock), dim3(kBlock), 0, st, dst, ring, R, per, step);
}

void launch_rho_partial(const SysView& s, const KView& k, hipStream_t st) {
  dim3 grid(rho_kblocks(s, k.P), k.nblk_rho);
#define RHO_LAUNCH(F32_, KB_) hipLaunchKernelGGL((k_rho_partial<F32_, KB_>), grid, dim3(kBlock), 0, st, s, k)
  RHO_DISPATCH(s.rho_kb, k.f32, RHO_LAUNCH);
#undef RHO_LAUNCH
}

void launch_rho_reduce(const SysView& s, const KView& k, hipStream_t st) {
  hipLaunchKernelGGL(k_rho_reduce, dim3((2 * k.P + kRedCols - 1) / kRedCols), dim3(kRedWaves * kWave), 0, st, s, k);
}

// Force pass.  One workgroup: 64 atoms (one per lane) against all P k
// vectors, split into KS = 2^LG contiguous ranges, one per wavefront (a tile
// choice, SysView::k_lanes_lg).  The lanes of a wavefront take the same k
// vector at the same time, so its record (KRec) is wave-uniform and comes
// through scalar loads; nothing of k passes through LDS.  The wavefronts'
// partial sums are added in wavefront order.
template <bool THERMO, bool PERATOM, bool F32, int LG>
__device__ __forceinline__ void kforce_body(const SysView& s, const KView& kv, double* rows, const int bid) {
#pragma clang fp contract(off)  // fused multiply-adds are written out: the same forces with or without the virial sums
  constexpr int KS = 1 << LG;
  constexpr int NX = PERATOM ? 9 : 3;  // per-atom sums the wavefronts hand over
  __shared__ double part[KS][NX][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int i = bid * kWave + lane;
  const bool act = i < s.n;
  double x = 0, y = 0, z = 0, q = 0, sx = 0, sy = 0, sz = 0;
  if (act) {
    x = s.x[i];
    y = s.y[i];
    z = s.z[i];
    q = s.q[i];
    sx = x * (1.0 / s.box[0]);
    sy = y * (1.0 / s.box[1]);
    sz = z * (1.0 / s.box[2]);
  }
  const double q4 = -4.0 * kPi * q;
  double fx = 0, fy = 0, fz = 0;
  double v[6] = {0, 0, 0, 0, 0, 0};
  const int per = (kv.P + KS - 1) / KS;
  const int l0 = w * per;
  const int l1 = l0 + per < kv.P ? l0 + per : kv.P;
  // constant address space: wave-uniform scalar loads (the records are not
  // written during this kernel)
  typedef const __attribute__((address_space(4))) KRec const_krec;
  const const_krec* rec = (const const_krec*)kv.krec;
  if constexpr (F32) {
    // Mixed precision: the term's coefficient pre Im rho cos - pre Re rho sin
    // from the record's products (FP64, so that the forces still sum to zero
    // over the atoms: sum_i q_i (a cos_i - b sin_i) = a Re rho - b Im rho
    // cancels only for a / b = Im rho / Re rho), the sums in FP64 and the
    // atom's -4 pi q applied to them.  The virial terms
    // (x k_a + x k_a) fs summed over k are 2 x F_a (and x F_b + y F_a off the
    // diagonal), formed from the wavefront's sums.  Two records per
    // iteration: their scalar loads are in flight together.
#pragma unroll 2
    for (int l = l0; l < l1; ++l) {
      const double t = fma(rec[l].mz, sz, fma(rec[l].my, sy, rec[l].mx * sx));
      float sn, cs;
      sincos_turns32_hw(t, &sn, &cs);
      const double g = fma((double)cs, rec[l].ri, -((double)sn * rec[l].rr));
      fx = fma(g, rec[l].kx, fx);
      fy = fma(g, rec[l].ky, fy);
      fz = fma(g, rec[l].kz, fz);
    }
    fx = q4 * fx;
    fy = q4 * fy;
    fz = q4 * fz;
    if (THERMO || PERATOM) {
      v[0] = (x + x) * fx;
      v[1] = (y + y) * fy;
      v[2] = (z + z) * fz;
      v[3] = x * fy + y * fx;
      v[4] = x * fz + z * fx;
      v[5] = y * fz + z * fy;
    }
  } else {
    for (int l = l0; l < l1; ++l) {
      const double kx = rec[l].kx, ky = rec[l].ky, kz = rec[l].kz;
      double sn, cs;
      phase_sincos<F32>(rec[l].mx, rec[l].my, rec[l].mz, sx, sy, sz, &sn, &cs);
      const double fs = (q4 * rec[l].pre) * fma(cs, rec[l].ri, -(sn * rec[l].rr));
      fx = fma(fs, kx, fx);
      fy = fma(fs, ky, fy);
      fz = fma(fs, kz, fz);
      if (THERMO || PERATOM) {
        v[0] += (x * kx + x * kx) * fs;
        v[1] += (y * ky + y * ky) * fs;
        v[2] += (z * kz + z * kz) * fs;
        v[3] += (x * ky + y * kx) * fs;
        v[4] += (x * kz + z * kx) * fs;
        v[5] += (y * kz + z * ky) * fs;
      }
    }
  }
  const double sc = kv.S_over_P;
  if (THERMO) {
    const double vals[kRowK] = {v[0] * sc, v[1] * sc, v[2] * sc, v[3] * sc, v[4] * sc, v[5] * sc};
    block_write_row<kRowK, KS>(vals, rows + (size_t)bid * kRowK);
  }
  // the atom's totals: wavefront partials added in order
  if (w > 0) {
    part[w][0][lane] = fx;
    part[w][1][lane] = fy;
    part[w][2][lane] = fz;
    if (PERATOM)
      for (int k = 0; k < 6; ++k) part[w][3 + k][lane] = v[k];
  }
  __syncthreads();
  if (w == 0 && act) {
    for (int u = 1; u < KS; ++u) {
      fx += part[u][0][lane];
      fy += part[u][1][lane];
      fz += part[u][2][lane];
      if (PERATOM)
        for (int k = 0; k < 6; ++k) v[k] += part[u][3 + k][lane];
    }
    s.fkx[i] = s.qqr2e * (fx * sc);
    s.fky[i] = s.qqr2e * (fy * sc);
    s.fkz[i] = s.qqr2e * (fz * sc);
    if (PERATOM)
      for (int k = 0; k < 6; ++k) s.vir_k[(size_t)6 * i + k] = v[k] * sc;
  }
}

template <bool THERMO, bool PERATOM, bool F32, int LG>
__global__ __launch_bounds__(kWave << LG) void k_kforce(SysView s, KView kv, double* rows) {
  kforce_body<THERMO, PERATOM, F32, LG>(s, kv, rows, (int)blockIdx.x);
}

int kforce_blocks(const SysView& s) { return (s.n + kWave - 1) / kWave; }

template <bool THERMO, bool PERATOM, bool F32>
void launch_kforce_lg(const SysView& s, const KView& k, hipStream_t st, double* rows, int grid) {
  if (s.k_lanes_lg == 1)
    hipLaunchKernelGGL((k_kforce<THERMO, PERATOM, F32, 1>), dim3(grid), dim3(kWave << 1), 0, st, s, k, rows);
  else if (s.k_lanes_lg == 3)
    hipLaunchKernelGGL((k_kforce<THERMO, PERATOM, F32, 3>), dim3(grid), dim3(kWave << 3), 0, st, s, k, rows);
  else
    hipLaunchKernelGGL((k_kforce<THERMO, PERATOM, F32, 2>), dim3(grid), dim3(kWave << 2), 0, st, s, k, rows);
}

void launch_kforce(const SysView& s, const KView& k, hipStream_t st, int thermo, int peratom, double* rows,
                   int* nblocks) {
  int grid = kforce_blocks(s);
  if (nblocks) *nblocks = grid;
  if (k.f32) {
    if (peratom)
      launch_kforce_lg<true, true, true>(s, k, st, rows, grid);
    else if (thermo)
      launch_kforce_lg<true, false, true>(s, k, st, rows, grid);
    else
      launch_kforce_lg<false, false, true>(s, k, st, rows, grid);
    return;
  }
  if (peratom)
    launch_kforce_lg<true, true, false>(s, k, st, rows, grid);
  else if (thermo)
    launch_kforce_lg<true, false, false>(s, k, st, rows, grid);
  else
    launch_kforce_lg<false, false, false>(s, k, st, rows, grid);
}

// ------------------------------------------------------------------ packs

__global__ __launch_bounds__(kBlock) void kp_kvec_slot(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  kvec_slot_body(const_cast<double*>(e.kv.kvec), e.kring, e.ring_R, 3 * e.kv.P, &e.nhc->step, bid, PACK_GRID());
}

template <bool F32, int KB>
__global__ __launch_bounds__(kBlock) void kp_rho_partial(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  const int kb = (e.kv.P + kWave * KB - 1) / (kWave * KB);  // k-vector blocks (rho_kblocks)
  rho_partial_body<F32, KB>(e.v, e.kv, bid % kb, bid / kb);
}

__global__ __launch_bounds__(kRedWaves * kWave) void kp_rho_reduce(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  rho_reduce_body(e.v, e.kv, bid);
}

template <bool F32, int LG, bool THERMO>
__global__ __launch_bounds__(kWave << LG) void kp_kforce(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  kforce_body<THERMO, false, F32, LG>(e.v, e.kv, e.rows_k, bid);
}

void launch_pack_kspace(int kernel, const PackLaunch& L, int nblocks, const PackSig& g, hipStream_t st) {
  if (nblocks <= 0) return;
  const dim3 grid(nblocks);
  switch (kernel) {
    case PK_KSLOT:
      hipLaunchKernelGGL(kp_kvec_slot, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      break;
    case PK_RHO:
#define KP_RHO_LAUNCH(F32_, KB_) hipLaunchKernelGGL((kp_rho_partial<F32_, KB_>), grid, dim3(kBlock), 0, st, PACK_ARGS(L))
      RHO_DISPATCH(g.rho_kb, g.f32, KP_RHO_LAUNCH);
#undef KP_RHO_LAUNCH
      break;
    case PK_RHO_RED:
      hipLaunchKernelGGL(kp_rho_reduce, grid, dim3(kRedWaves * kWave), 0, st, PACK_ARGS(L));
      break;
    case PK_KFORCE:
#define KP_KF_LAUNCH(F32_, LG_)                                                                    \
  do {                                                                                             \
    if (g.thermo)                                                                                  \
      hipLaunchKernelGGL((kp_kforce<F32_, LG_, true>), grid, dim3(kWave << LG_), 0, st, PACK_ARGS(L));  \
    else                                                                                           \
      hipLaunchKernelGGL((kp_kforce<F32_, LG_, false>), grid, dim3(kWave << LG_), 0, st, PACK_ARGS(L)); \
  } while (0)
      if (g.k_lanes_lg == 1) {
        if (g.f32) KP_KF_LAUNCH(true, 1); else KP_KF_LAUNCH(false, 1);
      } else if (g.k_lanes_lg == 3) {
        if (g.f32) KP_KF_LAUNCH(true, 3); else KP_KF_LAUNCH(false, 3);
      } else {
        if (g.f32) KP_KF_LAUNCH(true, 2); else KP_KF_LAUNCH(false, 2);
      }
#undef KP_KF_LAUNCH
      break;
    default:
      break;
  }
}

// The grid of each kernel of the step for one system, as its own launches
// size it (k_first_half .. k_rattle above and in the other kernel files).
int pack_blocks(const SysView& s, const IntView& iv, int P, int kernel) {
  const int atoms = (s.n + kBlock - 1) / kBlock;
  const int mols = (iv.nmol + kBlock - 1) / kBlock;
  switch (kernel) {
    case PK_NHC:
    case PK_SCAN:
    case PK_ONE:
      return 1;
    case PK_KSLOT:
      return (3 * P + kBlock - 1) / kBlock;
    case PK_FIRST_HALF:
    case PK_SCATTER:
    case PK_BONDED:
    case PK_SECOND:
    case PK_ATOMS:
      return atoms;
    case PK_CELL_SORT:
      return s.ncell;
    case PK_SHAKE:
    case PK_RATTLE:
      return mols;
    case PK_NBR:
      return kNbrParts * (s.n_items + kMisfitBlocks);
    case PK_PAIR:
      return pair_blocks(s);
    case PK_RHO:
      return rho_kblocks(s, P) * ((s.n + rho_atoms_per_block(s.n, rho_kblocks(s, P)) - 1) /
                                  rho_atoms_per_block(s.n, rho_kblocks(s, P)));
    case PK_RHO_RED:
      return (2 * P + kRedCols - 1) / kRedCols;
    case PK_KFORCE:
      return kforce_blocks(s);
    default:
      return 0;
  }
}

}  // namespace rbmd
