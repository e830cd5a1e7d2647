// pair.hpp -- the real-space pair term (LJ + erfc Coulomb) shared by the
// neighbour-list pair kernels and the fused cell-stencil force kernel.
#pragma once
#include "dev.hpp"
#include "cells.hpp"

namespace rbmd {

__device__ __forceinline__ double gnear(double alpha, double sqrt_alpha, double dis) {
  return erfc(sqrt_alpha * dis) / (dis * dis) + 2.0 * sqrt_alpha * exp(-alpha * dis * dis) / (sqrt(kPi) * dis);
}

// Step table of the RBL path: the reference's tabulated G at alpha = 0.1
// (ERFTable.h; erf_table.inc is generated from it), looked up as
// ExecStaticTable::Extract + TableGnearValue do (ExecStaticTable.h:38-39,
// 96-135): bucket = (float bits & (2^26 - 2^15)) >> 15, minus the first index
// of the table chosen by dis < 2.  The reference reads out of bounds below
// dis = 0.0998 and from 12 on (hazard Q3): here the bucket is clamped to the
// table, and table mode with rc >= 12 is refused when the system is created.
#define RBMD_ERF_STORAGE static __device__ const
#include "erf_table.inc"
#undef RBMD_ERF_STORAGE

__device__ __forceinline__ double gnear_table(double dis) {
  const float fd = (float)dis;
  const int idx = (int)((__float_as_uint(fd) & ((1u << 26) - (1u << 15))) >> 15);
  if (fd < 2.0f) {
    const int k = min(max(idx - RBMD_ERF_LEQ2_INDEX0, 0), RBMD_ERF_LEQ2_N - 1);
    return (double)kErfGnearLeq2[k];
  }
  const int k = min(max(idx - RBMD_ERF_GEQ2_INDEX0, 0), RBMD_ERF_GEQ2_N - 1);
  return (double)kErfGnearGeq2[k];
}

__device__ __forceinline__ void min_image(const SysView& s, double& rx, double& ry, double& rz) {
  if (fabs(rx) > s.box[0] * 0.5) rx -= (rx > 0 ? s.box[0] : -s.box[0]);
  if (fabs(ry) > s.box[1] * 0.5) ry -= (ry > 0 ? s.box[1] : -s.box[1]);
  if (fabs(rz) > s.box[2] * 0.5) rz -= (rz > 0 ? s.box[2] : -s.box[2]);
}


// decode() for lists over cells at least rc wide (VERLET, stencil half width
// 1): every image shift is -1, 0 or +1 box lengths, so the offset is the shift
// times shift_len, which equals decode()'s dxdydz * (shift * dims) bit for bit
// (a product with -1, 0 or 1 is exact) without its integer multiplies.
__device__ __forceinline__ void decode1(const SysView& s, int e, int& j, double& ox, double& oy, double& oz) {
  j = e >> kShiftBits;
  ox = (double)((e & 7) - 3) * s.shift_len[0];
  oy = (double)(((e >> 3) & 7) - 3) * s.shift_len[1];
  oz = (double)(((e >> 6) & 7) - 3) * s.shift_len[2];
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

// P (ISQ false) or Q (ISQ true) of the Coulomb table on interval k at the
// local coordinate u: Horner over the coefficient pairs of the LDS copy.
template <bool ISQ>
__device__ __forceinline__ double coul_poly(const double2* ctab, int k, double u) {
  const double2* c = ctab + (ISQ ? kCoulPairs * kCoulM : 0) + k;
  const double2 top = c[(kCoulPairs - 1) * kCoulM];
  double p = fma(top.y, u, top.x);
#pragma unroll
  for (int dp = kCoulPairs - 2; dp >= 0; --dp) {
    const double2 cc = c[dp * kCoulM];
    p = fma(p, u, cc.y);
    p = fma(p, u, cc.x);
  }
  return p;
}

// One pair term; returns the scalar f such that the force on i is f * r.
// Reference arithmetic: LJ ExecForceFunction.h:412-435 / 818-841, G(r) :682-689,
// special weights RunWorklet.cpp:2827-2857.  With rinv = 1/r:
//   LJ     f = -24 eps (2 s6 - 1) s6 rinv^2,    s6 = sigma^6 rinv^6
//   coul   f = -qq G(r)/r = -qq P(r) rinv^3,     e = qq Q(r) rinv / 2
//   weight f += (1 - w) qq rinv^3,               e -= (1 - w) qq rinv / 2
// P and Q come from the LDS copy of the piecewise polynomial table.  CHECK:
// 0 no test (pairs the list predicate accepted with the same vector); 1 the
// force function's own test 1e-4 < d2 < rc^2 (minimum-image vectors); 2 the
// list predicate rc^2 - d2 > 1e-4 and d2 > 1e-4 (entries of a skin list,
// which holds pairs out to rc + skin).  A pair that fails contributes zero.
//
// The force functions compile with FMA contraction off and their fused
// multiply-adds written out: the compiler would otherwise choose what to fuse
// per kernel (by what else reads a product), and the kernels with and without
// the energy and virial sums, alone and packed, would round the same force
// differently.
// 1/sqrt(x) for a positive finite x (a pair's d2 > 1e-4): the hardware
// estimate y refined by the second-order series of (1 - e)^(-1/2) in the
// residual e = 1 - x y^2, y (1 + e/2 + 3 e^2 / 8).  It is the device
// library's rsqrt without that function's select for zero and infinite
// arguments, which cannot occur here, so the value is the same to the bit.
__device__ __forceinline__ double rsqrt_pos(double x) {
#pragma clang fp contract(off)
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(y * -x, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

template <int CHECK>
__device__ __forceinline__ bool pair_in(const SysView& s, double d2) {
  if (CHECK == 0) return true;
  if (CHECK == 1) return d2 < s.cut2 && d2 > 0.0001;
  return s.cut2 - d2 > 0.0001 && d2 > 0.0001;
}

template <int MODE, bool EN, int CHECK = 1>
__device__ __forceinline__ double pair_term(const SysView& s, const double2* ctab, double d2, double qq, double epsij,
                                            double sig6, double w, double& elj, double& ec) {
#pragma clang fp contract(off)
  const bool in = pair_in<CHECK>(s, d2);
  double flj = 0, fc = 0;
  elj = 0;
  ec = 0;
  if (in) {
    const double rinv = rsqrt_pos(d2);
    const double rinv2 = rinv * rinv;
    const double sd6 = sig6 * (rinv2 * rinv2 * rinv2);
    flj = -24.0 * epsij * (fma(2.0, sd6, -1.0) * sd6) * rinv2;
    if (EN) elj = 0.5 * (4.0 * epsij * (sd6 - 1.0) * sd6);
    if (MODE != 2) {
      const double fk = (d2 * rinv) * s.ctab_inv_h;
      int k = (int)fk;
      k = k < kCoulM - 1 ? k : kCoulM - 1;
      const double u = fk - (double)k - 0.5;
      const double P = coul_poly<false>(ctab, k, u);
      const double rinv3 = rinv2 * rinv;
      if (MODE == 0) {
        fc = s.qqr2e * (-qq * P * rinv3);
      } else {
        fc = s.qqr2e * (qq * rinv3 * ((1.0 - w) - P));
      }
      if (EN) {
        const double Q = coul_poly<true>(ctab, k, u);
        ec = MODE == 0 ? s.qqr2e * (0.5 * qq * Q * rinv) : s.qqr2e * (0.5 * qq * rinv * (Q - (1.0 - w)));
      }
    }
  }
  if (MODE == 1) {
    flj *= w;
    elj *= w;
  }
  return fc + flj;
}

// FP32 Coulomb table polynomial (degree kCoulD32), same layout as coul_poly
template <bool ISQ>
__device__ __forceinline__ float coul_poly32(const float2* ctab, int k, float u) {
  const float2* c = ctab + (ISQ ? kCoulPairs32 * kCoulM : 0) + k;
  const float2 top = c[(kCoulPairs32 - 1) * kCoulM];
  float p = fmaf(top.y, u, top.x);
#pragma unroll
  for (int dp = kCoulPairs32 - 2; dp >= 0; --dp) {
    const float2 cc = c[dp * kCoulM];
    p = fmaf(p, u, cc.y);
    p = fmaf(p, u, cc.x);
  }
  return p;
}

// pair_term with the scalar arithmetic in FP32 (mixed precision): the caller
// forms the pair vector and d2 in FP64 and scales and sums in FP64.  1/sqrt
// is the hardware reciprocal square root (v_rsq_f32; a Newton step on it left
// the force error unchanged at ~3e-7 RMS, so none is taken).  The Coulomb
// factors come from the degree-kCoulD32 FP32 table.
template <int MODE, bool EN, int CHECK = 1>
__device__ __forceinline__ float pair_term32(const SysView& s, const float2* ctab, double d2d, float qq, float epsij,
                                             float sig6, float w, float& elj, float& ec) {
#pragma clang fp contract(off)
  const bool in = pair_in<CHECK>(s, d2d);
  const float d2 = (float)d2d;
  float flj = 0.f, fc = 0.f;
  elj = 0.f;
  ec = 0.f;
  if (in) {
    const float rinv = __builtin_amdgcn_rsqf(d2);
    const float rinv2 = rinv * rinv;
    const float sd6 = sig6 * (rinv2 * rinv2 * rinv2);
    flj = -24.0f * epsij * (fmaf(2.0f, sd6, -1.0f) * sd6) * rinv2;
    if (EN) elj = 0.5f * (4.0f * epsij * (sd6 - 1.0f) * sd6);
    if (MODE != 2) {
      const float fk = (d2 * rinv) * (float)s.ctab_inv_h;
      int k = (int)fk;
      k = k < kCoulM - 1 ? k : kCoulM - 1;
      const float u = fk - (float)k - 0.5f;
      const float P = coul_poly32<false>(ctab, k, u);
      const float rinv3 = rinv2 * rinv;
      const float qqc = (float)s.qqr2e * qq;
      if (MODE == 0) {
        fc = -qqc * P * rinv3;
      } else {
        fc = qqc * rinv3 * ((1.0f - w) - P);
      }
      if (EN) {
        const float Q = coul_poly32<true>(ctab, k, u);
        ec = MODE == 0 ? 0.5f * qqc * Q * rinv : 0.5f * qqc * rinv * (Q - (1.0f - w));
      }
    }
  }
  if (MODE == 1) {
    flj *= w;
    elj *= w;
  }
  return fc + flj;
}

}  // namespace rbmd
