// sincos.hpp -- sin and cos of a phase given in turns, for the structure
// factor and k-space force kernels (and host checks of the same code).
//
// Method: the phase t (turns, |t| < 2^49) is reduced exactly in FP64 to a
// quadrant q = rint(4 t) mod 4 and a remainder r = t - rint(4 t)/4 (exact by
// Sterbenz's lemma) with |r| <= 1/8 turn (|2 pi r| <=
// pi/4); sin(2 pi r) and cos(2 pi r) are Taylor polynomials in r truncated
// where the next term is below the target precision's rounding, with
// coefficients (-1)^k (2 pi)^n / n! evaluated in double at compile time (the
// same bits on the host and the device); the quadrant rotates (s, c) by
// q * 90 degrees.
//   FP64: sin to r^17, cos to r^16 (next terms < 5e-17 and < 3e-18)
//   FP32: sin to r^9,  cos to r^10 (next terms < 2e-9 and < 1.2e-10);
//         r is rounded to FP32 after the FP64 reduction (error <= 2^-28 turn)
// On the device the FP32 variant the kernels use is sincos_turns32_hw: the
// FP64 reduction to r = t - rint(t) (exact, |r| <= 1/2 turn), rounded to
// FP32, then the hardware's v_sin_f32 / v_cos_f32, which take their argument
// in turns.  Measured on an MI355X over 2^24 points of [-1/2, 1/2]: at most
// 1.2e-7 absolute error against sincos_turns, the same bound as the FP32
// polynomial, in 2 instructions instead of ~20.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define RBMD_HD __host__ __device__
#else
#define RBMD_HD
#endif

namespace rbmd {
namespace sincos_detail {

constexpr double kTwoPi = 6.283185307179586476925286766559005768;

// (-1)^k (2 pi)^n / n! with n = 2k (+1 for sin)
constexpr double taylor(int n) {
  double v = 1.0;
  for (int i = 1; i <= n; ++i) v = v * kTwoPi / (double)i;
  return ((n / 2) % 2) ? -v : v;
}

}  // namespace sincos_detail

// quadrant rotation: (s, c) of the remainder -> (sin, cos) of the phase
template <typename R>
RBMD_HD inline void sincos_rotate(int q, R s, R c, R* sn, R* cs) {
  const R a = (q & 1) ? c : s;
  const R b = (q & 1) ? s : c;
  *sn = (q & 2) ? -a : a;
  *cs = ((q + 1) & 2) ? -b : b;
}

// FP64: sin(2 pi t), cos(2 pi t)
RBMD_HD inline void sincos_turns(double t, double* sn, double* cs) {
  using namespace sincos_detail;
  const double qd = rint(4.0 * t);
  const double r = fma(-0.25, qd, t);
  const int q = (int)fma(-4.0, floor(0.25 * qd), qd);  // qd mod 4, exact
  const double z = r * r;
  constexpr double S1 = taylor(1), S3 = taylor(3), S5 = taylor(5), S7 = taylor(7),
                   S9 = taylor(9), S11 = taylor(11), S13 = taylor(13),
                   S15 = taylor(15), S17 = taylor(17);
  constexpr double C2 = taylor(2), C4 = taylor(4), C6 = taylor(6), C8 = taylor(8),
                   C10 = taylor(10), C12 = taylor(12), C14 = taylor(14),
                   C16 = taylor(16);
  double ps = fma(z, S17, S15);
  ps = fma(z, ps, S13);
  ps = fma(z, ps, S11);
  ps = fma(z, ps, S9);
  ps = fma(z, ps, S7);
  ps = fma(z, ps, S5);
  ps = fma(z, ps, S3);
  ps = fma(z, ps, S1);
  double pc = fma(z, C16, C14);
  pc = fma(z, pc, C12);
  pc = fma(z, pc, C10);
  pc = fma(z, pc, C8);
  pc = fma(z, pc, C6);
  pc = fma(z, pc, C4);
  pc = fma(z, pc, C2);
  pc = fma(z, pc, 1.0);
  sincos_rotate<double>(q, r * ps, pc, sn, cs);
}

// FP32 polynomial on the FP64-reduced phase: sin(2 pi t), cos(2 pi t)
RBMD_HD inline void sincos_turns32(double t, float* sn, float* cs) {
  using namespace sincos_detail;
  const double qd = rint(4.0 * t);
  const float r = (float)fma(-0.25, qd, t);
  const int q = (int)fma(-4.0, floor(0.25 * qd), qd);  // qd mod 4, exact
  const float z = r * r;
  constexpr float S1 = (float)taylor(1), S3 = (float)taylor(3), S5 = (float)taylor(5), S7 = (float)taylor(7),
                  S9 = (float)taylor(9);
  constexpr float C2 = (float)taylor(2), C4 = (float)taylor(4), C6 = (float)taylor(6), C8 = (float)taylor(8),
                  C10 = (float)taylor(10);
  float ps = fmaf(z, S9, S7);
  ps = fmaf(z, ps, S5);
  ps = fmaf(z, ps, S3);
  ps = fmaf(z, ps, S1);
  float pc = fmaf(z, C10, C8);
  pc = fmaf(z, pc, C6);
  pc = fmaf(z, pc, C4);
  pc = fmaf(z, pc, C2);
  pc = fmaf(z, pc, 1.0f);
  sincos_rotate<float>(q, r * ps, pc, sn, cs);
}

#if defined(__HIPCC__)
// FP32 on the device: hardware sin / cos of the FP64-reduced phase
__device__ inline void sincos_turns32_hw(double t, float* sn, float* cs) {
  const float r = (float)(t - rint(t));
  *sn = __builtin_amdgcn_sinf(r);
  *cs = __builtin_amdgcn_cosf(r);
}
#endif

}  // namespace rbmd
