// sampler.hpp -- seeded RBE k-vector sampler (host).
//
// Restates RBEPSAMPLE (framework/model/RBEPSample.h:40-177): a Metropolis-
// Hastings chain per axis with rounded Box-Muller proposals of width
// sigma_d = sqrt(alpha/2) L_d / pi (ExecutionNVT.cpp:336-345), target
// P(m) = exp(-(2 pi m / L)^2 / 4 alpha) / H_d, proposal probability q(m) from
// erf, zero vectors rejected.  The reference seeds a fresh std::mt19937 from
// std::random_device on every call (framework/math/Utiles.h:32-39), so its draws
// cannot be replayed; this sampler uses a counter-based splitmix64 stream per
// (seed, stream) instead, with the same call order (U1 loop, U2, accept u).
// P(m) and q(m) are tabulated per axis; a table entry is the same double the
// direct formula gives, so the chain is identical to evaluating it directly.
#pragma once
#include <stdint.h>

#include <vector>

namespace rbmd {

class KSampler {
 public:
  KSampler() = default;
  KSampler(double alpha, const double box[3], int P);
  // m_out: [3P] integer-valued doubles
  void draw(uint64_t seed, uint64_t stream, double* m_out) const;
  // k = 2 pi m / L_d, the reference's conversion (RunWorklet.cpp:2124-2126)
  void to_k(const double* m, double* k) const;
  int P() const { return P_; }
  double S() const { return S_; }

 private:
  double P_of(double m, int d) const;
  double q_of(double m, int d) const;
  double alpha_ = 0, box_[3] = {1, 1, 1}, H_[3] = {1, 1, 1}, sigma_[3] = {1, 1, 1}, S_ = 0;
  int P_ = 0, M_ = 0;
  std::vector<double> tabP_[3], tabq_[3];
};

double rbe_S(double alpha, const double box[3]);

}  // namespace rbmd
