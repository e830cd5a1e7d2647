// sampler.cpp -- seeded RBE k-vector sampler (see sampler.hpp).
#include "sampler.hpp"

#include <algorithm>
#include <cmath>

namespace rbmd {

namespace {
const double kPiH = 3.14159265358979323846;

struct SplitMix {
  uint64_t s;
  SplitMix(uint64_t seed, uint64_t stream) : s(seed ^ (stream * 0xD1B54A32D192ED03ull)) { next(); }
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// FetchSample_1D, RBEPSample.h:117-143
double box_muller(SplitMix& g, double sigma) {
  double U1;
  do {
    U1 = g.uniform();
  } while (U1 < 1e-6);
  double U2 = g.uniform();
  return sigma * std::sqrt(-2.0 * std::log(U1)) * std::cos(2 * kPiH * U2) + 0.0;
}
}  // namespace

// Compute_H / Compute_S, RBEPSample.h:40-63
static void compute_H(double alpha, const double box[3], double H[3]) {
  for (int d = 0; d < 3; ++d) {
    double factor = -(alpha * box[d] * box[d]);
    H[d] = 0;
    for (int m = -10; m <= 10; ++m) H[d] += std::exp(m * m * factor);
    H[d] *= std::sqrt(-factor / kPiH);
  }
}

double rbe_S(double alpha, const double box[3]) {
  double H[3];
  compute_H(alpha, box, H);
  return H[0] * H[1] * H[2] - 1;
}

KSampler::KSampler(double alpha, const double box[3], int P) : alpha_(alpha), P_(P) {
  for (int d = 0; d < 3; ++d) box_[d] = box[d];
  compute_H(alpha, box, H_);
  S_ = H_[0] * H_[1] * H_[2] - 1;
  double smax = 0;
  for (int d = 0; d < 3; ++d) {
    sigma_[d] = std::sqrt(alpha / 2.0) * box[d] / kPiH;
    smax = std::max(smax, sigma_[d]);
  }
  M_ = (int)std::ceil(12.0 * smax) + 8;
  for (int d = 0; d < 3; ++d) {
    tabP_[d].resize(2 * M_ + 1);
    tabq_[d].resize(2 * M_ + 1);
    for (int m = -M_; m <= M_; ++m) {
      tabP_[d][m + M_] = P_of((double)m, d);
      tabq_[d][m + M_] = q_of((double)m, d);
    }
  }
}

// Distribution_P, RBEPSample.h:89-95
double KSampler::P_of(double x, int d) const {
  double t = 2 * kPiH * x / box_[d];
  return std::exp(-(t * t) / (4 * alpha_)) / H_[d];
}

// Distribution_q, RBEPSample.h:97-115
double KSampler::q_of(double x, int d) const {
  double s = std::sqrt(alpha_ * box_[d] * box_[d] / (kPiH * kPiH));
  if (x == 0) return std::erf((1.0 / 2) / s);
  return (std::erf(((1.0 / 2) + std::fabs(x)) / s) - std::erf((std::fabs(x) - (1.0 / 2)) / s)) / 2;
}

void KSampler::draw(uint64_t seed, uint64_t stream, double* out) const {
  SplitMix g(seed, stream);
  auto P = [&](double m, int d) {
    long k = (long)m;
    return (k >= -M_ && k <= M_) ? tabP_[d][k + M_] : P_of(m, d);
  };
  auto q = [&](double m, int d) {
    long k = (long)m;
    return (k >= -M_ && k <= M_) ? tabq_[d][k + M_] : q_of(m, d);
  };
  double X0[3];
  do {
    for (int d = 0; d < 3; ++d) X0[d] = std::round(box_muller(g, sigma_[d]));
  } while (std::fabs(X0[0]) < 1e-6 && std::fabs(X0[1]) < 1e-6 && std::fabs(X0[2]) < 1e-6);
  for (int d = 0; d < 3; ++d) out[d] = X0[d];
  // MH_Algorithm (RBEPSample.h:65-87) and Fetch_P_Sample (:145-176)
  for (int i = 1; i < P_; i++) {
    double X1[3];
    for (int d = 0; d < 3; ++d) {
      double m = X0[d];
      double mw = std::round(box_muller(g, sigma_[d]));
      double prob = (P(mw, d) / P(m, d)) * (q(m, d) / q(mw, d));
      prob = std::min(prob, 1.0);
      double u = g.uniform();
      if (u <= prob) m = mw;
      X1[d] = m;
    }
    for (int d = 0; d < 3; ++d) out[3 * i + d] = X1[d];
    for (int d = 0; d < 3; ++d) X0[d] = X1[d];
    if (std::fabs(X1[0]) < 1e-6 && std::fabs(X1[1]) < 1e-6 && std::fabs(X1[2]) < 1e-6) {
      i = i - 1;
      continue;
    }
  }
}

void KSampler::to_k(const double* m, double* k) const {
  for (int l = 0; l < P_; ++l)
    for (int d = 0; d < 3; ++d) k[3 * l + d] = 2 * kPiH * m[3 * l + d] / box_[d];
}

}  // namespace rbmd
