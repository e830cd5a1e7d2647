// Host build of rbmd_amd/csrc/sincos.hpp (the code the k-space kernels run):
// max |error| of sin(2 pi t), cos(2 pi t) against long double sinl/cosl over
// a dense grid of t in [-1/2, 1/2] and random t in [-50, 50].
#include <cmath>
#include <cstdio>
#include <random>

#include "sincos.hpp"

int main() {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> u(-50.0, 50.0);
  long double e64 = 0, e32 = 0;
  for (int i = 0; i < 3000000; ++i) {
    const double t = i < 1000001 ? (i - 500000) / 1000000.0 : u(g);
    double s, c;
    float sf, cf;
    rbmd::sincos_turns(t, &s, &c);
    rbmd::sincos_turns32(t, &sf, &cf);
    const long double th = two_pi * (long double)t;
    const long double S = sinl(th), C = cosl(th);
    e64 = fmaxl(e64, fmaxl(fabsl(s - S), fabsl(c - C)));
    e32 = fmaxl(e32, fmaxl(fabsl(sf - S), fabsl(cf - C)));
  }
  printf("%.6Le %.6Le\n", e64, e32);
  return 0;
}
