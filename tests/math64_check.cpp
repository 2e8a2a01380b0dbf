// Accuracy of the FP64 noise's log_unit and sincos_2pi (clv_math64.h) on the
// uniforms u = (k + 1/2) 2^-32 the noise feeds them, against long-double
// references.  Built and run by tests/test_math64.py (host compiler, no FMA
// contraction: the same operation sequence as the device code).
// Prints: log_ulp <max> sin_ulp <max> cos_ulp <max> sincos_abs <max> n <count>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "clv_math64.h"

static long double ulp_of(double ref) {
  int e;
  (void)std::frexp(ref, &e);  // ref = m 2^e, m in [1/2, 1)
  return std::ldexp(1.0L, e - 53);
}

int main() {
  const long double PI = 3.141592653589793238462643383279502884L;
  std::vector<uint32_t> ks;
  // edges: ends, quarter turns, the sqrt(1/2) split of the logarithm, powers of two
  const uint64_t K = 1ull << 32;
  for (uint64_t j = 0; j <= 4; ++j)
    for (int64_t d = -40; d <= 40; ++d) {
      const int64_t k = int64_t(j * (K / 4)) + d;
      if (k >= 0 && k < int64_t(K)) ks.push_back(uint32_t(k));
    }
  for (int64_t d = -40; d <= 40; ++d) ks.push_back(uint32_t(int64_t(0xB504F333ull) + d));
  for (int b = 0; b < 32; ++b)
    for (int d = -2; d <= 2; ++d) ks.push_back(uint32_t((1ull << b) + d));
  uint64_t x = 0x9E3779B97F4A7C15ull;  // splitmix64
  for (int i = 0; i < (1 << 21); ++i) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    ks.push_back(uint32_t(z ^ (z >> 31)));
  }
  double lmax = 0, smax = 0, cmax = 0, amax = 0;
  for (uint32_t k : ks) {
    const double u = (double(k) + 0.5) * 2.3283064365386963e-10;  // (k + 1/2) 2^-32, exact
    const double l = clv::log_unit(u);
    const long double lr = std::log((long double)u);
    lmax = std::fmax(lmax, double(std::fabs((long double)l - lr) / ulp_of(double(lr))));
    double s, c;
    clv::sincos_2pi(u, &s, &c);
    // independent reference: the long-double angle (absolute error only)
    const long double ang = 2 * PI * (long double)u;
    amax = std::fmax(amax, double(std::fmax(std::fabs((long double)s - std::sin(ang)),
                                            std::fabs((long double)c - std::cos(ang)))));
    // relative reference: the exact reduction 2 pi u = (pi/2) j + pi r in long double
    const long double jl = std::rint(4.0L * u);
    const long double rl = 2.0L * ((long double)u - jl / 4.0L);
    const int j = int(jl);
    const long double sl = std::sin(PI * rl), cl = std::cos(PI * rl);
    long double sref = (j & 1) ? cl : sl, cref = (j & 1) ? sl : cl;
    if (j & 2) sref = -sref;
    if ((j + 1) & 2) cref = -cref;
    smax = std::fmax(smax, double(std::fabs((long double)s - sref) / ulp_of(double(sref))));
    cmax = std::fmax(cmax, double(std::fabs((long double)c - cref) / ulp_of(double(cref))));
  }
  std::printf("log_ulp %.4f sin_ulp %.4f cos_ulp %.4f sincos_abs %.3e n %zu\n", lmax, smax, cmax, amax,
              ks.size());
  return 0;
}
