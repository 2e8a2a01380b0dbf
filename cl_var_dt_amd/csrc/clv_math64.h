// clv_math64.h -- the FP64 logarithm and sine/cosine of the Box-Muller noise.
//
// Box-Muller needs log(u) and sin/cos(2 pi v) of uniforms in (0, 1).  The
// general-purpose library routines spend most of their instructions on
// arguments these never take (and on last-bit rounding, with double-double
// arithmetic in the logarithm); these two take the uniforms' range only and
// come within 2 ulp of the exact values:
//   log_unit:  u = m 2^e with m in [sqrt(1/2), sqrt(2)),
//              log m = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...), s = (m-1)/(m+1)
//              (|s| <= 0.1716; the series to s^21 leaves < 1e-18 relative),
//              log u = e ln2 + log m with ln2 split into two doubles;
//   sincos_2pi: 2 pi v = (pi/2) j + pi r with j = rint(4v) and |r| <= 1/4
//              (exact in binary), sin(pi r) and cos(pi r) by their Taylor
//              series to r^17 and r^16 (the first omitted terms are at
//              most 2e-18), then the quarter-turn rotation by j.
// Every multiply-add is an explicit fma and nothing else is contracted, and
// the divide, frexp, rint and fma are IEEE-exact on both sides, so a host
// build performs the device's operation sequence: tests/test_math64.py checks
// that build against long-double references.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define CLV_HD __host__ __device__
#else
#define CLV_HD
#endif

namespace clv {

// log(u) for a positive normal double u (the noise passes u in [2^-33, 1)).
CLV_HD inline double log_unit(double u) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  int e;
  double m = frexp(u, &e);  // u = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;  // exact (Sterbenz)
  const double s = f / (m + 1.0);
  const double z = s * s;
  double q = 1.0 / 21.0;
  q = fma(q, z, 1.0 / 19.0);
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  const double s2 = s + s;
  const double lm = fma(s2 * z, q, s2);  // log m = 2s + 2s z q(z)
  const double ed = double(e);
  return fma(ed, 0.6931471805599453, fma(ed, 2.3190468138462996e-17, lm));
}

// sin(2 pi v) and cos(2 pi v) for v in [0, 1] with 4v not a half-integer
// (the noise's v = (k + 1/2) 2^-32 never is).
CLV_HD inline void sincos_2pi(double v, double *sp, double *cp) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const double jd = rint(4.0 * v);          // quarter turns
  const double r = 2.0 * fma(jd, -0.25, v);  // 2 pi v = (pi/2) j + pi r, exact
  const double w = r * r;
  // sin(pi r) = r S(w), S_k = (-1)^k pi^(2k+1) / (2k+1)!
  double S = 7.952054001475513e-07;
  S = fma(S, w, -2.1915353447830217e-05);
  S = fma(S, w, 0.00046630280576761255);
  S = fma(S, w, -0.0073704309457143504);
  S = fma(S, w, 0.08214588661112823);
  S = fma(S, w, -0.5992645293207921);
  S = fma(S, w, 2.5501640398773455);
  S = fma(S, w, -5.16771278004997);
  S = fma(S, w, 3.141592653589793);
  const double sr = r * S;
  // cos(pi r) = C(w), C_k = (-1)^k pi^(2k) / (2k)!
  double C = 4.303069587032947e-06;
  C = fma(C, w, -0.0001046381049248457);
  C = fma(C, w, 0.0019295743094039231);
  C = fma(C, w, -0.02580689139001406);
  C = fma(C, w, 0.2353306303588932);
  C = fma(C, w, -1.3352627688545895);
  C = fma(C, w, 4.0587121264167685);
  C = fma(C, w, -4.934802200544679);
  const double cr = fma(C, w, 1.0);
  // rotate by j quarter turns: (sin, cos) -> (cos, -sin) per turn
  const int j = int(jd);
  const double a = (j & 1) ? cr : sr;
  const double b = (j & 1) ? sr : cr;
  *sp = (j & 2) ? -a : a;
  *cp = ((j + 1) & 2) ? -b : b;
}

}  // namespace clv
